#!/usr/bin/env python3
"""Freeze the VALU op counts used by bench.py's roofline (writes opcount.json).

Builds probe kernels around this package's own device functions (Threefry2x64-20,
FP64 and FP32 Box-Muller) for gfx950, reads the VALU instructions of each
kernel body from the compiler's assembly listing, subtracts the probe's
load/store scaffolding, and drops the blocks that only run for arguments
outside Box-Muller's range (the sine/cosine large-argument reduction).  Counts
are wave instructions, i.e. lane-ops per lane.

Per-sample algorithmic costs of the shift model (one normal per inner sample,
so half a Threefry block and half a Box-Muller pair):
  fp64  inner = (threefry + bm64) / 2 + 4    (Y + z, s1 +=, s2 += d*d)
  mixed inner = (threefry + bm32) / 2 + 16   (Y + z, d*d, |Y|+|z|, three compensated adds)
  outer = threefry + bm64                    (one outer normal in FP64)

Per-sample costs of an Exact-GBM inner sample of the sub-sampled portfolio
(two Threefry blocks: the option uniform, then the Box-Muller pair):
  mixed = 2 threefry + bm32 + exact_mix + search + 14
          (exact_mix: the FP32 loss and its error scale from the 16-byte
          record; search: the option uniform to its guide bucket, the path of
          most draws; 14: three compensated adds, the squared error scale and
          the work add)
  fp64  = 2 threefry + bm64 + exact_d + search + 2
          (exact_d: the FP64 loss divided by the option's probability)

Keys already in opcount.json are kept (the counts are frozen once, before
tuning); --refresh recomputes every key.
"""
import json
import os
import re
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

PROBE = r'''
#include "%s/csrc/rng.hpp"
using namespace nmlmc;
extern "C" __global__ void op_base(const uint64_t* in, uint64_t* out) {
    int i = threadIdx.x; out[i] = in[i] ^ in[i + 64] ^ in[i + 128]; out[i + 64] = in[i]; }
extern "C" __global__ void op_threefry(const uint64_t* in, uint64_t* out) {
    int i = threadIdx.x; u64x2 r = threefry2x64_20(in[i], 0, in[i + 64], in[i + 128]);
    out[i] = r.a; out[i + 64] = r.b; }
extern "C" __global__ void op_bm64(const uint64_t* in, double* out) {
    int i = threadIdx.x; double x, y; boxmuller_f64(in[i], in[i + 64], x, y);
    out[i] = x; out[i + 64] = y; }
extern "C" __global__ void op_bm32(const uint64_t* in, float* out) {
    int i = threadIdx.x; float x, y; boxmuller_f32(in[i], in[i + 64], x, y);
    out[i] = x; out[i + 64] = y; }
'''

PROBE_PF = r'''
#include "%s/csrc/portfolio_dev.hpp"
using namespace nmlmc;
extern "C" __global__ void op_base_mix(const float* f, const uint32_t* u, float* out) {
    int i = threadIdx.x;
    float a = f[i] + f[i + 64] + f[i + 128] + (float)u[i] + f[i + 192] + f[i + 256] + f[i + 320] + f[i + 384]
              + f[i + 448] + f[i + 512] + f[i + 576] + f[i + 640];
    out[i] = a; out[i + 64] = a; }
extern "C" __global__ void op_exact_mix(const float* f, const uint32_t* u, float* out) {
    int i = threadIdx.x;
    MixRec q{f[i], f[i + 64], f[i + 128], u[i]};
    MixStock st{f[i + 192], f[i + 256], f[i + 320], f[i + 384], f[i + 448], {0.f, 0.f, 0.f}};
    float mag;
    float d = inner_exact_mix(q, st, true, true, f[i + 512], f[i + 576], f[i + 640], mag);
    out[i] = d; out[i + 64] = mag; }
extern "C" __global__ void op_base_d(const double* f, const PortfolioDevice pf, double* out) {
    int i = threadIdx.x;
    double a = f[i] + f[i + 64] + f[i + 128] + f[i + 192] + f[i + 256] + f[i + 320];
    out[i] = a; out[i + 64] = a; }
extern "C" __global__ void op_exact_d(const double* f, const PortfolioDevice pf, double* out) {
    int i = threadIdx.x;
    OptionRec o{};
    o.strike = f[i]; o.maturity = f[i + 64]; o.w = f[i + 128]; o.prob = f[i + 192];
    o.type = 1; o.stock_ind = 0; o.W_est = 1.;
    double w = 0.;
    const double R[1] = {f[i + 256]};
    double v = inner_exact_d(pf, o, R, f[i + 320], f[i + 320] * 0.5, w) / (o.prob / pf.prob_norm);
    out[i] = v; out[i + 64] = w; }
extern "C" __global__ void op_base_search(const uint64_t* in, const PortfolioDevice pf, uint32_t* out) {
    int i = threadIdx.x; out[i] = (uint32_t)in[i]; }
extern "C" __global__ void op_search(const uint64_t* in, const PortfolioDevice pf, uint32_t* out) {
    int i = threadIdx.x;
    const uint32_t nb = pf.guide_nb[1];
    const double fb = u01_f64(in[i]) * pf.compute_prob[1] * pf.guide_scale[1];
    out[i] = fb < (double)(nb - 1) ? (uint32_t)fb : nb - 1; }
'''


def kernel_bodies(asm):
    res, cur = {}, None
    for line in asm.split("\n"):
        t = line.split(";")[0].strip()
        m = re.match(r"^(op_\w+):$", t)
        if m:
            cur = m.group(1)
            res[cur] = []
            continue
        if cur and t.startswith("s_endpgm"):
            cur = None
            continue
        if cur and t and (t.endswith(":") or not t.startswith((".", "/"))):
            res[cur].append(t)
    return res


def valu_count(body, skip_large_arg=False):
    """VALU instructions; optionally skip the block guarded by the large-argument
    test of the trig reduction (entered for |x| >= 2^30 only)."""
    n, skipping, skip_label = 0, False, None
    for i, t in enumerate(body):
        if t.endswith(":"):
            if skipping and t[:-1] == skip_label:
                skipping = False
            continue
        if skip_large_arg and skip_label is None and t.startswith("s_cbranch_execz"):
            # The first divergent branch of op_bm64 guards the large-argument path.
            skip_label = t.split()[1]
            skipping = True
            continue
        if not skipping and t.startswith("v_"):
            n += 1
    return n


def compile_asm(text):
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.hip")
        with open(src, "w") as f:
            f.write(text)
        out = os.path.join(d, "probe.s")
        subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17",
                        "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops", "--cuda-device-only", "-S", src, "-o", out],
                       check=True, stderr=subprocess.DEVNULL)
        return open(out).read()


def main():
    import sys
    b = kernel_bodies(compile_asm(PROBE % HERE))
    bp = kernel_bodies(compile_asm(PROBE_PF % HERE))
    base = valu_count(b["op_base"])
    tf = valu_count(b["op_threefry"]) - base
    bm64 = valu_count(b["op_bm64"], skip_large_arg=True) - base
    # The FP32 pair's first divergent branch is the words near 1 (2^-8 of them).
    bm32 = valu_count(b["op_bm32"], skip_large_arg=True) - base
    # The base probes combine their loads with 11 (FP32) and 5 (FP64) adds.
    exact_mix = valu_count(bp["op_exact_mix"]) - valu_count(bp["op_base_mix"]) + 11
    exact_d = valu_count(bp["op_exact_d"]) - valu_count(bp["op_base_d"]) + 5
    search = valu_count(bp["op_search"]) - valu_count(bp["op_base_search"])
    res = {
        "isa": "gfx950",
        "unit": "VALU wave instructions (lane-ops per lane)",
        "threefry2x64_20": tf,
        "boxmuller_f64": bm64,
        "boxmuller_f32": bm32,
        "shift_model": {
            "fp64": {"inner": (tf + bm64) / 2 + 4, "outer": tf + bm64},
            "mixed": {"inner": (tf + bm32) / 2 + 16, "outer": tf + bm64},
        },
        "exact_mix_term": exact_mix,
        "exact_fp64_term": exact_d,
        "option_search_guide_hit": search,
        "portfolio_exact": {
            "mixed": {"inner": 2 * tf + bm32 + exact_mix + search + 14},
            "fp64": {"inner": 2 * tf + bm64 + exact_d + search + 2},
        },
    }
    path = os.path.join(HERE, "opcount.json")
    if os.path.exists(path) and "--refresh" not in sys.argv:
        with open(path) as f:
            old = json.load(f)
        old.update({k: v for k, v in res.items() if k not in old})
        res = old
    with open(path, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
