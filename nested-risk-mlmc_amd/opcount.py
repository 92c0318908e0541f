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


def main():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.hip")
        with open(src, "w") as f:
            f.write(PROBE % HERE)
        out = os.path.join(d, "probe.s")
        subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17",
                        "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops", "--cuda-device-only", "-S", src, "-o", out],
                       check=True, stderr=subprocess.DEVNULL)
        asm = open(out).read()
    b = kernel_bodies(asm)
    base = valu_count(b["op_base"])
    tf = valu_count(b["op_threefry"]) - base
    bm64 = valu_count(b["op_bm64"], skip_large_arg=True) - base
    bm32 = valu_count(b["op_bm32"]) - base
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
    }
    with open(os.path.join(HERE, "opcount.json"), "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
