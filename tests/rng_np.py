"""Counter-based draws in numpy, written from the published definitions and not
from the oracle: Threefry2x64-20 (Salmon et al., SC'11; the Skein 1.3 rotation
and key-schedule constants), Random123's u01/uneg11 for 64-bit words and the
Box-Muller map.  Used by the tests that pin the oracle's outer draws and its
refinement rule: an independent restatement that a slip in the oracle's own
code would not share.  Vectorised over arrays of counters and keys."""
import numpy as np

_ROT = (16, 42, 12, 31, 16, 32, 24, 21)
_PARITY = np.uint64(0x1BD11BDAA9FC1A22)


def _rotl(x, r):
    return (x << np.uint64(r)) | (x >> np.uint64(64 - r))


def threefry2x64_20(c0, c1, k0, k1):
    """Blocks for counters (c0, c1) under keys (k0, k1); all broadcast."""
    c0, c1, k0, k1 = (np.asarray(v, dtype=np.uint64) for v in (c0, c1, k0, k1))
    with np.errstate(over="ignore"):  # arithmetic mod 2^64 is the definition
        ks = (k0, k1, _PARITY ^ k0 ^ k1)
        x0 = c0 + ks[0]
        x1 = c1 + ks[1]
        for rnd in range(20):
            x0 = x0 + x1
            x1 = _rotl(x1, _ROT[rnd % 8]) ^ x0
            if rnd % 4 == 3:
                s = rnd // 4 + 1
                x0 = x0 + ks[s % 3]
                x1 = x1 + ks[(s + 1) % 3] + np.uint64(s)
    return x0, x1


def u01(u):
    """(0, 1]: u 2^-64 + 2^-65."""
    return np.asarray(u, np.uint64).astype(np.float64) * 2.0 ** -64 + 2.0 ** -65


def uneg11(u):
    """[-1, 1]: (signed u) 2^-63 + 2^-64."""
    return np.asarray(u, np.uint64).view(np.int64).astype(np.float64) * 2.0 ** -63 + 2.0 ** -64


def boxmuller(u0, u1):
    """(x, y) = (sin, cos)(pi uneg11(u0)) sqrt(-2 ln u01(u1))."""
    a = np.pi * uneg11(u0)
    r = np.sqrt(-2.0 * np.log(u01(u1)))
    return np.sin(a) * r, np.cos(a) * r


def stream_normals(k0, k1, n):
    """The first n normals of the stream keyed (k0, k1) whose blocks sit at
    counters {j, 0}: the first normal of a pair is boxmuller(second word,
    first word) of block j, where the first word handed out is the block's
    element 1; the second normal is the pair's y.  Returns [len(k1), n]."""
    k1 = np.atleast_1d(np.asarray(k1, np.uint64))
    out = np.empty((len(k1), n))
    for j in range((n + 1) // 2):
        a, b = threefry2x64_20(np.uint64(j), np.uint64(0), k0, k1)
        x, y = boxmuller(a, b)
        out[:, 2 * j] = x
        if 2 * j + 1 < n:
            out[:, 2 * j + 1] = y
    return out
