"""Bit-exact CPU restatement of the int8-sliced FP64 GEMM (CORR_PREC_INT8X5,
corr_amd/csrc/ozaki.hip), used by the tests as the checker of the GPU kernel.

Row scales of A and column scales of B are powers of two 2^e with
max |x| / 2^e < 127.5/128; the scaled values are cut into 5 int8 slices by
round-to-nearest (q_s = rint(128 r_s), r_{s+1} = 128 r_s - q_s); the 15 slice
products with s + t <= 4 are exact integer sums, combined per diagonal by
Horner's rule in 2^-7 and scaled by sig[m] tau[n] 2^-14, in the kernel's order.
"""
from __future__ import annotations

import numpy as np

S = 5


def scale(mx: np.ndarray, bad: np.ndarray) -> np.ndarray:
    f, e = np.frexp(np.where(mx > 0, mx, 1.0))
    e = e + (f * 128.0 >= 127.5)
    with np.errstate(over="ignore"):
        out = np.ldexp(1.0, np.minimum(e, 1024))
    out = np.where(mx > 0, out, 1.0)
    return np.where(bad | (e > 1023), np.nan, out)


def inv_scale(sc: np.ndarray) -> np.ndarray:
    _, e = np.frexp(np.where(np.isnan(sc), 1.0, sc))
    return np.where(np.isnan(sc), 0.0, np.ldexp(1.0, 1 - e))


def slices(x: np.ndarray) -> list[np.ndarray]:
    """Round-to-nearest int8 slices of x in (-1, 1): x = sum_s q_s 2^-7(s+1) + r."""
    r = np.array(x, dtype=np.float64)
    out = []
    for _ in range(S):
        r = r * 128.0
        t = np.rint(r)
        r = r - t
        out.append(t.astype(np.int64))
    return out


def row_scales(A: np.ndarray) -> np.ndarray:
    a = np.abs(A)
    bad = ~np.all(a <= np.finfo(np.float64).max, axis=1)
    mx = np.max(np.where(np.isfinite(a), a, 0.0), axis=1) if A.shape[1] else np.zeros(A.shape[0])
    return scale(mx, bad)


def gemm_int8x5(A: np.ndarray, B: np.ndarray, alpha: float = 1.0) -> np.ndarray:
    """alpha A B as the GPU kernel computes it, bit for bit."""
    A = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    sig = row_scales(A)
    tau = row_scales(B.T)
    with np.errstate(invalid="ignore"):
        qa = slices(np.where(np.isfinite(A), A, 0.0) * inv_scale(sig)[:, None])
        qb = slices(np.where(np.isfinite(B), B, 0.0) * inv_scale(tau)[None, :])
    acc = [sum(qa[s] @ qb[d - s] for s in range(d + 1)) for d in range(S)]
    h = acc[4].astype(np.float64)
    for d in (3, 2, 1, 0):
        h = h * (1.0 / 128.0) + acc[d].astype(np.float64)
    with np.errstate(invalid="ignore"):
        return alpha * ((h * (1.0 / 16384.0)) * sig[:, None] * tau[None, :])
