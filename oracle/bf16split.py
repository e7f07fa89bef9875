"""numpy restatement of the BF16-split product, for tests only.

TEST INFRASTRUCTURE ONLY: the product path never imports this module.

Exponent-aligned splitting (Ozaki, Ogita, Oishi and Rump, Numer. Algorithms
59, 2012) with one exponent per matrix row of A and per matrix column of B:
E = exponent with |x| < 2^E over the row (column); slice t (1-based) is
rint(r / 2^(E-8t)) * 2^(E-8t) of the remainder r; the product keeps the slice
pairs with t + u <= S + 1. Every quantity here is exact in float64 except the
final matmul sums, so `emulate` is the value the GPU kernel must reproduce up
to FP64 summation order (and the 2^-40-relative rounding of its level 4+
FP32 accumulator).
"""
from __future__ import annotations

import numpy as np

NO_EXP = -30000


def exponents(x: np.ndarray, axis: int) -> np.ndarray:
    """Smallest E with |x| < 2^E along axis (NO_EXP where all zero)."""
    m = np.max(np.abs(x), axis=axis)
    _, e = np.frexp(m)
    e = e.astype(np.int64)
    e[m == 0] = NO_EXP
    return e


def split(x: np.ndarray, e: np.ndarray, slices: int):
    """Slices of x (e broadcast against x) and the final remainder."""
    r = np.array(x, dtype=np.float64, copy=True)
    out = []
    for t in range(1, slices + 1):
        ue = e - 8 * t
        live = (e != NO_EXP) & (ue >= -120)
        ue = np.where(live, ue, 0)
        unit = np.where(live, np.ldexp(1.0, ue), 0.0)
        inv = np.where(live, np.ldexp(1.0, -ue), 0.0)
        sl = np.rint(r * inv) * unit
        r = r - sl
        out.append(sl)
    return out, r


def emulate(a: np.ndarray, b: np.ndarray, slices: int) -> np.ndarray:
    """Dense BF16-split product of dense a (rows sliced) and b (columns sliced)."""
    ea = exponents(a, axis=1)[:, None]
    eb = exponents(b, axis=0)[None, :]
    sa, _ = split(a, ea, slices)
    sb, _ = split(b, eb, slices)
    c = np.zeros((a.shape[0], b.shape[1]))
    for t in range(slices):
        for u in range(slices - t):
            c += sa[t] @ sb[u]
    return c
