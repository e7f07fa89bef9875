"""The BF16 splitting restated in oracle/bf16split.py, checked against what
the mathematics fixes: exact BF16 slices, exact reconstruction, the remainder
bound, the FP32 exactness bound of one accumulation chain, and the product
error shrinking by about 2^-8 per slice against the FP64 product.
"""
import numpy as np
import pytest

from lamb_amd import systems as S
from oracle import bf16split as B


def bf16_exact(x):
    f = x.astype(np.float32)
    return np.array_equal(f.astype(np.float64), x) and not np.any(f.view(np.uint32) & 0xFFFF)


@pytest.mark.parametrize("slices", [3, 4, 5, 6])
def test_slices_are_exact_bf16_and_reconstruct(slices):
    rng = np.random.default_rng(slices)
    x = rng.uniform(-1, 1, (40, 30)) * np.exp(rng.uniform(-12, 3, (40, 1)))
    x[3] = 0.0
    x[5, ::2] = 0.0
    e = B.exponents(x, axis=1)[:, None]
    sl, r = B.split(x, e, slices)
    total = np.zeros_like(x)
    for t, s in enumerate(sl, start=1):
        assert bf16_exact(s)
        m = s / np.where(e == B.NO_EXP, 1.0, np.ldexp(1.0, np.where(e == B.NO_EXP, 0, e - 8 * t)))
        assert np.array_equal(m, np.rint(m))  # integer multiples of the slice unit
        assert np.max(np.abs(m)) <= (256 if t == 1 else 128)
        total += s
    assert np.array_equal(total + r, x)  # exact split
    bound = np.ldexp(1.0, np.where(e == B.NO_EXP, -2000, e - 8 * slices - 1))
    assert np.all(np.abs(r) <= bound)
    assert np.all(r[3] == 0)


def test_chain_bound_keeps_fp32_sums_exact():
    # level 2: 256 products below 2^16 units; level 3: 2 x 256 products below 2^15 units
    assert 256 * 256 * 256 == 2 ** 24
    assert 2 * 256 * 256 * 128 == 2 ** 24
    # worst case in FP32: sums of integer products of that size are exact
    rng = np.random.default_rng(3)
    m1 = rng.integers(-256, 257, (4, 256)).astype(np.float32)
    m2 = rng.integers(-256, 257, (256, 4)).astype(np.float32)
    acc = np.zeros((4, 4), np.float32)
    for k in range(256):
        acc += np.outer(m1[:, k], m2[k, :]).astype(np.float32)
    assert np.array_equal(acc.astype(np.int64), m1.astype(np.int64) @ m2.astype(np.int64))


def test_product_error_falls_per_slice():
    atoms = S.water_system(96, seed=5)
    a = S.decay_matrix(atoms, 8.0, 0.4, 6).to_dense()
    b = S.decay_matrix(atoms, 8.0, 0.4, 7).to_dense()
    ref = a @ b
    errs = [np.linalg.norm(B.emulate(a, b, s) - ref) / np.linalg.norm(ref) for s in (3, 4, 5, 6)]
    print("relative Frobenius error by slice count 3..6:", errs)
    for e0, e1 in zip(errs, errs[1:]):
        assert e1 < e0 / 32
    assert errs[2] <= 1e-9  # five slices meet 1e-9 on this box
