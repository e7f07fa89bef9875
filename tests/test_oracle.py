"""Pins the CPU oracle (oracle/lamb_oracle.c) against what mathematics fixes.

Each check is chosen so that a plausible mistake in the oracle fails it:
dense FP64 products of the densified operands (a transposed operand, a
dropped k term or a wrong block offset changes the values), brute-force
enumeration of kept triples in Python with an exact rational sum of squares
(a wrong prefilter comparison or norm order changes the count), identities
(A = I, B = I), symmetry, and integer-valued golden fixtures whose FP64
products are exact.
"""
import fractions
import itertools

import numpy as np
import pytest

import oracle
from lamb_amd import systems as S


def dense_product(a, b):
    return a.to_dense() @ b.to_dense()


def blocks_of(csr):
    rows = csr.rows()
    return {(int(rows[t]), int(csr.col[t])): csr.block(t) for t in range(csr.n_blocks)}


def exact_fma_nsq(x):
    """nsq with each step fma(x, x, s) evaluated exactly and rounded once."""
    s = 0.0
    for v in np.asarray(x, dtype=np.float64).reshape(-1, order="F"):
        s = float(fractions.Fraction(v) * fractions.Fraction(v) + fractions.Fraction(s))
    return s


@pytest.mark.parametrize("seed,n,sizes", [(1, 16, (5, 13)), (2, 24, (5, 13, 23)), (3, 12, (1, 2, 3))])
def test_product_matches_dense(seed, n, sizes):
    atoms, a = S.small_system(n, seed, sizes=sizes)
    b = S.decay_matrix(atoms, 6.0, 0.5, seed + 200)
    c, st = oracle.multiply(a, b, eps=0.0)
    ref = dense_product(a, b)
    got = c.to_dense()
    assert np.linalg.norm(got - ref) <= 1e-14 * np.linalg.norm(ref)
    # every stored block lies in the structural pattern and no block of it is missing
    Pa = {(i, j) for (i, j) in blocks_of(a)}
    Pb = {(i, j) for (i, j) in blocks_of(b)}
    struct = {(i, j) for (i, k) in Pa for (k2, j) in Pb if k == k2}
    assert set(blocks_of(c)) == struct
    # flop count = sum over triples of 2 m k n
    blk = a.blk
    fl = sum(2.0 * blk[i] * blk[k] * blk[j] for (i, k) in Pa for (k2, j) in Pb if k == k2)
    assert st["flops"] == fl


def test_identity_operands():
    atoms, a = S.small_system(20, 5)
    n = a.n_atoms
    # identity in the same block structure
    sizes = a.blk.astype(np.int64) ** 2
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    data = np.concatenate([np.eye(int(m)).reshape(-1, order="F") for m in a.blk])
    eye = S.BlockCSR(a.blk.copy(), np.arange(n + 1, dtype=np.int64),
                     np.arange(n, dtype=np.int32), off.astype(np.int64), data)
    c1, _ = oracle.multiply(eye, a)
    c2, _ = oracle.multiply(a, eye)
    for c in (c1, c2):
        assert np.array_equal(c.row_p, a.row_p) and np.array_equal(c.col, a.col)
        assert np.array_equal(c.data, a.data)


def test_symmetric_square_is_symmetric():
    atoms, a = S.small_system(30, 9, symmetric=True)
    d = a.to_dense()
    assert np.array_equal(d, d.T)
    c, _ = oracle.multiply(a, a)
    cd = c.to_dense()
    assert np.linalg.norm(cd - cd.T) <= 1e-14 * np.linalg.norm(cd)


def test_nsq_is_sequential_fma():
    rng = np.random.default_rng(4)
    for m, n in [(5, 5), (13, 5), (23, 13), (1, 7)]:
        x = rng.standard_normal((m, n)) * 10.0 ** rng.integers(-3, 3)
        assert oracle.nsq1(x.reshape(-1, order="F")) == exact_fma_nsq(x)


def test_prefilter_matches_brute_force():
    # a decay so steep that the prefilter drops triples
    atoms, a = S.small_system(14, 21, cutoff=9.0, decay=1.5)
    b = S.decay_matrix(atoms, 9.0, 1.2, 22)
    eps = 1e-3
    na = {key: exact_fma_nsq(v) for key, v in blocks_of(a).items()}
    nb = {key: exact_fma_nsq(v) for key, v in blocks_of(b).items()}
    kept, dropped = [], 0
    for (i, k), x in na.items():
        for (k2, j), y in nb.items():
            if k2 != k:
                continue
            if x * y >= eps * eps:
                kept.append((i, k, j))
            else:
                dropped += 1
    c, st = oracle.multiply(a, b, eps=eps)
    assert dropped > 0 and st["n_triples_filtered"] == dropped
    assert st["n_triples_kept"] == len(kept)
    # values: sum over kept k only, post-filtered by nsq(C_ij) >= eps^2
    A, B = blocks_of(a), blocks_of(b)
    acc = {}
    for (i, k, j) in kept:
        acc[(i, j)] = acc.get((i, j), 0) + A[(i, k)] @ B[(k, j)]
    expect = {key: v for key, v in acc.items() if exact_fma_nsq(v) >= eps * eps}
    got = blocks_of(c)
    assert set(got) == set(expect)
    for key, v in expect.items():
        assert np.allclose(got[key], v, rtol=1e-13, atol=1e-16)


def test_alpha_beta_union_pattern():
    atoms, a = S.small_system(18, 31)
    b = S.decay_matrix(atoms, 4.0, 0.5, 32)
    old = S.decay_matrix(atoms, 7.0, 0.5, 33)
    c, _ = oracle.multiply(a, b, alpha=-2.0, beta=3.0, c_old=old)
    ref = -2.0 * dense_product(a, b) + 3.0 * old.to_dense()
    assert np.linalg.norm(c.to_dense() - ref) <= 1e-14 * np.linalg.norm(ref)
    pat_prod = set(blocks_of(oracle.multiply(a, b)[0]))
    assert set(blocks_of(c)) == pat_prod | set(blocks_of(old))


def test_trace_and_gershgorin():
    atoms, a = S.small_system(22, 41, symmetric=True)
    d = a.to_dense()
    assert abs(oracle.trace(a) - np.trace(d)) <= 1e-13 * np.abs(np.diag(d)).sum()
    lo, hi = oracle.gershgorin(a)
    ev = np.linalg.eigvalsh(d)
    assert lo <= ev.min() + 1e-12 and hi >= ev.max() - 1e-12
    r = np.abs(d).sum(1) - np.abs(np.diag(d))
    assert np.isclose(lo, (np.diag(d) - r).min()) and np.isclose(hi, (np.diag(d) + r).max())


def test_adjoint_matches_dense():
    atoms, a = S.small_system(16, 51)
    b = S.decay_matrix(atoms, 6.0, 0.5, 52)
    c, _ = oracle.multiply(a, b)
    rng = np.random.default_rng(53)
    g = S.BlockCSR(c.blk, c.row_p, c.col, c.blk_off, rng.standard_normal(len(c.data)))
    da, db = oracle.adjoint(a, b, g)
    G = g.to_dense()
    dA = G @ b.to_dense().T
    dB = a.to_dense().T @ G
    ga = S.BlockCSR(a.blk, a.row_p, a.col, a.blk_off, da).to_dense()
    gb = S.BlockCSR(b.blk, b.row_p, b.col, b.blk_off, db).to_dense()
    mask_a = a.to_dense() != 0
    mask_b = b.to_dense() != 0
    assert np.allclose(ga[mask_a], dA[mask_a], rtol=1e-12, atol=1e-12)
    assert np.allclose(gb[mask_b], dB[mask_b], rtol=1e-12, atol=1e-12)


def test_golden_integer_fixture():
    """Integer-valued operands: every FP64 product and sum is exact, so the
    expected C (committed, computed with Python integers) is bit-exact."""
    fx = np.load("tests/golden/integer_product.npz")
    a = S.BlockCSR(fx["blk"], fx["a_row_p"], fx["a_col"], fx["a_off"], fx["a_data"])
    b = S.BlockCSR(fx["blk"], fx["b_row_p"], fx["b_col"], fx["b_off"], fx["b_data"])
    c, st = oracle.multiply(a, b, eps=float(fx["eps"]))
    assert np.array_equal(c.row_p, fx["c_row_p"])
    assert np.array_equal(c.col, fx["c_col"])
    assert np.array_equal(c.data, fx["c_data"])
    assert st["n_triples_kept"] == int(fx["kept"])
