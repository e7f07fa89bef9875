"""The oracle's McWeeny chain and its helpers, pinned by dense linear
algebra: the chain against a dense numpy chain (no filtering), fixed points
(an idempotent P maps to itself), convergence of the trace to the count of
eigenvalues above 1/2 of P0, Gershgorin bounds enclosing the spectrum, and
the batch split of a block-diagonal matrix.
"""
import numpy as np

import oracle
from lamb_amd import systems as S


def dense_chain(p0, steps):
    p = p0.copy()
    tr, idem = [], []
    for _ in range(steps):
        p2 = p @ p
        idem.append(np.linalg.norm(p2 - p))
        p = 3 * p2 - 2 * p2 @ p
        tr.append(np.trace(p))
    return p, np.array(tr), np.array(idem)


def test_chain_matches_dense():
    atoms = S.water_system(60, seed=3)
    h = S.decay_matrix(atoms, 8.0, 0.4, 4, symmetric=True)
    p0 = oracle.purification_start(h)
    d0 = p0.to_dense()
    w = np.linalg.eigvalsh(d0)
    assert w.min() >= -1e-12 and w.max() <= 1 + 1e-12  # Gershgorin scaling into [0, 1]
    p, tr, idem = oracle.mcweeny(p0, 6)
    dp, dtr, didem = dense_chain(d0, 6)
    assert np.linalg.norm(p.to_dense() - dp) <= 1e-12 * np.linalg.norm(dp)
    assert np.allclose(tr, dtr, rtol=1e-12) and np.allclose(idem, didem, rtol=1e-10)


def test_projector_is_a_fixed_point():
    atoms = S.water_system(30, seed=5)
    h = S.decay_matrix(atoms, 30.0, 0.4, 6, symmetric=True)  # cutoff beyond the box: dense
    d = h.to_dense()
    w, v = np.linalg.eigh(d)
    proj = v[:, w > np.median(w)] @ v[:, w > np.median(w)].T
    pr = S.BlockCSR(h.blk, h.row_p, h.col, h.blk_off, np.zeros_like(h.data))
    # write the projector into the (full) block pattern
    rows = h.rows()
    off = np.concatenate([[0], np.cumsum(h.blk.astype(np.int64))])
    for k in range(h.n_blocks):
        i, j = rows[k], h.col[k]
        blk = proj[off[i]:off[i + 1], off[j]:off[j + 1]]
        pr.data[h.blk_off[k]:h.blk_off[k] + blk.size] = blk.reshape(-1, order="F")
    p, tr, idem = oracle.mcweeny(pr, 1)
    assert np.linalg.norm(p.to_dense() - proj) <= 1e-12 * np.linalg.norm(proj)
    assert idem[0] <= 1e-12 * np.linalg.norm(proj)


def test_trace_converges_to_occupied_count():
    atoms = S.water_system(45, seed=7)
    h = S.decay_matrix(atoms, 8.0, 0.4, 8, symmetric=True)
    p0 = oracle.purification_start(h)
    occupied = int(np.sum(np.linalg.eigvalsh(p0.to_dense()) > 0.5))
    p, tr, idem = oracle.mcweeny(p0, 40)
    assert abs(tr[-1] - occupied) < 1e-8
    assert idem[-1] < 1e-8


def test_batch_is_block_diagonal():
    b = S.config4(n_systems=3, lo=30, hi=60)
    d = b.h.to_dense()
    off = 0
    offs = [0]
    for atoms, h in b.systems:
        off += int(atoms.blk.sum())
        offs.append(off)
    for s, (atoms, h) in enumerate(b.systems):
        blk = d[offs[s]:offs[s + 1], offs[s]:offs[s + 1]]
        assert np.array_equal(blk, h.to_dense())
        d[offs[s]:offs[s + 1], offs[s]:offs[s + 1]] = 0.0
    assert not d.any()
