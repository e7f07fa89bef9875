"""GPU parity of the differentiation form.

- The adjoint kernels (dA = G B^T on A's pattern, dB = A^T G on B's
  pattern) against the CPU FP64 oracle adjoint: 1e-12 relative (FP64),
  1e-9 (BF16 split).
- The gradient of L = tr(P_k H) after k McWeeny steps (torch.autograd
  through the binding, checkpointed every 2 steps) against dense torch CPU
  FP64 autograd of the same loss, restricted to H's pattern: 1e-10
  relative; and against central finite differences (h = 1e-6) on 32 random
  blocks: |FD - dL.d| <= 1e-5 max(|dL.d|, 1) per unit direction d within a
  block (the FD noise floor is ~eps |L| / h, about 4e-7 at |L| ~ 1.6e3).
"""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from __graft_entry__ import compare
from lamb_amd import _native as N
from lamb_amd import autograd as AG
from lamb_amd import device as D
from lamb_amd import systems as S
from lamb_amd.matrix import Distribution, LambMatrix

pytestmark = pytest.mark.gpu


def rand_like_pattern(m: S.BlockCSR, seed: int) -> S.BlockCSR:
    rng = np.random.default_rng(seed)
    return S.BlockCSR(m.blk, m.row_p, m.col, m.blk_off, rng.uniform(-1, 1, m.data.size))


@pytest.mark.parametrize("flags,tol", [(D.FP64, 1e-12), (D.bf16_split(5), 1e-9)])
def test_adjoint_matches_oracle(ctx, flags, tol):
    atoms = S.water_system(150, seed=61, heavy_fraction=0.1)
    a = S.decay_matrix(atoms, 7.0, 0.4, 62)
    b = S.decay_matrix(atoms, 6.0, 0.4, 63)
    c, _ = oracle.multiply(a, b)
    g = rand_like_pattern(c, 64)
    print(f"case: adjoint atoms={atoms.n_atoms} flags={flags}", flush=True)
    want_da, want_db = oracle.adjoint(a, b, g)
    dist = Distribution(atoms.blk, atoms.pos, atoms.box)
    dd = D.DeviceDistribution(ctx, dist)
    da_, db_, dg_ = (D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, x)) for x in (a, b, g))
    dA, dB = D.DeviceMatrix(dd), D.DeviceMatrix(dd)
    st = N.Stats()
    rc = ctx.lib.lamb_hip_multiply_adjoint(ctx.h, da_.h, db_.h, dg_.h, dA.h, dB.h, flags,
                                           C.byref(st))
    ctx.check(rc, "multiply_adjoint")
    got_a = dA.download().to_global()
    got_b = dB.download().to_global()
    assert np.array_equal(got_a.col, a.col) and np.array_equal(got_b.col, b.col)
    ea = np.linalg.norm(got_a.data - want_da) / np.linalg.norm(want_da)
    eb = np.linalg.norm(got_b.data - want_db) / np.linalg.norm(want_db)
    print(f"  dA error {ea:.3e}, dB error {eb:.3e}", flush=True)
    assert ea <= tol and eb <= tol
    for x in (dA, dB, dg_, db_, da_, dd):
        x.close()


def transposed(m: S.BlockCSR) -> S.BlockCSR:
    """X^T in numpy: block (i, j) becomes (j, i), transposed, rows' columns
    ascending."""
    rows = m.rows()
    order = np.lexsort((rows, m.col))  # by new row (old col), then new col
    sizes = m.blk[rows].astype(np.int64) * m.blk[m.col]
    t_rows, t_col = m.col[order], rows[order]
    t_off = np.concatenate([[0], np.cumsum(sizes[order])[:-1]]).astype(np.int64)
    data = np.empty_like(m.data)
    for p, b in enumerate(order):
        blk = m.block(int(b))  # m_i x n_j
        data[t_off[p]:t_off[p] + blk.size] = blk.T.reshape(-1, order="F")
    row_p = np.zeros(m.n_atoms + 1, dtype=np.int64)
    np.add.at(row_p, t_rows.astype(np.int64) + 1, 1)
    return S.BlockCSR(m.blk, np.cumsum(row_p), t_col.astype(np.int32), t_off, data)


@pytest.mark.parametrize("n_procs", [1, 4])
def test_transpose_on_device(ctx, n_procs):
    """lamb_hip_matrix_transpose (index by a stable sort on the device):
    exactly X^T, pattern and values, against numpy; twice gives X back bit
    for bit; the transpose is a valid operand (X^T B against the oracle); an
    empty matrix transposes to an empty one."""
    atoms = S.water_system(150, seed=65, heavy_fraction=0.1)
    a = S.decay_matrix(atoms, 7.0, 0.4, 66)
    b = S.decay_matrix(atoms, 6.0, 0.4, 67)
    want = transposed(a)
    print(f"case: transpose atoms={atoms.n_atoms} blocks={a.n_blocks} n_procs={n_procs}",
          flush=True)
    dist = Distribution(atoms.blk, atoms.pos, atoms.box, n_procs=n_procs)
    dd = D.DeviceDistribution(ctx, dist)
    da = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, a))
    db = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, b))
    dt, dtt = D.DeviceMatrix(dd), D.DeviceMatrix(dd)
    ctx.check(ctx.lib.lamb_hip_matrix_transpose(da.h, dt.h), "transpose")
    got = dt.download().to_global()
    assert np.array_equal(got.row_p, want.row_p) and np.array_equal(got.col, want.col)
    assert np.array_equal(got.to_dense(), a.to_dense().T)
    assert np.array_equal(got.data, want.data)
    ctx.check(ctx.lib.lamb_hip_matrix_transpose(dt.h, dtt.h), "transpose")
    back = dtt.download().to_global()
    ref_a = da.download().to_global()
    assert np.array_equal(back.row_p, ref_a.row_p) and np.array_equal(back.col, ref_a.col)
    assert np.array_equal(back.data, ref_a.data)
    dc = D.DeviceMatrix(dd)
    D.multiply(dt, db, dc, eps_filter=1e-7)
    ref, _ = oracle.multiply(want, b, eps=1e-7)
    same, err = compare(dc.download().to_global(), ref)
    print(f"  X^T B vs oracle: pattern {same}, error {err:.3e}", flush=True)
    assert same and err <= 1e-12
    empty = S.BlockCSR(a.blk, np.zeros(a.n_atoms + 1, dtype=np.int64), np.zeros(0, np.int32),
                       np.zeros(0, np.int64), np.zeros(0))
    de = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, empty))
    ctx.check(ctx.lib.lamb_hip_matrix_transpose(de.h, dt.h), "transpose (empty)")
    got = dt.download().to_global()
    assert got.n_blocks == 0 and not got.row_p.any()
    for x in (de, dc, dtt, dt, db, da, dd):
        x.close()


def setup_h(ctx, atoms, h):
    dist = Distribution(atoms.blk, atoms.pos, atoms.box)
    dd = D.DeviceDistribution(ctx, dist)
    dh = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, h))
    lo, hi = dh.group_gershgorin()
    hv, ph = AG.leaf(dh)
    return dd, hv, ph, float(lo[0]), float(hi[0])


def grad_blocks(ph, g, blk):
    """{(row atom, col atom): gradient block (column-major flat)}."""
    rows, cols = ph.block_atoms()
    sizes = blk[rows].astype(np.int64) * blk[cols]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    g = g.detach().cpu().numpy()
    return {(int(r), int(c)): g[offs[k]:offs[k + 1]] for k, (r, c) in enumerate(zip(rows, cols))}


def test_chain_gradient_matches_dense_torch(ctx):
    atoms = S.water_system(60, seed=71)
    h = S.decay_matrix(atoms, 8.0, 0.4, 72, symmetric=True)
    steps = 6
    dd, hv, ph, lo, hi = setup_h(ctx, atoms, h)
    print(f"case: chain gradient atoms={atoms.n_atoms} steps={steps}", flush=True)
    hv = hv.clone().requires_grad_(True)
    loss = AG.purification_loss(hv, ph, atoms.blk, lo, hi, steps)
    loss.backward()
    # dense CPU FP64 reference of the same loss
    H = torch.tensor(h.to_dense(), dtype=torch.float64, requires_grad=True)
    n = H.shape[0]
    P = (H - lo * torch.eye(n, dtype=torch.float64)) / (hi - lo)
    for _ in range(steps):
        P2 = P @ P
        P = 3 * P2 - 2 * P2 @ P
    ref_loss = (P * H).sum()
    ref_loss.backward()
    assert loss.item() == pytest.approx(ref_loss.item(), rel=1e-11)
    G = H.grad.numpy()
    off = np.concatenate([[0], np.cumsum(atoms.blk.astype(np.int64))])
    got = grad_blocks(ph, hv.grad, atoms.blk)
    num = den = 0.0
    for (i, j), gb in got.items():
        want = G[off[i]:off[i + 1], off[j]:off[j + 1]].reshape(-1, order="F")
        num += float(np.sum((gb - want) ** 2))
        den += float(np.sum(want ** 2))
    err = np.sqrt(num / den)
    print(f"  loss {loss.item():.12g}, gradient error vs dense torch {err:.3e}", flush=True)
    assert err <= 1e-10
    dd.close()


def test_chain_gradient_finite_differences(ctx):
    atoms = S.water_system(240, seed=81)
    h = S.decay_matrix(atoms, 8.0, 0.4, 82, symmetric=True)
    steps = 10
    dd, hv, ph, lo, hi = setup_h(ctx, atoms, h)
    print(f"case: chain FD atoms={atoms.n_atoms} steps={steps}", flush=True)
    hg = hv.clone().requires_grad_(True)
    AG.purification_loss(hg, ph, atoms.blk, lo, hi, steps).backward()
    grad = hg.grad.detach()
    rows, cols = ph.block_atoms()
    sizes = atoms.blk[rows].astype(np.int64) * atoms.blk[cols]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    rng = np.random.default_rng(83)
    picks = rng.choice(len(rows), size=32, replace=False)
    step = 1e-6
    worst = worst_abs = 0.0
    with torch.no_grad():
        for k in picks:
            d = torch.zeros_like(hv)
            v = rng.normal(size=int(sizes[k]))
            d[offs[k]:offs[k + 1]] = torch.from_numpy(v / np.linalg.norm(v)).to(d.device)
            lp = AG.purification_loss(hv + step * d, ph, atoms.blk, lo, hi, steps).item()
            lm = AG.purification_loss(hv - step * d, ph, atoms.blk, lo, hi, steps).item()
            fd = (lp - lm) / (2 * step)
            an = float(torch.dot(grad, d))
            worst = max(worst, abs(fd - an) / max(abs(an), 1.0))
            worst_abs = max(worst_abs, abs(fd - an))
    print(f"  32 blocks: worst |FD - dL.d| {worst_abs:.3e}, scaled {worst:.3e}", flush=True)
    assert worst <= 1e-5
    dd.close()


def test_chain_gradient_bf16_split(ctx):
    atoms = S.water_system(90, seed=91)
    h = S.decay_matrix(atoms, 8.0, 0.4, 92, symmetric=True)
    steps = 6
    dd, hv, ph, lo, hi = setup_h(ctx, atoms, h)
    g = {}
    for name, flags in (("fp64", D.FP64), ("bf16", D.bf16_split(5))):
        x = hv.clone().requires_grad_(True)
        AG.purification_loss(x, ph, atoms.blk, lo, hi, steps, flags=flags).backward()
        g[name] = x.grad.detach()
    err = float(torch.linalg.norm(g["bf16"] - g["fp64"]) / torch.linalg.norm(g["fp64"]))
    print(f"case: bf16 chain gradient relative difference {err:.3e}", flush=True)
    assert err <= 1e-8
    dd.close()
