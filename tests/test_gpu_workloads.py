"""GPU parity at the named workloads' full sizes.

- 6,144-atom water box (2,048 H2O, 39.4 A, O 13 / H 5, r < 8 A, A != B,
  eps 1e-7, seed 2): the HIP multiply against the oracle on the same inputs.
  Bar: identical stored pattern (row_p, col), equal kept and filtered triple
  counts and flops, relative Frobenius error <= 1e-12.
- 64-system batch (water-like systems of 96..1,024 atoms, log-uniform, O 13 /
  H 5 / 23 for 10 % of atoms, r < 8 A, eps 1e-7), 20 McWeeny steps: the
  chain's graph replay equals the eager chain bit for bit; every system's
  idempotency error falls along the chain or its trace settles on an
  integer; the three smallest systems match the oracle chain run on each
  alone (pattern identical, 1e-10 relative).
- 3,072-atom water box, L = tr(P_10 H), checkpoints every 2 steps: the
  gradient through the adjoint kernels against a central finite difference
  (h = 1e-6) along one direction spanning 32 drawn blocks (bar: 1e-5 of
  dL.d plus four times the rounding floor eps |L| / h), and the BF16-split
  gradient against the FP64 one (1e-7 relative).
"""
import numpy as np
import pytest
import torch

import oracle
from __graft_entry__ import compare
from lamb_amd import autograd as AG
from lamb_amd import device as D
from lamb_amd import systems as S
from lamb_amd.matrix import Distribution, LambMatrix

pytestmark = pytest.mark.gpu


def test_water6144_matches_oracle(ctx, monkeypatch):
    """The default kernel, then the opt-in vector-ALU (lanes) and 4 x 4 x 4
    (fine) variants, each against the one oracle product."""
    atoms, a, b, eps = S.config2(2)
    print(f"case: 6,144-atom water box, A {a.n_blocks} blocks, eps {eps}", flush=True)
    dist = Distribution(atoms.blk, atoms.pos, atoms.box)
    dd = D.DeviceDistribution(ctx, dist)
    da = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, a))
    db = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, b))
    runs = []
    for variant in (None, "lanes", "fine"):
        if variant is None:
            monkeypatch.delenv("LAMB_HIP_F64_KERNEL", raising=False)
        else:
            monkeypatch.setenv("LAMB_HIP_F64_KERNEL", variant)
        dc = D.DeviceMatrix(dd)
        st = D.multiply(da, db, dc, eps_filter=eps)
        runs.append((variant or "default", dc.download().to_global(), st))
        dc.close()
        print(f"  gpu ({runs[-1][0]}): {st['n_triples_kept']} triples kept, "
              f"{st['n_c_blocks_kept']} C blocks, numeric {st['ms_numeric']:.2f} ms", flush=True)
    monkeypatch.delenv("LAMB_HIP_F64_KERNEL", raising=False)
    for h in (db, da, dd):
        h.close()
    ref, rst = oracle.multiply(a, b, eps=eps)
    for name, got, st in runs:
        same, err = compare(got, ref)
        print(f"  oracle vs {name}: {rst['n_triples_kept']} triples kept, {ref.n_blocks} C "
              f"blocks; pattern identical {same}, relative error {err:.3e}", flush=True)
        assert same, name
        assert err <= 1e-12, (name, err)
        assert st["n_triples_kept"] == rst["n_triples_kept"]
        assert st["n_triples_filtered"] == rst["n_triples_filtered"]
        assert st["flops"] == rst["flops"]
        assert st["n_c_blocks_kept"] == ref.n_blocks


def _batch_distribution(atoms):
    """The spatial partition with the fewest subsets whose local indices fit
    int16 (a batch of many systems can exceed 32,767 atoms in one subset)."""
    for n_procs in (1, 2, 4, 8, 16, 32, 64):
        try:
            return Distribution(atoms.blk, atoms.pos, atoms.box, n_procs=n_procs)
        except ValueError:
            continue
    raise ValueError("no partition fits the int16 local index")


def test_batch64_mcweeny_graph_and_oracle(ctx):
    steps, eps = 20, 1e-7
    bt = S.config4(seed=4, n_systems=64, lo=96, hi=1024)
    sizes = np.array([x.n_atoms for x, _ in bt.systems])
    print(f"case: batch of {bt.n_systems} systems, {sizes.min()}..{sizes.max()} atoms "
          f"({bt.atoms.n_atoms} in all), {steps} steps", flush=True)
    dist = _batch_distribution(bt.atoms)
    dd = D.DeviceDistribution(ctx, dist)
    dd.set_groups(bt.group, bt.n_systems)
    dh = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, bt.h))
    p = D.purification_start(dh)
    tr, idem, st = D.mcweeny(p, steps, eps_filter=eps)
    print(f"  eager chain: {st['flops']:.3e} flop", flush=True)
    p0 = D.purification_start(dh)
    g = D.ChainGraph(p0, steps, eps_filter=eps)
    out = D.DeviceMatrix(dd)
    tg, ig, ms = g.replay(p0, out)
    print(f"  graph replay {ms:.0f} ms", flush=True)
    assert np.array_equal(tg, tr) and np.array_equal(ig, idem)
    got_eager = p.download().to_global()
    got_graph = out.download().to_global()
    assert np.array_equal(got_eager.row_p, got_graph.row_p)
    assert np.array_equal(got_eager.col, got_graph.col)
    assert np.array_equal(got_eager.data, got_graph.data)
    # each system: idempotency error falls along the chain, or the trace has
    # settled on an integer (converged)
    frac = np.abs(tr[-1] - np.round(tr[-1]))
    falls = idem[-1] < idem[0]
    assert np.all(falls | (frac < 1e-8)), (idem[0], idem[-1], tr[-1])
    print(f"  converged systems: {int(np.sum(idem[-1] < 1e-8))} of {bt.n_systems}; "
          f"largest final idempotency error {idem[-1].max():.3e}", flush=True)
    parts = bt.split(got_eager)
    for sys_i in np.argsort(sizes, kind="stable")[:3]:
        atoms_s, h_s = bt.systems[sys_i]
        want, otr, oidem = oracle.mcweeny(oracle.purification_start(h_s), steps, eps=eps)
        same, err = compare(parts[sys_i], want)
        print(f"  system {sys_i} ({atoms_s.n_atoms} atoms): pattern identical {same}, "
              f"error {err:.3e}", flush=True)
        assert same and err <= 1e-10
        assert np.allclose(tr[:, sys_i], otr, rtol=1e-10, atol=1e-10)
    g.close()
    for h in (out, p0, p, dh, dd):
        h.close()


@pytest.fixture(scope="module")
def water3072(ctx):
    atoms = S.water_system(3072, seed=5)
    h = S.decay_matrix(atoms, 8.0, 0.4, 6, symmetric=True)
    dist = Distribution(atoms.blk, atoms.pos, atoms.box)
    dd = D.DeviceDistribution(ctx, dist)
    dh = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, h))
    lo, hi = (float(x[0]) for x in dh.group_gershgorin())
    hv, ph = AG.leaf(dh)  # dh becomes the pattern of H
    state = {}
    yield atoms, hv, ph, lo, hi, state
    state.clear()
    dh.close()
    dd.close()


def _gradient(water3072, flags):
    atoms, hv, ph, lo, hi, state = water3072
    key = ("grad", flags)
    if key not in state:
        x = hv.clone().requires_grad_(True)
        loss = AG.purification_loss(x, ph, atoms.blk, lo, hi, 10, flags=flags)
        loss.backward()
        state[key] = (float(loss.detach()), x.grad.detach())
    return state[key]


def test_water3072_gradient_finite_difference(water3072):
    atoms, hv, ph, lo, hi, _ = water3072
    print(f"case: 3,072-atom gradient of tr(P_10 H), FD along 32 blocks", flush=True)
    loss, grad = _gradient(water3072, D.FP64)
    rows, cols = ph.block_atoms()
    sz = atoms.blk[rows].astype(np.int64) * atoms.blk[cols]
    offs = np.concatenate([[0], np.cumsum(sz)])
    rng = np.random.default_rng(7)
    picks = rng.choice(len(rows), size=32, replace=False)
    # one direction over the 32 blocks, a unit vector within each
    d = torch.zeros_like(hv)
    for k in picks:
        v = rng.normal(size=int(sz[k]))
        d[offs[k]:offs[k + 1]] = torch.from_numpy(v / np.linalg.norm(v)).to(d.device)
    step = 1e-6
    with torch.no_grad():
        lp = AG.purification_loss(hv + step * d, ph, atoms.blk, lo, hi, 10).item()
        lm = AG.purification_loss(hv - step * d, ph, atoms.blk, lo, hi, 10).item()
    fd = (lp - lm) / (2 * step)
    an = float(torch.dot(grad, d))
    # rounding of the two losses (about eps |L| each) bounds what the
    # difference quotient resolves: eps |L| / h, 7.8e-6 at |L| = 3.5e4
    floor = 4 * np.finfo(np.float64).eps * abs(loss) / step
    print(f"  loss {loss:.12g}; dL.d {an:.10e}, FD {fd:.10e}, |diff| {abs(fd - an):.3e}, "
          f"noise floor {floor:.3e}", flush=True)
    assert abs(fd - an) <= 1e-5 * abs(an) + floor


def test_water3072_gradient_bf16_split(water3072):
    print("case: 3,072-atom gradient, BF16 split (5 slices) against FP64", flush=True)
    _, g64 = _gradient(water3072, D.FP64)
    _, g16 = _gradient(water3072, D.bf16_split(5))
    err = float(torch.linalg.norm(g16 - g64) / torch.linalg.norm(g64))
    print(f"  relative gradient difference {err:.3e}", flush=True)
    assert err <= 1e-7
