"""GPU parity of the batch / purification functions against the oracle.

A batch of independent systems runs as one block-diagonal problem (one
launch per kernel for all systems); every per-group result is checked
against the oracle run on that system alone. Bar: FP64 group functions to
1e-12 relative; McWeeny chains with the same stored pattern per system and
values / traces / idempotency errors to 1e-10 relative after the steps
(rounding differences grow through the chain); BF16-split chains within
1e-8 of the FP64 chain.
"""
import numpy as np
import pytest

import oracle
from __graft_entry__ import compare
from lamb_amd import device as D
from lamb_amd import systems as S
from lamb_amd.matrix import Distribution, LambMatrix

pytestmark = pytest.mark.gpu


def load_batch(ctx, b: S.Batch):
    dist = Distribution(b.atoms.blk, b.atoms.pos, b.atoms.box)
    dd = D.DeviceDistribution(ctx, dist)
    dd.set_groups(b.group, b.n_systems)
    dh = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, b.h))
    return dist, dd, dh


def test_group_functions(ctx):
    b = S.config4(seed=11, n_systems=3, lo=60, hi=200)
    dist, dd, dh = load_batch(ctx, b)
    print(f"case: batch of {b.n_systems}, {b.atoms.n_atoms} atoms", flush=True)
    tr = dh.group_trace()
    nrm = dh.group_norm()
    lo, hi = dh.group_gershgorin()
    for g, (atoms, h) in enumerate(b.systems):
        d = h.to_dense()
        assert tr[g] == pytest.approx(np.trace(d), rel=1e-12, abs=1e-12)
        assert nrm[g] == pytest.approx(np.linalg.norm(d), rel=1e-12)
        olo, ohi = oracle.gershgorin(h)
        assert lo[g] == pytest.approx(olo, rel=1e-12) and hi[g] == pytest.approx(ohi, rel=1e-12)
    p = D.purification_start(dh)
    got = b.split(p.download().to_global())
    for g, (atoms, h) in enumerate(b.systems):
        want = oracle.purification_start(h)
        same, err = compare(got[g], want)
        assert same and err <= 1e-14
    dn = p.group_diff_norm(dh)
    for g, (atoms, h) in enumerate(b.systems):
        want = oracle.diff_norm(oracle.purification_start(h), h)
        assert dn[g] == pytest.approx(want, rel=1e-12)
    for x in (p, dh, dd):
        x.close()


@pytest.mark.parametrize("eps", [0.0, 1e-7])
def test_mcweeny_matches_oracle(ctx, eps):
    b = S.config4(seed=12, n_systems=3, lo=60, hi=150)
    steps = 5
    dist, dd, dh = load_batch(ctx, b)
    print(f"case: mcweeny batch of {b.n_systems}, {b.atoms.n_atoms} atoms, eps={eps}", flush=True)
    p = D.purification_start(dh)
    traces, idem, st = D.mcweeny(p, steps, eps_filter=eps)
    got = b.split(p.download().to_global())
    for g, (atoms, h) in enumerate(b.systems):
        want, otr, oidem = oracle.mcweeny(oracle.purification_start(h), steps, eps=eps)
        same, err = compare(got[g], want)
        assert same, f"system {g}: pattern differs"
        assert err <= 1e-10, (g, err)
        assert np.allclose(traces[:, g], otr, rtol=1e-10, atol=1e-10)
        assert np.allclose(idem[:, g], oidem, rtol=1e-8, atol=1e-12)
    assert st["flops"] > 0
    for x in (p, dh, dd):
        x.close()


def test_mcweeny_converges(ctx):
    atoms = S.water_system(90, seed=13)
    h = S.decay_matrix(atoms, 8.0, 0.4, 14, symmetric=True)
    b = S.batch([(atoms, h)])
    dist, dd, dh = load_batch(ctx, b)
    p = D.purification_start(dh)
    occupied = int(np.sum(np.linalg.eigvalsh(oracle.purification_start(h).to_dense()) > 0.5))
    traces, idem, st = D.mcweeny(p, 40)
    print("trace", traces[-1, 0], "occupied", occupied, "idem", idem[-1, 0], flush=True)
    assert abs(traces[-1, 0] - occupied) < 1e-8
    assert idem[-1, 0] < 1e-8
    for x in (p, dh, dd):
        x.close()


def test_mcweeny_bf16_split(ctx):
    b = S.config4(seed=15, n_systems=2, lo=60, hi=150)
    steps = 5
    dist, dd, dh = load_batch(ctx, b)
    p64 = D.purification_start(dh)
    p16 = D.purification_start(dh)
    t64, i64, _ = D.mcweeny(p64, steps, eps_filter=1e-9)
    t16, i16, _ = D.mcweeny(p16, steps, eps_filter=1e-9, flags=D.bf16_split(5))
    rel = np.max(np.abs(t16 - t64) / np.abs(t64))
    print("bf16 chain: max relative trace difference", rel, flush=True)
    assert rel <= 1e-8
    d = p16.group_diff_norm(p64) / p64.group_norm()
    assert np.all(d <= 1e-8), d
    for x in (p64, p16, dh, dd):
        x.close()
