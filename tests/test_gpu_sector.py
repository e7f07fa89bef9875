"""Sector ownership (plexus) and per-sector products through the C-ABI.

- lamb_hip_plexus_* against the restatement oracle/plexus.py (itself pinned
  by hand-derived answers in tests/test_plexus.py) for P in
  {1, 3, 6, 15, 28}: pools, realm map, burdens, catalogs, orbits, sectors
  and schedules, field by field.
- P = 1 (two subsets): the reference's sector-pair loop
  (src/lamb_multiplication.F90:180-201) -- for every A sector (m, k) and B
  sector (k, n), C(m, n) += A(m, k) B(k, n) -- through lamb_hip_sector_multiply,
  then lamb_hip_matrix_filter, equals lamb_hip_multiply and the oracle
  (pattern identical, 1e-12).
- Forced P = 3 and 6: every process runs the sector products its schedule
  names for its own C sectors; together they give the whole product.
"""
import numpy as np
import pytest

import oracle
from __graft_entry__ import compare
from lamb_amd import device as D
from lamb_amd import systems as S
from lamb_amd.matrix import Distribution, LambMatrix
from oracle import plexus as PX

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_procs", [1, 3, 6, 15, 28])
def test_plexus_matches_restatement(n_procs):
    for p in sorted({0, 1 % n_procs, n_procs // 2, n_procs - 1}):
        got = D.Plexus(n_procs, p)
        want = PX.Plexus(n_procs, p)
        try:
            print(f"case: plexus P={n_procs} process {p}: {got.n_regions} regions, "
                  f"{len(got.sectors)} sectors, orbit {len(got.orbit)}", flush=True)
            assert got.n_regions == want.n_regions
            assert got.pools == want.pools
            assert np.array_equal(got.realm_map, want.map[1:, 1:])
            assert got.burden == want.burden
            assert got.catalog == want.catalog
            assert got.orbit == want.orbit
            assert got.sectors == want.sectors
            assert got.sched == want.schedule()
        finally:
            got.close()


def test_plexus_refuses_bad_arguments():
    for args in ((0, 0, 0), (3, 3, 0), (3, -1, 0), (3, 0, -2), (3, 0, 40000)):
        with pytest.raises(D.LambError):
            D.Plexus(*args)


def _setup(ctx, n_procs, seed):
    atoms = S.water_system(300, seed=seed)
    a = S.decay_matrix(atoms, 7.0, 0.4, seed + 1)
    b = S.decay_matrix(atoms, 7.0, 0.4, seed + 2)
    dist = Distribution(atoms.blk, atoms.pos, atoms.box, n_procs=n_procs)
    dd = D.DeviceDistribution(ctx, dist)
    da = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, a))
    db = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, b))
    return atoms, a, b, dist, dd, da, db


def _check_against_full(dd, da, db, dc, a, b, eps):
    dropped = D.matrix_filter(dc, eps)
    got = dc.download().to_global()
    full = D.DeviceMatrix(dd)
    D.multiply(da, db, full, eps_filter=eps)
    want = full.download().to_global()
    full.close()
    same, err = compare(got, want)
    print(f"  dropped by the filter {dropped}; vs lamb_hip_multiply: pattern {same}, "
          f"error {err:.3e}", flush=True)
    assert same and err <= 1e-12
    ref, _ = oracle.multiply(a, b, eps=eps)
    same, err = compare(got, ref)
    assert same and err <= 1e-12
    return dropped


@pytest.mark.parametrize("eps", [1e-7, 0.1])
def test_p1_sector_pair_loop_equals_multiply(ctx, eps):
    atoms, a, b, dist, dd, da, db, = _setup(ctx, 1, 131)
    assert dist.n_subsets == 2
    sectors = D.Plexus(1, 0).sectors
    assert sectors == dist.sectors == [(1, 1), (1, 2), (2, 1), (2, 2)]
    dc = D.DeviceMatrix(dd)
    pairs = 0
    flops = 0.0
    for (m, k) in sectors:           # A sectors
        for (k2, n) in sectors:      # B sectors
            if k2 != k:
                continue
            st = D.sector_multiply(da, db, dc, m, k, n, eps_filter=eps)
            flops += st["flops"]
            pairs += 1
    print(f"case: P=1 sector-pair loop, eps {eps}: {pairs} sector products, "
          f"{flops:.3e} flop", flush=True)
    assert pairs == 8
    dropped = _check_against_full(dd, da, db, dc, a, b, eps)
    if eps > 1e-6:
        assert dropped > 0  # the filter binds on this case
    for h in (dc, db, da, dd):
        h.close()


@pytest.mark.parametrize("n_procs", [3, 6])
def test_forced_process_schedules_cover_the_product(ctx, n_procs):
    atoms, a, b, dist, dd, da, db = _setup(ctx, n_procs, 141)
    plexi = [D.Plexus(n_procs, p) for p in range(n_procs)]
    assert dist.n_subsets == plexi[0].n_regions
    print(f"case: P={n_procs}, {dist.n_subsets} subsets, sectors per process "
          f"{[len(px.sectors) for px in plexi]}", flush=True)
    dc = D.DeviceMatrix(dd)
    owned = []
    for px in plexi:
        owned += px.sectors
        # one era = the A and B operand sectors of one product
        entries = {}
        for era, src, dst, which, r, s in px.sched:
            assert dst == px.my_proc
            assert src == px.realm_map[r - 1, s - 1]
            entries.setdefault(era, {})[which] = (r, s)
        for era in sorted(entries):
            (m, k), (k2, n) = entries[era][1], entries[era][2]
            assert k == k2 and (m, n) in px.sectors
            D.sector_multiply(da, db, dc, m, k, n, eps_filter=1e-7)
    assert sorted(owned) == [(r, s) for r in range(1, dist.n_subsets + 1)
                             for s in range(1, dist.n_subsets + 1)]
    _check_against_full(dd, da, db, dc, a, b, 1e-7)
    for px in plexi:
        px.close()
    for h in (dc, db, da, dd):
        h.close()
