"""The plexus restatement (oracle/plexus.py) against hand-derived answers
and structural invariants.

The reference's own check, lamb_plexus_verify (src/lamb_plexus_types.F90:
416-428), is empty, and its loop over P in {1, 3, 6, 15, 28}
(src/lamb_plexus_tests.F90:80-111) is switched off. Known answers below are
worked by hand from src/lamb_plexus_operations.F90:87-571:

P = 1: two regions, realms (1,1), (1,2), (2,2) all on process 0; sectors
(1,1), (1,2), (2,1), (2,2) -- the order the P = 1 multiply loop sees.

P = 3: four regions, ten realms in walk order (1,1) (1,2) (2,2) (1,3) (1,4)
(2,4) (2,3) (3,3) (3,4) (4,4); realm t goes to int((t - 1) * 0.3f): t = 1..4
to process 0, 5..7 to 1, 8..10 to 2. Pools {0,1} {0,1} {0,1,2} {1,2}; the
greedy map (ties to the highest process) and the sectors follow.
"""
import numpy as np
import pytest

from lamb_amd.matrix import n_cubes_for
from oracle import plexus as PX


def test_p1_known_answer():
    px = PX.Plexus(1, 0)
    assert px.n_regions == 2
    assert [(r, s) for r, s, _ in px.realms] == [(1, 1), (1, 2), (2, 2)]
    assert px.pools == {1: [0], 2: [0]}
    assert px.catalog == [1, 2] and px.orbit == [0]
    assert px.sectors == [(1, 1), (1, 2), (2, 1), (2, 2)]
    assert px.burden == [6]


def test_p3_known_answer():
    assert PX.n_regions_for(3) == 4
    walk = [(1, 1), (1, 2), (2, 2), (1, 3), (1, 4), (2, 4), (2, 3), (3, 3), (3, 4), (4, 4)]
    procs = [0, 0, 0, 0, 1, 1, 1, 2, 2, 2]
    want_map = np.array([[1, 0, 1, 1],
                         [0, 0, 0, 1],
                         [1, 0, 2, 2],
                         [1, 1, 2, 2]])
    want_sectors = {0: [(1, 2), (2, 1), (2, 2), (2, 3), (3, 2)],
                    1: [(1, 1), (1, 3), (1, 4), (2, 4), (3, 1), (4, 1), (4, 2)],
                    2: [(3, 3), (3, 4), (4, 3), (4, 4)]}
    want_catalog = {0: [1, 2, 3], 1: [1, 2, 3, 4], 2: [3, 4]}
    for p in range(3):
        px = PX.Plexus(3, p)
        assert [(r, s) for r, s, _ in px.realms] == walk
        assert [q for _, _, q in px.realms] == procs
        assert px.pools == {1: [0, 1], 2: [0, 1], 3: [0, 1, 2], 4: [1, 2]}
        assert np.array_equal(px.map[1:, 1:], want_map)
        assert px.burden == [6, 8, 6]
        assert px.catalog == want_catalog[p]
        assert px.orbit == [0, 1, 2]
        assert px.sectors == want_sectors[p]
    # P = 3, process 0: sector (1,2) fetches A(1,k) and B(k,2) for k = 1..4
    sched = PX.Plexus(3, 0).schedule()
    assert sched[:8] == [(1, 1, 0, 1, 1, 1), (1, 0, 0, 2, 1, 2),
                         (2, 0, 0, 1, 1, 2), (2, 0, 0, 2, 2, 2),
                         (3, 1, 0, 1, 1, 3), (3, 0, 0, 2, 3, 2),
                         (4, 1, 0, 1, 1, 4), (4, 1, 0, 2, 4, 2)]
    assert len(sched) == 2 * 5 * 4


@pytest.mark.parametrize("n_procs", [1, 3, 6, 15, 28, 2, 7, 40])
def test_invariants(n_procs):
    nr = PX.n_regions_for(n_procs)
    assert nr == n_cubes_for(n_procs)
    plexi = [PX.Plexus(n_procs, p) for p in range(n_procs)]
    px0 = plexi[0]
    # every realm once, in walk order, on a process of both pools
    assert len(px0.realms) == nr * (nr + 1) // 2
    assert sorted((r, s) for r, s, _ in px0.realms) == [(r, s) for r in range(1, nr + 1)
                                                         for s in range(r, nr + 1)]
    for r in range(1, nr + 1):
        pool = px0.pools[r]
        assert pool == sorted(set(pool)) and pool
        for s in range(1, nr + 1):
            q = px0.map[r, s]
            assert q == px0.map[s, r]
            assert q in px0.pools[r] and q in px0.pools[s]
    # burdens: 2 per realm owned
    owned = np.bincount([px0.map[r, s] for r in range(1, nr + 1) for s in range(r, nr + 1)],
                        minlength=n_procs)
    assert px0.burden == list(2 * owned)
    # catalogs and pools are dual; orbits are the union of the catalog's pools
    for p, px in enumerate(plexi):
        assert px.catalog == [r for r in range(1, nr + 1) if p in px0.pools[r]]
        assert px.orbit == sorted(set(q for r in px.catalog for q in px0.pools[r]))
        assert all(s == t for s, t in zip(px.sectors, sorted(px.sectors)))
    # the processes' sectors cover every (r, s) exactly once
    allsec = [s for px in plexi for s in px.sectors]
    assert sorted(allsec) == [(r, s) for r in range(1, nr + 1) for s in range(1, nr + 1)]
    # the schedule asks the owner of every operand sector
    for p, px in enumerate(plexi[:3]):
        for era, src, dst, which, a, b in px.schedule():
            assert dst == p and src == px0.map[a, b]
