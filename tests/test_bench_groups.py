"""bench.py's memory-bounded grouping of a batch (config 4 at its full sizes):
every system in exactly one group, every group's estimate within the
budget, largest systems placed first, a system above the budget refused."""
import numpy as np
import pytest

import bench
from lamb_amd import systems as S


def fake_systems(dims):
    """(AtomBox, None) pairs whose matrix dimension is dims[i] (one atom)."""
    out = []
    for d in dims:
        atoms = S.AtomBox(np.zeros((1, 3), np.float32), np.array([d], np.int32), 10.0)
        out.append((atoms, None))
    return out


def need(d):
    return bench.CHAIN_BYTES_PER_DENSE_BYTE * 8.0 * d * d


def test_groups_cover_and_fit():
    rng = np.random.default_rng(5)
    dims = rng.integers(1, 120, 64).tolist()
    budget = 4.0 * need(100)
    groups, est = bench.memory_groups(fake_systems(dims), budget)
    flat = sorted(i for g in groups for i in g)
    assert flat == list(range(len(dims)))
    for g in groups:
        assert g == sorted(g)
        assert sum(est[i] for i in g) <= budget
    assert est == [need(d) for d in dims]


def test_first_fit_decreasing_known_answer():
    # needs 4, 9, 9, 1, 4, 4 (units of need(1)) into groups of 10, largest
    # first, each into the first group with room: 9 -> G0; 9 -> G1; 4 (atom
    # 0) -> G2; 4 (atom 4) -> G2 (room 2 left); 4 (atom 5) -> G3; 1 -> G0
    dims = [2, 3, 3, 1, 2, 2]
    groups, _ = bench.memory_groups(fake_systems(dims), 10.0 * need(1))
    assert groups == [[1, 3], [2], [0, 4], [5]]


def test_system_above_budget_refused():
    with pytest.raises(ValueError):
        bench.memory_groups(fake_systems([10, 50]), need(40))
