"""GPU atom partition (lamb_hip_partition_atoms) against the host mirror of
lamb_set_partition_update (lamb_amd/matrix.py Distribution, pinned by the
known answers in test_layout.py): identical subset and order maps, exactly.
"""
import numpy as np
import pytest

from lamb_amd import device as D
from lamb_amd import matrix as M
from lamb_amd import systems as S

pytestmark = pytest.mark.gpu


def test_known_answer(ctx):
    coords = np.array([[1.0, 0.0, 0.0], [9.0, 1.0, 1.0], [2.0, 5.0, 5.0],
                       [9.9, 9.9, 9.9], [5.0, 0.0, 0.0]], np.float32)
    subset, order = D.partition_atoms(ctx, coords, 10.0, (2, 1, 1))
    assert subset.tolist() == [1, 2, 1, 2, 2]
    assert order.tolist() == [1, 1, 2, 2, 3]


@pytest.mark.parametrize("n_procs", [1, 8, 60, 500])
def test_matches_host_partition(ctx, n_procs):
    atoms = S.water_system(3000, seed=51)
    dist = M.Distribution(atoms.blk, atoms.pos, atoms.box, n_procs=n_procs)
    print(f"case: partition n_procs={n_procs} dims={dist.dims}", flush=True)
    subset, order = D.partition_atoms(ctx, dist.coords, dist.box, dist.dims)
    assert np.array_equal(subset, dist.subset_map)
    assert np.array_equal(order, dist.order_map)


def test_unwrapped_coordinates_are_refused(ctx):
    # x = -7: int(-7 * 0.2) = -1, mod(-1, 2) = -1 -> subset 0 (outside 1..2)
    coords = np.array([[1.0, 1.0, 1.0], [-7.0, 1.0, 1.0]], np.float32)
    with pytest.raises(D.LambError) as e:
        D.partition_atoms(ctx, coords, 10.0, (2, 1, 1))
    assert e.value.type == 100
