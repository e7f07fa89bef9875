"""Host-side mirror of the reference's layout, pinned by known answers read
off the reference's source text, and the C-ABI symbol table."""
import ctypes
import re

import numpy as np
import pytest

from lamb_amd import matrix as M
from lamb_amd import systems as S


def test_index_compress_known_answer():
    # rows [1,1,3] over 3 rows -> row_p [0,2,2,3] (src/lamb_index_methods.F90:248-274)
    assert M.index_compress(np.array([1, 1, 3]), 3).tolist() == [0, 2, 2, 3]
    assert M.index_compress(np.array([], dtype=np.int64), 2).tolist() == [0, 0, 0]


@pytest.mark.parametrize("n,dims", [(1, (1, 1, 1)), (2, (2, 1, 1)), (8, (2, 2, 2)),
                                    (12, (2, 2, 3)), (60, (10, 2, 3))])
def test_partition_dims(n, dims):
    # cube counts used by src/lamb_partition_tests.F90:72-73
    assert M.partition_dims(n) == dims


@pytest.mark.parametrize("p,n", [(1, 2), (3, 4), (6, 6), (10, 8)])
def test_n_cubes(p, n):
    assert M.n_cubes_for(p) == n


def test_which_cube_and_order_map():
    coords = np.array([[1.0, 0.0, 0.0], [9.0, 1.0, 1.0], [2.0, 5.0, 5.0],
                       [9.9, 9.9, 9.9], [5.0, 0.0, 0.0]], np.float32)
    dist = M.Distribution(np.array([3, 2, 1, 4, 5]), coords, 10.0)
    # P = 1 -> 2 cubes along x: x < 5 -> subset 1, x >= 5 -> subset 2
    assert dist.subset_map.tolist() == [1, 2, 1, 2, 2]
    # local index = arrival order inside each subset
    assert dist.order_map.tolist() == [1, 1, 2, 2, 3]
    assert dist.sectors == [(1, 1), (1, 2), (2, 1), (2, 2)]
    assert dist.sector_index((2, 1)) == 3 and dist.sector_index((3, 1)) == 0


def test_get_blk_n():
    row_p = np.array([0, 3, 3, 5])
    col_i = np.array([1, 4, 7, 2, 9], np.int16)
    assert M.get_blk_n(row_p, col_i, 1, 4) == (2, True)
    assert M.get_blk_n(row_p, col_i, 3, 9) == (5, True)
    assert M.get_blk_n(row_p, col_i, 1, 5)[1] is False
    assert M.get_blk_n(row_p, col_i, 2, 1)[1] is False


def test_sector_round_trip_is_exact():
    atoms, a = S.small_system(40, 7, sizes=(5, 13, 23))
    dist = M.Distribution(atoms.blk, atoms.pos, atoms.box)
    lm = M.LambMatrix.from_global(dist, a)
    assert len(lm.sectors) == 4
    for sec in lm.sectors:
        # reference conventions: 0-based exclusive row_p, ascending 1-based col_i,
        # 1-based inclusive extents back to back
        assert sec.row_p[0] == 0 and sec.row_p[-1] == sec.n_blks
        for r in range(len(sec.row_p) - 1):
            row = sec.col_i[sec.row_p[r]:sec.row_p[r + 1]]
            assert np.all(np.diff(row) > 0) and (len(row) == 0 or row.min() >= 1)
        if sec.n_blks:
            assert sec.extents[0, 0] == 1 and sec.extents[-1, 1] == len(sec.data)
    back = lm.to_global()
    assert np.array_equal(back.row_p, a.row_p) and np.array_equal(back.col, a.col)
    assert np.array_equal(back.data, a.data)


def test_c_abi_exports_every_declared_symbol():
    from lamb_amd import _native
    header = open("include/lamb_hip.h").read()
    declared = set(re.findall(r"\b(lamb_hip_\w+)\s*\(", header))
    assert declared, "no declarations found"
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    assert declared == set(_native.SIGNATURES)
