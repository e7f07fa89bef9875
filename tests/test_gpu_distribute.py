"""GPU: the block-list entry (lamb_hip_distribute, the reference's
internal_distribute, src/lamb_matrix_operations.F90:88-375) and the checks
that sector_upload / goods_unpack / commit apply to caller arrays.

- Blocks in shuffled order spread over three data stores (three calls, as
  data_areas(:) with datastore_i at src/lamb_matrix_operations.F90:88-97)
  commit to the same matrix as a sector upload, bit for bit.
- Out-of-range subset or local indices, extents outside their store and two
  blocks sharing values in one store are refused with type 100
  (lamb_wrong_args, src/lamb_error.F90:141-159).
- A row_p that is not a non-decreasing sequence inside [0, n_blks]
  (the CSR contract of src/lamb_index_methods.F90:73-107,248-274) is refused
  by sector_upload and, through it, by goods_unpack.
"""
import ctypes as C

import numpy as np
import pytest

from __graft_entry__ import compare
from lamb_amd import _native as N
from lamb_amd import device as D
from lamb_amd import systems as S
from lamb_amd.matrix import Distribution, LambMatrix
from oracle import goods as G

pytestmark = pytest.mark.gpu

WRONG_ARGS = 100


@pytest.fixture(scope="module")
def water(ctx):
    atoms = S.water_system(150, seed=91)
    a = S.decay_matrix(atoms, 6.0, 0.4, 92)
    dist = Distribution(atoms.blk, atoms.pos, atoms.box, n_procs=8)
    dd = D.DeviceDistribution(ctx, dist)
    yield dist, dd, a
    dd.close()


def _block_list(dist, a):
    """(subset_idx [n,2] int32, local_idx [n,2] int16, sizes, offsets) of a
    global BlockCSR's blocks in CSR order."""
    rows = a.rows().astype(np.int64)
    cols = a.col.astype(np.int64)
    sub = np.stack([dist.subset_map[rows], dist.subset_map[cols]], 1).astype(np.int32)
    loc = np.stack([dist.order_map[rows], dist.order_map[cols]], 1).astype(np.int16)
    sizes = (a.blk[rows] * a.blk[cols]).astype(np.int64)
    return sub, loc, sizes, a.blk_off.astype(np.int64)


def _distribute(dm, sub, loc, ext, data):
    lib = dm.ctx.lib
    sub = np.ascontiguousarray(sub, np.int32).reshape(-1)
    loc = np.ascontiguousarray(loc, np.int16).reshape(-1)
    ext = np.ascontiguousarray(ext, np.int64).reshape(-1)
    data = np.ascontiguousarray(data, np.float64)
    rc = lib.lamb_hip_distribute(dm.h, len(loc) // 2, N.ptr(sub, C.c_int32),
                                 N.ptr(loc, C.c_int16), N.ptr(ext, C.c_int64),
                                 N.ptr(data, C.c_double), len(data))
    dm.ctx.check(rc, "distribute")


def _stores(dist, a, n_stores, seed):
    """The blocks of a, shuffled and dealt over n_stores data areas, each
    with its own 1-based extents."""
    sub, loc, sizes, offs = _block_list(dist, a)
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(sizes))
    which = rng.integers(0, n_stores, len(sizes))
    out = []
    for s in range(n_stores):
        sel = order[which[order] == s]
        sz = sizes[sel]
        ends = np.cumsum(sz)
        ext = np.stack([ends - sz + 1, ends], 1)
        data = (np.concatenate([a.data[o:o + z] for o, z in zip(offs[sel], sz)])
                if len(sel) else np.zeros(0))
        out.append((sub[sel], loc[sel], ext, data))
    return out


def test_distribute_three_stores_shuffled(water):
    dist, dd, a = water
    dm = D.DeviceMatrix(dd)
    try:
        stores = _stores(dist, a, 3, seed=5)
        print(f"case: {a.n_blocks} blocks in stores of "
              f"{[len(s[0]) for s in stores]}", flush=True)
        for st in stores:
            _distribute(dm, *st)
        dm.ctx.check(dm.ctx.lib.lamb_hip_matrix_commit(dm.h), "commit")
        same, err = compare(dm.download().to_global(), a)
        assert same and err == 0.0
    finally:
        dm.close()


@pytest.mark.parametrize("what", ["local_high", "local_zero", "subset_high", "extent_past",
                                  "extent_short", "overlap"])
def test_distribute_refuses(water, what):
    dist, dd, a = water
    sub, loc, ext, data = _stores(dist, a, 1, seed=6)[0]
    sub, loc, ext = sub.copy(), loc.copy(), ext.copy()
    if what == "local_high":
        loc[3, 1] = len(dist.subset_atoms[sub[3, 1] - 1]) + 1
    elif what == "local_zero":
        loc[3, 0] = 0
    elif what == "subset_high":
        sub[3, 0] = dist.n_subsets + 1
    elif what == "extent_past":
        ext[-1] += 1
    elif what == "extent_short":
        ext[3, 1] -= 1
    elif what == "overlap":
        # block 1 takes block 0's values (same size: the extent stays whole)
        same = np.flatnonzero((ext[:, 1] - ext[:, 0]) == (ext[0, 1] - ext[0, 0]))
        ext[same[1]] = ext[0]
    dm = D.DeviceMatrix(dd)
    try:
        print(f"case: distribute {what}", flush=True)
        _distribute(dm, sub, loc, ext, data)
        with pytest.raises(D.LambError) as e:
            dm.ctx.check(dm.ctx.lib.lamb_hip_matrix_commit(dm.h), "commit")
        assert e.value.type == WRONG_ARGS
    finally:
        dm.close()


def _upload(dm, sec, row_p):
    lib = dm.ctx.lib
    r, s = sec.subsets
    row_p = np.ascontiguousarray(row_p, np.int64)
    col_i = np.ascontiguousarray(sec.col_i, np.int16)
    ext = np.ascontiguousarray(sec.extents, np.int64).reshape(-1)
    data = np.ascontiguousarray(sec.data, np.float64)
    return lib.lamb_hip_sector_upload(dm.h, r, s, sec.n_blks, N.ptr(row_p, C.c_int64),
                                      N.ptr(col_i, C.c_int16), N.ptr(ext, C.c_int64),
                                      N.ptr(data, C.c_double), len(data))


def _bad_row_ps(row_p, n_blks):
    """row_p variants that keep row_p[0] = 0 and row_p[-1] = n_blks."""
    n = len(row_p) - 1
    assert n >= 3 and n_blks >= 2
    up = row_p.copy()
    up[1] = n_blks + 5            # a middle entry above n_blks
    down = row_p.copy()
    down[1], down[2] = n_blks, 1  # decreasing
    neg = row_p.copy()
    neg[1] = -1
    return {"above": up, "decreasing": down, "negative": neg}


def test_upload_refuses_bad_row_p(water):
    dist, dd, a = water
    lm = LambMatrix.from_global(dist, a)
    sec = max(lm.sectors, key=lambda t: t.n_blks)
    dm = D.DeviceMatrix(dd)
    try:
        for name, rp in _bad_row_ps(sec.row_p, sec.n_blks).items():
            print(f"case: upload row_p {name}", flush=True)
            rc = _upload(dm, sec, rp)
            assert rc != 0
            lvl, typ = C.c_int32(), C.c_int32()
            dm.ctx.lib.lamb_hip_last_error(dm.ctx.h, C.byref(lvl), C.byref(typ), None, 0)
            assert typ.value == WRONG_ARGS
        # the good one still goes through and commits
        assert _upload(dm, sec, sec.row_p) == 0
        dm.ctx.check(dm.ctx.lib.lamb_hip_matrix_commit(dm.h), "commit")
        assert dm.info()[0] == sec.n_blks
    finally:
        dm.close()


def test_goods_unpack_refuses_bad_row_p(water):
    dist, dd, a = water
    lm = LambMatrix.from_global(dist, a)
    sec = max(lm.sectors, key=lambda t: t.n_blks)
    r, s = sec.subsets
    dm = D.DeviceMatrix(dd)
    try:
        for name, rp in _bad_row_ps(sec.row_p, sec.n_blks).items():
            print(f"case: goods row_p {name}", flush=True)
            index, data = G.pack(rp, sec.col_i, sec.extents, sec.data)
            with pytest.raises(D.LambError) as e:
                dm.goods_unpack(r, s, index, data, commit=False)
            assert e.value.type == WRONG_ARGS
    finally:
        dm.close()
