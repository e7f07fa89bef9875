"""GPU: the sector wire format (lamb_hip_goods_size / _pack / _unpack)
against the numpy restatement of lamb_goods_X_create / lamb_goods_X_unpack.

- Every sector of a seeded water-box matrix over 7 spatial subsets (with a
  7 A cutoff, every sector filled; with 3 A, 14 of 49 empty) packs to exactly the index and data words that the
  restatement builds from the downloaded sector.
- Unpacking those words into a fresh matrix and committing gives back the
  matrix bit for bit, and a multiply over it the same product bit for bit.
- Words packed by the restatement from the host-side sectors unpack to the
  same matrix.
- A header that disagrees with the subset's size, a wrong data length and a
  short pack buffer are refused with an error, touching no device memory.
"""
import ctypes as C

import numpy as np
import pytest

from __graft_entry__ import compare
from lamb_amd import device as D
from lamb_amd import systems as S
from lamb_amd.matrix import Distribution, LambMatrix
from oracle import goods as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[7.0, 3.0], ids=["dense_sectors", "empty_sectors"])
def water(ctx, request):
    atoms = S.water_system(200, seed=71)
    a = S.decay_matrix(atoms, request.param, 0.4, 72)
    dist = Distribution(atoms.blk, atoms.pos, atoms.box, n_procs=8)
    dd = D.DeviceDistribution(ctx, dist)
    lm = LambMatrix.from_global(dist, a)
    dm = D.DeviceMatrix.from_lamb(dd, lm)
    yield dist, dd, a, lm, dm
    dm.close()
    dd.close()


def test_pack_matches_restatement(water):
    dist, dd, a, lm, dm = water
    host = dm.download()
    n_empty = 0
    for sec in host.sectors:
        r, s = sec.subsets
        print(f"case: sector ({r},{s}) blocks={sec.n_blks}", flush=True)
        index, data = dm.goods_pack(r, s)
        ri, rd = G.pack(sec.row_p, sec.col_i, sec.extents, sec.data)
        assert np.array_equal(index, ri)
        assert np.array_equal(data.view(np.int64), rd.view(np.int64))
        n_empty += sec.n_blks == 0
    assert len(host.sectors) > 1
    print(f"sectors={len(host.sectors)} empty={n_empty}", flush=True)


def test_unpack_round_trip_and_multiply(ctx, water):
    dist, dd, a, lm, dm = water
    dm2 = D.DeviceMatrix(dd)
    try:
        for (r, s) in dist.sectors:
            index, data = dm.goods_pack(r, s)
            dm2.goods_unpack(r, s, index, data, commit=False)
        ctx.check(ctx.lib.lamb_hip_matrix_commit(dm2.h), "matrix_commit")
        back, again = dm.download(), dm2.download()
        for s0, s1 in zip(back.sectors, again.sectors):
            assert s0.subsets == s1.subsets
            assert np.array_equal(s0.row_p, s1.row_p) and np.array_equal(s0.col_i, s1.col_i)
            assert np.array_equal(s0.extents, s1.extents)
            assert np.array_equal(s0.data.view(np.int64), s1.data.view(np.int64))
        outs = []
        for src in (dm, dm2):
            dc = D.DeviceMatrix(dd)
            D.multiply(src, src, dc, eps_filter=1e-8)
            outs.append(dc.download().to_global())
            dc.close()
        same, err = compare(outs[1], outs[0])
        assert same and err == 0.0
    finally:
        dm2.close()


def test_unpack_restatement_words(water):
    dist, dd, a, lm, dm = water
    dm2 = D.DeviceMatrix(dd)
    try:
        for sec in lm.sectors:
            r, s = sec.subsets
            dm2.goods_unpack(r, s, *G.pack(sec.row_p, sec.col_i, sec.extents, sec.data),
                             commit=False)
        dm2.ctx.check(dm2.ctx.lib.lamb_hip_matrix_commit(dm2.h), "matrix_commit")
        same, err = compare(dm2.download().to_global(), a)
        assert same and err == 0.0
    finally:
        dm2.close()


def test_malformed_words_refused(water):
    dist, dd, a, lm, dm = water
    sec = max(lm.sectors, key=lambda t: t.n_blks)
    r, s = sec.subsets
    index, data = G.pack(sec.row_p, sec.col_i, sec.extents, sec.data)
    dm2 = D.DeviceMatrix(dd)
    try:
        bad = index.copy()
        bad[5] += 1  # row count of another subset
        with pytest.raises(Exception):
            dm2.goods_unpack(r, s, bad, data, commit=False)
        with pytest.raises(Exception):
            dm2.goods_unpack(r, s, index, data[:-1], commit=False)
        bad = index.copy()
        bad[6] += 1  # col_i offset
        with pytest.raises(Exception):
            dm2.goods_unpack(r, s, bad, data, commit=False)
        lib = dm.ctx.lib
        short = np.zeros(len(index) - 1, np.int64)
        out = np.zeros(len(data))
        rc = lib.lamb_hip_goods_pack(dm.h, r, s, short.ctypes.data_as(C.POINTER(C.c_int64)),
                                     len(short), out.ctypes.data_as(C.POINTER(C.c_double)),
                                     len(out))
        assert rc != 0
    finally:
        dm2.close()
