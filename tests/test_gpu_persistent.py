"""Persistent launch forms of the numeric kernels: a grid of the device's
resident workgroups, each looping over tile slots of the tile order, in place
of one workgroup per tile; the XCDs either share one window of the order or
each walk their own range of it. Every tile is computed the same way in all
forms, so each kernel variant must give the same bits in each."""
import numpy as np
import pytest

import oracle
from __graft_entry__ import compare
from lamb_amd import device as D
from lamb_amd import systems as S
from lamb_amd.matrix import Distribution, LambMatrix

pytestmark = pytest.mark.gpu


def run(ctx, atoms, a, b, eps, flags):
    dist = Distribution(atoms.blk, atoms.pos, atoms.box)
    dd = D.DeviceDistribution(ctx, dist)
    da = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, a))
    db = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, b))
    dc = D.DeviceMatrix(dd)
    st = D.multiply(da, db, dc, eps_filter=eps, flags=flags)
    got = dc.download().to_global()
    launch = D.tuning_launch(ctx)
    for h in (dc, db, da, dd):
        h.close()
    return got, st, launch


def cases():
    # whole entries (unmasked walks): about 2,000 water atoms, more tiles
    # than the device holds workgroups at once
    atoms = S.water_system(2000, seed=201)
    yield "whole", atoms, S.decay_matrix(atoms, 8.0, 0.4, 202), S.decay_matrix(atoms, 8.0, 0.4, 203), 1e-7
    # the prefilter cuts entries (masked walks)
    atoms, a = S.small_system(1500, 31, cutoff=9.0, decay=1.5, sizes=(1, 5, 13))
    yield "partial", atoms, a, S.decay_matrix(atoms, 9.0, 1.3, 32), 2e-3


@pytest.mark.parametrize("variant", ["quadrant", "half", "ksplit", "bf16"])
def test_persistent_forms_give_the_same_bits(ctx, monkeypatch, variant):
    flags = D.BF16_SPLIT if variant == "bf16" else D.FP64
    if variant != "bf16":
        monkeypatch.setenv("LAMB_HIP_F64_KERNEL", variant)
    for name, atoms, a, b, eps in cases():
        print(f"case: {variant} {name} atoms={atoms.n_atoms} eps={eps}", flush=True)
        monkeypatch.setenv("LAMB_HIP_PERSISTENT", "0")
        g0, st0, (f0, w0, t0) = run(ctx, atoms, a, b, eps, flags)
        assert f0 == 0 and w0 == t0
        for form in (1, 2):
            monkeypatch.setenv("LAMB_HIP_PERSISTENT", str(form))
            g1, st1, (f1, w1, t1) = run(ctx, atoms, a, b, eps, flags)
            print(f"  form {form}: {w1} workgroups for {t1} tiles (grid form {w0})", flush=True)
            assert f1 == form and 0 < w1 < t1 == t0  # the resident workgroups loop
            assert np.array_equal(g0.row_p, g1.row_p) and np.array_equal(g0.col, g1.col)
            assert np.array_equal(g0.data, g1.data)
            assert st0["n_triples_kept"] == st1["n_triples_kept"]
        if name == "partial":
            assert st0["n_triples_filtered"] > 0
        ref, rst = oracle.multiply(a, b, eps=eps)
        same, err = compare(g1, ref)
        assert same and st1["n_triples_kept"] == rst["n_triples_kept"]
        assert err <= (1e-9 if variant == "bf16" else 1e-12), err


@pytest.mark.parametrize("n_atoms", [3, 20, 60])
def test_xcd_ranges_with_fewer_tiles_than_xcds(ctx, monkeypatch, n_atoms):
    """Products of 1 to a few tiles: most XCD ranges are empty, and the grid
    is raised to one workgroup per residue."""
    atoms = S.water_system(n_atoms, seed=221)
    a = S.decay_matrix(atoms, 8.0, 0.4, 222)
    b = S.decay_matrix(atoms, 8.0, 0.4, 223)
    print(f"case: {atoms.n_atoms} atoms", flush=True)
    monkeypatch.setenv("LAMB_HIP_PERSISTENT", "0")
    g0, _, (_, _, t0) = run(ctx, atoms, a, b, 1e-9, D.FP64)
    monkeypatch.setenv("LAMB_HIP_PERSISTENT", "2")
    g2, _, (f2, w2, t2) = run(ctx, atoms, a, b, 1e-9, D.FP64)
    print(f"  {t2} tiles, {w2} workgroups", flush=True)
    assert f2 == 2 and t2 == t0 and w2 >= 8
    assert np.array_equal(g0.data, g2.data) and np.array_equal(g0.col, g2.col)
    ref, _ = oracle.multiply(a, b, eps=1e-9)
    same, err = compare(g2, ref)
    assert same and err <= 1e-12, err


def test_autotuner_keeps_a_launch_form(ctx):
    """The first product of a class times tile orders and launch forms; the
    context then reuses the choice, and the result does not depend on it."""
    atoms = S.water_system(2000, seed=211)
    a = S.decay_matrix(atoms, 8.0, 0.4, 212)
    b = S.decay_matrix(atoms, 8.0, 0.4, 213)
    fresh = D.Context(0)
    try:
        g1, _, l1 = run(fresh, atoms, a, b, 1e-7, D.FP64)
        g2, _, l2 = run(fresh, atoms, a, b, 1e-7, D.FP64)
        print(f"tuned launch: form={l1[0]} workgroups={l1[1]} tiles={l1[2]}", flush=True)
        assert l1 == l2 and D.tuning_info(fresh)[1] == 1
        assert np.array_equal(g1.data, g2.data)
    finally:
        fresh.close()
