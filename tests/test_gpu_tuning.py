"""Tile-shape tuning and the persisted tuning table.

With shape tuning on, the first product of a workload class times the
FP64 per-wave tile shapes and keeps the fastest for the class; the class key
includes the distribution's set of block sizes. A table saved from one
context and loaded into another reproduces every choice, so the second
context's results are bit-identical to the first's without tuning."""
import numpy as np
import pytest

import oracle
from __graft_entry__ import compare
from lamb_amd import device as D
from lamb_amd import systems as S
from lamb_amd.matrix import Distribution, LambMatrix

pytestmark = pytest.mark.gpu


def products():
    # water with a third kind of 23 x 23 blocks (5 / 13 / 23), sparse
    atoms = S.water_system(900, seed=301, heavy_fraction=0.1)
    yield atoms, S.decay_matrix(atoms, 8.0, 0.4, 302), S.decay_matrix(atoms, 8.0, 0.4, 303), 1e-7
    # a denser product of the same kinds (longer cutoff)
    atoms = S.water_system(300, seed=311, heavy_fraction=0.1)
    yield atoms, S.decay_matrix(atoms, 12.0, 0.3, 312), S.decay_matrix(atoms, 12.0, 0.3, 313), 0.0
    # prefilter binds (masked walks), 1 / 5 / 13 blocks
    atoms, a = S.small_system(400, 41, cutoff=9.0, decay=1.5, sizes=(1, 5, 13))
    yield atoms, a, S.decay_matrix(atoms, 9.0, 1.3, 42), 2e-3


def run(ctx, atoms, a, b, eps):
    dist = Distribution(atoms.blk, atoms.pos, atoms.box)
    dd = D.DeviceDistribution(ctx, dist)
    da = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, a))
    db = D.DeviceMatrix.from_lamb(dd, LambMatrix.from_global(dist, b))
    dc = D.DeviceMatrix(dd)
    st = D.multiply(da, db, dc, eps_filter=eps)
    got = dc.download().to_global()
    for h in (dc, db, da, dd):
        h.close()
    return got, st, D.tuning_info(ctx)[0]


def test_shape_tuning_and_table_round_trip(tmp_path):
    path = tmp_path / "tuning.txt"
    first = D.Context(0)
    results = []
    try:
        D.tuning_mode(first, True)
        for atoms, a, b, eps in products():
            got, st, variant = run(first, atoms, a, b, eps)
            again, _, v2 = run(first, atoms, a, b, eps)
            print(f"case: atoms={atoms.n_atoms} eps={eps} shape={variant} tiles={st['n_tiles']}",
                  flush=True)
            assert v2 == variant and np.array_equal(got.data, again.data)
            ref, rst = oracle.multiply(a, b, eps=eps)
            same, err = compare(got, ref)
            assert same and err <= 1e-12, err
            assert st["n_triples_kept"] == rst["n_triples_kept"]
            results.append((got, variant))
        D.tuning_save(first, path)
    finally:
        first.close()
    text = path.read_text().splitlines()
    assert text[0] == "lamb_hip tuning table 1"
    assert sum(line.startswith("shape ") for line in text) == 3
    assert sum(line.startswith("order ") for line in text) >= 3
    second = D.Context(0)  # shape tuning off: the table decides
    try:
        D.tuning_load(second, path)
        n0 = D.tuning_info(second)[1]
        for (atoms, a, b, eps), (ref_got, ref_variant) in zip(products(), results):
            got, _, variant = run(second, atoms, a, b, eps)
            assert variant == ref_variant
            assert np.array_equal(got.row_p, ref_got.row_p) and np.array_equal(got.col, ref_got.col)
            assert np.array_equal(got.data, ref_got.data)
        assert D.tuning_info(second)[1] == n0  # no class needed tuning
    finally:
        second.close()


def test_malformed_tables_refused(tmp_path):
    ctx = D.Context(0)  # an empty tuning table
    try:
        check_refusals(ctx, tmp_path)
    finally:
        ctx.close()


def check_refusals(ctx, tmp_path):
    bad = {
        "header": "not a table\nshape 0 0 0 0 0 2\n",
        "variant": "lamb_hip tuning table 1\nshape 0 1 2 3 4 7\n",
        "rowblock": "lamb_hip tuning table 1\norder 0 1 2 3 0 0\n",
        "form": "lamb_hip tuning table 1\norder 0 1 2 3 8 3\n",
        "fields": "lamb_hip tuning table 1\nshape 0 1 2 3 2\n",
        "trailing": "lamb_hip tuning table 1\norder 0 1 2 3 8 0 9\n",
        "empty": "",
        # a good entry before a bad one: the whole table is refused
        "partly": "lamb_hip tuning table 1\norder 0 1 2 3 8 0\norder 0 1 2 x 8 0\n",
    }
    for name, text in bad.items():
        p = tmp_path / f"{name}.txt"
        p.write_text(text)
        with pytest.raises(D.LambError) as err:
            D.tuning_load(ctx, p)
        assert err.value.type == 100, name  # wrong args
    with pytest.raises(D.LambError) as err:
        D.tuning_load(ctx, tmp_path / "missing.txt")
    assert err.value.type == -2000
    assert D.tuning_info(ctx)[1] == 0  # nothing was loaded
    with pytest.raises(D.LambError) as err:
        ctx.check(ctx.lib.lamb_hip_tuning_mode(ctx.h, 2), "tuning_mode")
    assert err.value.type == 100


def test_shape_tuning_in_the_xcd_ranges_form(monkeypatch):
    """Shape tuning launches each FP64 shape in the forced launch form; in the
    XCD ranges form each of those launches splits the tile order first."""
    monkeypatch.setenv("LAMB_HIP_PERSISTENT", "2")
    ctx = D.Context(0)
    try:
        D.tuning_mode(ctx, True)
        for atoms, a, b, eps in products():
            got, st, variant = run(ctx, atoms, a, b, eps)
            form = D.tuning_launch(ctx)[0]
            print(f"case: atoms={atoms.n_atoms} eps={eps} shape={variant} form={form}", flush=True)
            assert form == 2
            ref, rst = oracle.multiply(a, b, eps=eps)
            same, err = compare(got, ref)
            assert same and err <= 1e-12, err
            assert st["n_triples_kept"] == rst["n_triples_kept"]
    finally:
        ctx.close()
