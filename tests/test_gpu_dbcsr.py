"""GPU: DBCSR dumps loaded into the device pool (lamb_hip_dbcsr_load).

- The byte-level fixture tests/golden/two_chunks.dbcsr, loaded over a
  distribution of its block sizes, downloads as exactly its blocks.
- Seeded water-box operands written in 4 and 5 chunks and loaded from the
  files multiply to the oracle's product of the original operands: same
  block pattern, 1e-12 relative Frobenius error.
- A file of a symmetric type is refused as unimplemented, as the
  reference's loader does; a file whose block sizes differ from the
  distribution's is refused as wrong arguments.
"""
import json
import pathlib

import numpy as np
import pytest

import oracle
from __graft_entry__ import compare
from lamb_amd import dbcsr as D
from lamb_amd import device as DV
from lamb_amd import systems as S
from lamb_amd.matrix import Distribution

pytestmark = pytest.mark.gpu
GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def test_fixture_loads_exactly(ctx):
    exp = json.loads((GOLDEN / "two_chunks.json").read_text())
    blk = np.array(exp["sizes"], np.int32)
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, 10, (len(blk), 3)).astype(np.float32)
    print("case: dbcsr fixture load", flush=True)
    dist = Distribution(blk, pos, 10.0)
    dd = DV.DeviceDistribution(ctx, dist)
    dm = DV.DeviceMatrix.from_dbcsr(dd, GOLDEN / "two_chunks.dbcsr")
    got = dm.download().to_global()
    want = D.read(GOLDEN / "two_chunks.dbcsr")
    assert np.array_equal(got.row_p, want.row_p) and np.array_equal(got.col, want.col)
    assert np.array_equal(got.data, want.data)
    dm.close()
    dd.close()


def test_loaded_operands_multiply_like_the_oracle(ctx, tmp_path):
    atoms = S.water_system(240, seed=21, heavy_fraction=0.1)
    a = S.decay_matrix(atoms, 7.0, 0.4, 22)
    b = S.decay_matrix(atoms, 7.0, 0.4, 23)
    D.write(tmp_path / "a.dbcsr", a, n_chunks=4)
    D.write(tmp_path / "b.dbcsr", b, n_chunks=5)
    print(f"case: dbcsr load + multiply atoms={atoms.n_atoms}", flush=True)
    dist = Distribution(atoms.blk, atoms.pos, atoms.box)
    dd = DV.DeviceDistribution(ctx, dist)
    da = DV.DeviceMatrix.from_dbcsr(dd, tmp_path / "a.dbcsr")
    db = DV.DeviceMatrix.from_dbcsr(dd, tmp_path / "b.dbcsr")
    dc = DV.DeviceMatrix(dd)
    DV.multiply(da, db, dc, eps_filter=1e-6)
    got = dc.download().to_global()
    ref, _ = oracle.multiply(a, b, eps=1e-6)
    same, err = compare(got, ref)
    print(f"  error {err:.3e}", flush=True)
    assert same and err <= 1e-12
    for h in (dc, db, da, dd):
        h.close()


def test_symmetric_type_and_wrong_sizes_are_refused(ctx, tmp_path):
    atoms = S.water_system(30, seed=31)
    m = S.decay_matrix(atoms, 6.0, 0.4, 32)
    D.write(tmp_path / "s.dbcsr", m, matrix_type="S")
    dist = Distribution(atoms.blk, atoms.pos, atoms.box)
    dd = DV.DeviceDistribution(ctx, dist)
    with pytest.raises(DV.LambError) as e:
        DV.DeviceMatrix.from_dbcsr(dd, tmp_path / "s.dbcsr")
    assert e.value.type == -1000
    other = S.water_system(30, seed=33, heavy_fraction=0.5)
    D.write(tmp_path / "o.dbcsr", S.decay_matrix(other, 6.0, 0.4, 34))
    assert not np.array_equal(other.blk, atoms.blk)
    with pytest.raises(DV.LambError) as e:
        DV.DeviceMatrix.from_dbcsr(dd, tmp_path / "o.dbcsr")
    assert e.value.type == 100
    dd.close()


@pytest.mark.parametrize("eps", [0.0, 1e-6])
def test_multiply_test_entry(ctx, tmp_path, eps):
    """lamb_hip_multiply_test: dump -> A, copies B and C, C = A*B. Its
    statistics equal the oracle's for A*A (kept and filtered triples, flops,
    stored C blocks); the fixture file multiplies too."""
    atoms = S.water_system(150, seed=51, heavy_fraction=0.1)
    a = S.decay_matrix(atoms, 7.0, 0.4, 52)
    D.write(tmp_path / "small.dbcsr", a, n_chunks=3)
    print(f"case: multiply_test atoms={atoms.n_atoms} eps={eps}", flush=True)
    st = DV.multiply_test(ctx, tmp_path / "small.dbcsr", eps_filter=eps)
    ref, rst = oracle.multiply(a, a, eps=eps)
    assert st["n_triples_kept"] == rst["n_triples_kept"]
    assert st["n_triples_filtered"] == rst["n_triples_filtered"]
    assert st["flops"] == rst["flops"]
    assert st["n_c_blocks_kept"] == ref.n_blocks
    st = DV.multiply_test(ctx, GOLDEN / "two_chunks.dbcsr")
    fixture = D.read(GOLDEN / "two_chunks.dbcsr")
    ref, rst = oracle.multiply(fixture, fixture)
    assert st["n_c_blocks_kept"] == ref.n_blocks and st["flops"] == rst["flops"]
    with pytest.raises(DV.LambError) as e:
        DV.multiply_test(ctx, tmp_path / "missing.dbcsr")
    assert e.value.type == -2000
