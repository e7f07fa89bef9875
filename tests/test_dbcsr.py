"""DBCSR dump reader (C-ABI, no GPU needed) against the byte-level fixture
and the numpy restatement of the reference's reader.

- tests/golden/two_chunks.dbcsr (written field by field by
  tests/golden/make_dbcsr_fixture.py) read by the oracle and by the library:
  the blocks of two_chunks.json, exactly, in file order.
- Round trips of seeded water-box matrices through lamb_amd.dbcsr.write, in
  1..7 chunks (some empty): the library reader returns the same BlockCSR
  bit for bit, and the oracle reader the same blocks.
- Malformed files (truncated, extent beyond the data area, block size 127,
  decreasing row_p): both readers refuse them.
"""
import json
import pathlib
import struct

import numpy as np
import pytest

from lamb_amd import dbcsr as D
from lamb_amd import systems as S
from oracle import dbcsr as O

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
FIXTURE = GOLDEN / "two_chunks.dbcsr"


def expected():
    return json.loads((GOLDEN / "two_chunks.json").read_text())


def test_oracle_reader_matches_fixture():
    head, blocks = O.read(FIXTURE)
    exp = expected()
    assert head["n_chunks"] == exp["n_chunks"] and head["matrix_type"] == "N"
    assert list(head["row_blk_size"]) == exp["sizes"]
    assert len(blocks) == len(exp["blocks"])
    for (r, c, v), (er, ec, ev) in zip(blocks, exp["blocks"]):
        assert (r, c) == (er, ec)
        assert np.array_equal(v, ev)


def test_library_reader_matches_fixture():
    info, rs, cs = D.read_info(FIXTURE)
    exp = expected()
    assert info["n_chunks"] == 2 and info["data_type"] == 3
    assert info["name"].rstrip() == exp["name"] and info["version"] == "DBCSRv_1.0"
    assert list(rs) == exp["sizes"] and list(cs) == exp["sizes"]
    rows, cols, vals, _ = D.read_blocks(FIXTURE)
    assert [[int(r), int(c)] for r, c in zip(rows, cols)] == [[b[0], b[1]] for b in exp["blocks"]]
    assert np.array_equal(vals, np.concatenate([b[2] for b in exp["blocks"]]))
    m = D.read(FIXTURE)
    assert m.n_blocks == len(exp["blocks"])


@pytest.mark.parametrize("n_chunks", [1, 3, 7])
def test_round_trip(tmp_path, n_chunks):
    atoms = S.water_system(60, seed=11, heavy_fraction=0.1)
    m = S.decay_matrix(atoms, 7.0, 0.4, 12)
    path = tmp_path / "m.dbcsr"
    D.write(path, m, n_chunks=n_chunks, name="round trip")
    got = D.read(path)
    assert np.array_equal(got.blk, m.blk)
    assert np.array_equal(got.row_p, m.row_p)
    assert np.array_equal(got.col, m.col)
    assert np.array_equal(got.data, m.data)
    head, blocks = O.read(path)
    assert head["n_chunks"] == n_chunks and len(blocks) == m.n_blocks
    rows = m.rows()
    want = {(int(rows[b]) + 1, int(m.col[b]) + 1): m.data[m.blk_off[b]:m.blk_off[b] +
                                                            m.blk[rows[b]] * m.blk[m.col[b]]]
            for b in range(m.n_blocks)}
    for r, c, v in blocks:
        assert np.array_equal(v, want[(r, c)])


def corrupt(tmp_path, edit):
    raw = bytearray(FIXTURE.read_bytes())
    edit(raw)
    p = tmp_path / "bad.dbcsr"
    p.write_bytes(bytes(raw))
    return p


H1 = 10 + 80 + 1 + 16 + 4 * 5 + 4 * 5


@pytest.mark.parametrize("case", ["truncated", "extent", "size127", "row_p"])
def test_malformed_files_are_refused(tmp_path, case):
    if case == "truncated":
        p = corrupt(tmp_path, lambda r: r.__delitem__(slice(len(r) - 8, len(r))))
    elif case == "extent":
        # chunk 1's data area declared 5 values short of its blocks
        p = corrupt(tmp_path, lambda r: r.__setitem__(slice(H1 + 4, H1 + 8),
                                                      struct.pack("<i", 21)))
    elif case == "size127":
        p = corrupt(tmp_path, lambda r: r.__setitem__(slice(107, 111), struct.pack("<i", 127)))
    else:
        # chunk 1 row_p[1] (after row 1) set above row_p[2]
        off3 = H1 + 8 * 2
        p = corrupt(tmp_path, lambda r: r.__setitem__(slice(off3 + 4, off3 + 8),
                                                      struct.pack("<i", 4)))
    with pytest.raises(D.DbcsrError):
        D.read_blocks(p)
    if case != "row_p":  # the reference's own reader does not check row_p order
        with pytest.raises(O.FormatError):
            O.read(p)
