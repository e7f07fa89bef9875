"""numpy restatement of the reference's DBCSR dump reader, for tests only.

TEST INFRASTRUCTURE ONLY: the product path never imports this module.

Follows lamb_load_dbcsr and load_dbcsr_chunks (src/lamb_io_operations.F90):
- sub-header1 at offset 0: version (10 bytes), name (80), matrix type (1),
  four int32 (numnodes, data type, block rows, block columns), then the row
  and column block sizes (:353-384);
- h1_size and the chunk offsets (:557-597): sub-header2 of chunk c at
  h1_size + 8c; sub-header3s after the last sub-header2, each
  4 (rows + 1) + 4 * 2 nblks bytes; data areas after the last sub-header3,
  each 8 * area bytes;
- the coordinate index (:642-656): global row from row_p ranges, global
  column col_i, extent [blk_p, blk_p - 1 + m n] into the chunk's area;
- the check that every extent lies inside its data area (:669-673) and
  that block sizes stay below huge(int8) = 127 (:391-400).

Pinned by tests/golden/two_chunks.dbcsr, which
tests/golden/make_dbcsr_fixture.py writes field by field from the format
description (:296-330) together with the blocks it holds.
"""
from __future__ import annotations

import numpy as np


class FormatError(ValueError):
    pass


def read(path):
    """(header, blocks): header a dict, blocks a list of (row, col, values)
    in file order (1-based block row / column, values column-major)."""
    raw = open(path, "rb").read()

    def i32(off, n):
        if off + 4 * n > len(raw):
            raise FormatError(f"file ends at {len(raw)}, needs {off + 4 * n}")
        return np.frombuffer(raw, "<i4", n, off).astype(np.int64)

    if len(raw) < 107:
        raise FormatError("file shorter than sub-header1")
    head = {"version": raw[0:10].decode("latin-1"), "name": raw[10:90].decode("latin-1"),
            "matrix_type": raw[90:91].decode("latin-1")}
    numnodes, data_type, nr, nc = (int(v) for v in i32(91, 4))
    head.update(n_chunks=numnodes, data_type=data_type, n_block_rows=nr, n_block_cols=nc)
    row_blk = i32(107, nr)
    col_blk = i32(107 + 4 * nr, nc)
    if np.any(row_blk >= 127) or np.any(col_blk >= 127):
        raise FormatError("block size too big for the int8 block-size kind")
    h1_size = 10 + 80 + 1 + 4 * 4 + 4 * nr + 4 * nc
    h2 = i32(h1_size, 2 * numnodes).reshape(numnodes, 2)
    off3 = np.zeros(numnodes, np.int64)
    offd = np.zeros(numnodes, np.int64)
    off3[0] = h1_size + 8 * numnodes
    for c in range(1, numnodes):
        off3[c] = off3[c - 1] + 4 * (nr + 1) + 4 * 2 * h2[c - 1, 0]
    offd[0] = off3[numnodes - 1] + 4 * (nr + 1) + 4 * 2 * h2[numnodes - 1, 0]
    for c in range(1, numnodes):
        offd[c] = offd[c - 1] + 8 * h2[c - 1, 1]
    head["row_blk_size"], head["col_blk_size"] = row_blk, col_blk
    blocks = []
    for c in range(numnodes):
        nblks, area = int(h2[c, 0]), int(h2[c, 1])
        row_p = i32(int(off3[c]), nr + 1)
        col_i = i32(int(off3[c]) + 4 * (nr + 1), nblks)
        blk_p = i32(int(off3[c]) + 4 * (nr + 1) + 4 * nblks, nblks)
        for r in range(nr):
            for b in range(int(row_p[r]), int(row_p[r + 1])):
                j = int(col_i[b])
                n = int(row_blk[r] * col_blk[j - 1])
                lo = int(blk_p[b])
                if n and lo - 1 + n > area:
                    raise FormatError("data block extent exceeds data area")
                start = int(offd[c]) + 8 * (lo - 1)
                if n and start + 8 * n > len(raw):
                    raise FormatError(f"file ends at {len(raw)}, needs {start + 8 * n}")
                vals = np.frombuffer(raw, "<f8", n, start) if n else np.zeros(0)
                blocks.append((r + 1, j, vals.copy()))
    return head, blocks
