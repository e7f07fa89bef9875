"""Write tests/golden/two_chunks.dbcsr and two_chunks.json.

The file is assembled field by field with struct, following the DBCSR dump
format as the reference documents and reads it
(src/lamb_io_operations.F90:296-330 and the offsets at :557-597):
sub-header1, one sub-header2 per chunk, one sub-header3 per chunk, one FP64
data area per chunk, little-endian int32 / float64.

Contents: 5 block rows and columns of sizes 2, 3, 1, 2, 3; type 'N'; two
chunks (the blocks of rows 1, 3, 5 and of rows 2, 4). Chunk 1 stores its
blocks out of order in its data area, with an unused value between two
blocks and two unused values at the end, so blk_p is exercised. Values are
integers (1000 row + 100 column + element), exact in FP64. The JSON holds
the blocks in file order: [row, column, values (column-major)].

Run: python tests/golden/make_dbcsr_fixture.py
"""
import json
import pathlib
import struct

HERE = pathlib.Path(__file__).resolve().parent
SIZES = [2, 3, 1, 2, 3]
# per chunk: blocks in (row, column) order, with the 1-based blk_p chosen below
CHUNKS = [
    [(1, 1), (1, 4), (3, 2), (5, 1), (5, 5)],
    [(2, 2), (2, 5), (4, 3)],
]


def values(r, c):
    n = SIZES[r - 1] * SIZES[c - 1]
    return [float(1000 * r + 100 * c + e) for e in range(n)]


def main():
    n = len(SIZES)
    out = bytearray()
    out += b"DBCSRv_1.0"                                   # version, 10 chars
    out += b"two_chunks_fixture".ljust(80)                 # name, 80 chars
    out += b"N"                                            # matrix type
    out += struct.pack("<4i", len(CHUNKS), 3, n, n)        # numnodes, data type, rows, cols
    out += struct.pack(f"<{n}i", *SIZES)                   # row block sizes
    out += struct.pack(f"<{n}i", *SIZES)                   # column block sizes
    headers3, areas, expected = [], [], []
    for ci, blocks in enumerate(CHUNKS):
        # data area layout: chunk 1 reversed with one pad value after its
        # first stored block and two at the end; chunk 2 in order, packed
        order = list(reversed(blocks)) if ci == 0 else list(blocks)
        area, blk_p = [], {}
        for k, (r, c) in enumerate(order):
            blk_p[(r, c)] = len(area) + 1
            area += values(r, c)
            if ci == 0 and k == 0:
                area.append(-7.0)
        if ci == 0:
            area += [-8.0, -9.0]
        row_p = [0] * (n + 1)
        for r, _ in blocks:
            row_p[r] += 1
        for r in range(n):
            row_p[r + 1] += row_p[r]
        h3 = struct.pack(f"<{n + 1}i", *row_p)
        h3 += struct.pack(f"<{len(blocks)}i", *[c for _, c in blocks])
        h3 += struct.pack(f"<{len(blocks)}i", *[blk_p[b] for b in blocks])
        headers3.append(h3)
        areas.append(struct.pack(f"<{len(area)}d", *area))
        out += struct.pack("<2i", len(blocks), len(area))  # sub-header2: nblks, area size
        expected += [[r, c, values(r, c)] for r, c in blocks]
    for h3 in headers3:
        out += h3
    for a in areas:
        out += a
    (HERE / "two_chunks.dbcsr").write_bytes(bytes(out))
    (HERE / "two_chunks.json").write_text(json.dumps(
        {"sizes": SIZES, "name": "two_chunks_fixture", "matrix_type": "N", "n_chunks": 2,
         "data_type": 3, "blocks": expected}, indent=1) + "\n")


if __name__ == "__main__":
    main()
