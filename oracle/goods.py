"""numpy restatement of the reference's sector wire format, for tests only.

TEST INFRASTRUCTURE ONLY: the product path never imports this module.

Follows lamb_goods_X_create (src/lamb_goods_X_types.F90:119-268) and
lamb_goods_X_unpack (src/lamb_goods_X_methods.F90:166-218):
- twelve header slots, 1-based (src/lamb_goods_X_types.F90:79-91): index
  offset, index size, data offset, data size, then offset and element count
  of row_p, col_i, extents and flags;
- each array occupies byte_to_word_count(bytes) = max(1, ceil(bytes / 8))
  int64 words (:262-267), with row_p n_rows int64 (blkn_k = int_8,
  src/lamb_kinds.F90:175), col_i n_blks int16 (locatm_k = int_2, :157),
  extents 2 n_blks int64 (dataptr_k = int_8, :177), flags n_blks bytes;
- offsets run 1, 13, then each array after the previous (:185-213);
- row_p carries the first n_rows entries of the n_rows + 1 CSR pointers
  (TRANSFER of n_words_rows words, :227-231); col_i is TRANSFER-packed
  int16 (:233-237; bytes past the last entry are left by the reference as
  TRANSFER makes them, zero here); extents are (lo, hi) per block
  (:239-252); flags are zeroed (:254-258);
- unpack points row_p, col_i, extents (2 x extents size) and flags into the
  index at the stated offsets (src/lamb_goods_X_methods.F90:189-214).

Pinned by the hand-computed layouts of tests/test_goods.py (a 3-row,
5-block sector and an empty one, every word written out) and by
unpack(pack(s)) == s over seeded sectors.
"""
from __future__ import annotations

import numpy as np

SLOTS = 12


def _words(n_bytes: int) -> int:
    return max(1, (n_bytes + 7) // 8)


def pack(row_p, col_i, extents, data):
    """(index int64[...], data float64[...]) of one sector."""
    row_p = np.asarray(row_p, np.int64)
    col_i = np.asarray(col_i, np.int16)
    extents = np.asarray(extents, np.int64).reshape(-1, 2)
    data = np.asarray(data, np.float64)
    n_rows = row_p.shape[0] - 1
    n_blks = col_i.shape[0]
    w_rows, w_cols = _words(8 * n_rows), _words(2 * n_blks)
    w_ext, w_flags = _words(16 * n_blks), _words(n_blks)
    total = SLOTS + w_rows + w_cols + w_ext + w_flags
    index = np.zeros(total, np.int64)
    o_rows = SLOTS + 1
    o_cols = o_rows + w_rows
    o_ext = o_cols + w_cols
    o_flags = o_ext + w_ext
    index[:SLOTS] = [1, total, 1, data.shape[0], o_rows, n_rows, o_cols, n_blks,
                     o_ext, n_blks, o_flags, n_blks]
    index[o_rows - 1:o_rows - 1 + n_rows] = row_p[:n_rows]
    cols = np.zeros(4 * w_cols, np.dtype("<i2"))
    cols[:n_blks] = col_i
    index[o_cols - 1:o_cols - 1 + w_cols] = cols.view("<i8")
    index[o_ext - 1:o_ext - 1 + 2 * n_blks] = extents.reshape(-1)
    return index, data.copy()


def unpack(index, data):
    """(row_p[n_rows + 1], col_i, extents[n_blks, 2], data) of one sector."""
    index = np.asarray(index, np.int64)
    n_rows = int(index[5])
    n_blks = int(index[9])
    a = int(index[4]) - 1
    row_p = np.empty(n_rows + 1, np.int64)
    row_p[:n_rows] = index[a:a + n_rows]
    row_p[n_rows] = n_blks
    a = int(index[6]) - 1
    col_i = index[a:a + _words(2 * n_blks)].astype("<i8").view("<i2")[:n_blks].astype(np.int16)
    a = int(index[8]) - 1
    extents = index[a:a + 2 * n_blks].reshape(n_blks, 2).copy()
    assert int(index[3]) == len(data)
    return row_p, col_i, extents, np.asarray(data, np.float64).copy()
