"""Sector wire format (lamb_goods_X_create / lamb_goods_X_unpack) on the CPU:
the numpy restatement against layouts computed by hand from
src/lamb_goods_X_types.F90:79-91 and :119-268, and its round trip.

- A 3-row, 5-block sector: header, row_p, col_i (int16 pairs packed
  little-endian, one over-large column to show the 16-bit field), extents and
  the zero flags word, every word written out.
- An empty sector of 2 rows: each of col_i, extents and flags still takes one
  word (byte_to_word_count's max(1, ...), :262-267).
- unpack(pack(s)) == s over seeded sectors of 0..70 blocks, so the col_i
  word count crosses several multiples of four.
The library's own pack and unpack are compared with these on the GPU
(tests/test_gpu_goods.py).
"""
import numpy as np
import pytest

from oracle import goods as G


def test_pack_three_rows_five_blocks():
    row_p = [0, 2, 2, 5]
    col_i = [1, 3, 2, 4, 300]
    ext = [[1, 4], [5, 6], [7, 12], [13, 16], [17, 20]]
    data = np.arange(20) + 0.5
    index, d = G.pack(row_p, col_i, ext, data)
    expected = [
        # index off, size, data off, size, row_p off, n, col_i off, n, ext off, n, flags off, n
        1, 28, 1, 20, 13, 3, 16, 5, 18, 5, 28, 5,
        0, 2, 2,                         # row_p without its last entry
        1125908496973825, 300,           # col_i 1,3,2,4 | 300,0,0,0 as int16 lanes
        1, 4, 5, 6, 7, 12, 13, 16, 17, 20,
        0,                               # five flag bytes in one word
    ]
    assert index.tolist() == expected
    assert np.array_equal(d, data)


def test_pack_empty_sector():
    index, d = G.pack([0, 0, 0], np.zeros(0, np.int16), np.zeros((0, 2), np.int64), [])
    assert index.tolist() == [1, 17, 1, 0, 13, 2, 15, 0, 16, 0, 17, 0, 0, 0, 0, 0, 0]
    assert d.shape == (0,)
    row_p, col_i, ext, dd = G.unpack(index, d)
    assert row_p.tolist() == [0, 0, 0] and col_i.shape == (0,) and ext.shape == (0, 2)


@pytest.mark.parametrize("seed", range(6))
def test_round_trip(seed):
    rng = np.random.default_rng(seed)
    n_rows = int(rng.integers(1, 9))
    counts = rng.integers(0, 10, n_rows)
    row_p = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    n = int(row_p[-1])
    col_i = rng.integers(1, 2000, n).astype(np.int16)
    sizes = rng.integers(1, 30, n)
    hi = np.cumsum(sizes)
    ext = np.stack([hi - sizes + 1, hi], 1).astype(np.int64)
    data = rng.standard_normal(int(hi[-1]) if n else 0)
    index, d = G.pack(row_p, col_i, ext, data)
    assert index[1] == len(index) and index[3] == len(data)
    r2, c2, e2, d2 = G.unpack(index, d)
    assert np.array_equal(r2, row_p) and np.array_equal(c2, col_i)
    assert np.array_equal(e2, ext) and np.array_equal(d2, data)
