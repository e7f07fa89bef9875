// Sector wire format: the serialized form in which one sector travels
// between processes (lamb_goods_X_create, src/lamb_goods_X_types.F90:119-268,
// and lamb_goods_X_unpack, src/lamb_goods_X_methods.F90:166-218), built from
// and into this library's device pool through the sector download / upload
// entry points.
//
// The index array is int64 words, slots 1-based as in the reference:
//   words 1..12: index offset (1), index size, data offset (1), data size,
//                row_p offset, n_rows, col_i offset, n_blks,
//                extents offset, n_blks, flags offset, n_blks
//   row_p:   the first n_rows entries of the sector's row_p (0-based
//            exclusive block offsets; the last, n_blks, is implied)
//   col_i:   n_blks int16 packed into words, little-endian, padding zero
//   extents: n_blks (lo, hi) int64 pairs, 1-based inclusive into data
//   flags:   one byte per block, all zero (the reference does not copy them)
// Each of the four arrays takes max(1, ceil(bytes / 8)) words. The data
// array is the sector's FP64 store (data size doubles).
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.h"
#include "internal.h"

using namespace lamb;

namespace {

constexpr int64_t kGoodsSlots = 12;

struct GoodsLayout {
  int64_t rows = 0, cols = 0, extents = 0, flags = 0, total = 0;
};

int64_t words_of(int64_t bytes) { return std::max<int64_t>(1, (bytes + 7) / 8); }

GoodsLayout goods_layout(int64_t n_rows, int64_t n_blks) {
  GoodsLayout g;
  g.rows = words_of(8 * n_rows);
  g.cols = words_of(2 * n_blks);
  g.extents = words_of(16 * n_blks);
  g.flags = words_of(n_blks);
  g.total = kGoodsSlots + g.rows + g.cols + g.extents + g.flags;
  return g;
}

}  // namespace

extern "C" {

int32_t lamb_hip_goods_size(lamb_hip_mat* m, int32_t sr, int32_t sc, int64_t* n_index_words,
                            int64_t* n_data) {
  if (!m) return LAMB_HIP_FAILURE;
  lamb_hip_dist* d = m->dist;
  int64_t n_blks = 0, len = 0;
  int32_t rc = lamb_hip_sector_query(m, sr, sc, &n_blks, &len);
  if (rc) return rc;
  const int64_t n_rows = (int64_t)d->subset_atoms[sr - 1].size();
  if (n_index_words) *n_index_words = goods_layout(n_rows, n_blks).total;
  if (n_data) *n_data = len;
  return LAMB_HIP_OK;
}

int32_t lamb_hip_goods_pack(lamb_hip_mat* m, int32_t sr, int32_t sc, int64_t* index,
                            int64_t n_index_words, double* data, int64_t n_data) {
  if (!m) return LAMB_HIP_FAILURE;
  lamb_hip_dist* d = m->dist;
  lamb_hip_ctx* ctx = d->ctx;
  int64_t n_blks = 0, len = 0;
  int32_t rc = lamb_hip_sector_query(m, sr, sc, &n_blks, &len);
  if (rc) return rc;
  const int64_t n_rows = (int64_t)d->subset_atoms[sr - 1].size();
  const GoodsLayout g = goods_layout(n_rows, n_blks);
  if (!index || n_index_words != g.total || n_data != len || (len > 0 && !data))
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "goods_pack: buffers of %lld index words and %lld values, the sector "
                     "needs %lld and %lld",
                     (long long)n_index_words, (long long)n_data, (long long)g.total,
                     (long long)len);
  std::vector<int64_t> row_p(n_rows + 1);
  std::vector<int16_t> col_i(std::max<int64_t>(n_blks, 1));
  std::vector<int64_t> ext(std::max<int64_t>(2 * n_blks, 1));
  rc = lamb_hip_sector_download(m, sr, sc, row_p.data(), col_i.data(), ext.data(), data);
  if (rc) return rc;
  std::fill(index, index + g.total, (int64_t)0);
  const int64_t o_rows = kGoodsSlots + 1;  // 1-based word offsets
  const int64_t o_cols = o_rows + g.rows;
  const int64_t o_ext = o_cols + g.cols;
  const int64_t o_flags = o_ext + g.extents;
  const int64_t header[kGoodsSlots] = {1,      g.total, 1,     len,    o_rows,  n_rows,
                                       o_cols, n_blks,  o_ext, n_blks, o_flags, n_blks};
  std::copy(header, header + kGoodsSlots, index);
  std::copy(row_p.begin(), row_p.begin() + n_rows, index + (o_rows - 1));
  if (n_blks) {
    std::memcpy(index + (o_cols - 1), col_i.data(), sizeof(int16_t) * (size_t)n_blks);
    std::copy(ext.begin(), ext.begin() + 2 * n_blks, index + (o_ext - 1));
  }
  return LAMB_HIP_OK;
}

int32_t lamb_hip_goods_unpack(lamb_hip_mat* m, int32_t sr, int32_t sc, const int64_t* index,
                              int64_t n_index_words, const double* data, int64_t n_data) {
  if (!m) return LAMB_HIP_FAILURE;
  lamb_hip_dist* d = m->dist;
  lamb_hip_ctx* ctx = d->ctx;
  if (sr < 1 || sr > d->n_subsets || sc < 1 || sc > d->n_subsets)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "goods_unpack: subsets (%d,%d) outside 1..%d", sr, sc, d->n_subsets);
  if (!index || n_index_words < kGoodsSlots)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "goods_unpack: index of %lld words has no header",
                     (long long)n_index_words);
  const int64_t n_rows = (int64_t)d->subset_atoms[sr - 1].size();
  const int64_t n_blks = index[7];
  // every block takes two extents words, so n_blks <= n_index_words bounds the
  // header's count before any size is computed from it
  if (index[5] != n_rows || n_blks < 0 || n_blks > n_index_words || index[9] != n_blks ||
      index[11] != n_blks)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "goods_unpack: header holds %lld rows and %lld / %lld / %lld blocks, "
                     "subset %d has %lld rows",
                     (long long)index[5], (long long)index[7], (long long)index[9],
                     (long long)index[11], sr, (long long)n_rows);
  const GoodsLayout g = goods_layout(n_rows, n_blks);
  const int64_t o_rows = kGoodsSlots + 1;
  const int64_t o_cols = o_rows + g.rows;
  const int64_t o_ext = o_cols + g.cols;
  const int64_t o_flags = o_ext + g.extents;
  if (index[0] != 1 || index[1] != g.total || n_index_words != g.total || index[2] != 1 ||
      index[3] != n_data || index[4] != o_rows || index[6] != o_cols || index[8] != o_ext ||
      index[10] != o_flags || (n_data > 0 && !data))
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "goods_unpack: header offsets or sizes do not match a sector of %lld rows "
                     "and %lld blocks with %lld values",
                     (long long)n_rows, (long long)n_blks, (long long)n_data);
  std::vector<int64_t> row_p(n_rows + 1);
  std::copy(index + (o_rows - 1), index + (o_rows - 1) + n_rows, row_p.begin());
  row_p[n_rows] = n_blks;
  std::vector<int16_t> col_i(std::max<int64_t>(n_blks, 1));
  if (n_blks) std::memcpy(col_i.data(), index + (o_cols - 1), sizeof(int16_t) * (size_t)n_blks);
  return lamb_hip_sector_upload(m, sr, sc, n_blks, row_p.data(), col_i.data(),
                                index + (o_ext - 1), data, n_data);
}

}  // extern "C"
