// DBCSR dump reader and loader (host side). The device sees the blocks
// through the staging and commit path of api.hip (lamb_hip_distribute,
// lamb_hip_matrix_commit).
//
// Offsets follow load_dbcsr_chunks (src/lamb_io_operations.F90:557-597):
//   h1   = 10 + 80 + 1 + 4*4 + 4*block_rows + 4*block_cols
//   sub-header2 of chunk c at h1 + 8*c,
//   sub-header3 of chunk 0 after the last sub-header2, each one
//     4*(block_rows + 1) + 4*2*nblks long,
//   data area of chunk 0 after the last sub-header3, each 8*area long.
// Checks kept from the reference: block sizes below 127 (:391-400, the
// int8 blksz_k), every block inside its chunk's data area (:669-673), type
// 'N' only for a load (:409-412).
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "internal.h"

using lamb::set_error;

namespace {

constexpr int kVersionLen = 10;   // dbcsr_version_len, :331
constexpr int kNameLen = 80;      // dbcsr_default_string_length, :332
constexpr int32_t kMaxBlk = 127;  // huge(int8): sizes must stay below it

struct Reader {
  FILE* f = nullptr;
  int64_t size = 0;
  std::string err;
  ~Reader() {
    if (f) fclose(f);
  }
  bool fail(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return false;
  }
  bool open(const char* path) {
    if (!path) return fail("dbcsr: no file name");
    f = fopen(path, "rb");
    if (!f) return fail("dbcsr: cannot open %s", path);
    if (fseeko(f, 0, SEEK_END) != 0) return fail("dbcsr: cannot seek in %s", path);
    size = (int64_t)ftello(f);
    return true;
  }
  bool read(int64_t off, void* dst, int64_t bytes, const char* what) {
    if (bytes == 0) return true;
    if (off < 0 || bytes < 0 || off + bytes > size)
      return fail("dbcsr: %s needs bytes [%lld, %lld) but the file has %lld", what,
                  (long long)off, (long long)(off + bytes), (long long)size);
    if (fseeko(f, (off_t)off, SEEK_SET) != 0 || fread(dst, 1, (size_t)bytes, f) != (size_t)bytes)
      return fail("dbcsr: read of %s failed", what);
    return true;
  }
};

struct Layout {
  char version[kVersionLen + 1] = {};
  char name[kNameLen + 1] = {};
  char type = 0;
  int32_t n_chunks = 0, data_type = 0, nr = 0, nc = 0;
  std::vector<int32_t> rbs, cbs;        // block sizes
  std::vector<int32_t> nblks, area;     // per chunk
  std::vector<int64_t> off3, offd;      // per chunk: sub-header3, data area
  int64_t n_blocks = 0;
};

// sub-header1, every sub-header2 and the chunk offsets
bool parse(Reader& R, Layout& L) {
  int64_t off = 0;
  if (!R.read(off, L.version, kVersionLen, "version")) return false;
  off += kVersionLen;
  if (!R.read(off, L.name, kNameLen, "matrix name")) return false;
  off += kNameLen;
  if (!R.read(off, &L.type, 1, "matrix type")) return false;
  off += 1;
  int32_t h1[4];
  if (!R.read(off, h1, sizeof h1, "sub-header1")) return false;
  off += sizeof h1;
  L.n_chunks = h1[0];
  L.data_type = h1[1];
  L.nr = h1[2];
  L.nc = h1[3];
  if (L.n_chunks < 0 || L.nr < 0 || L.nc < 0)
    return R.fail("dbcsr: negative count in sub-header1 (chunks %d, rows %d, columns %d)",
                  L.n_chunks, L.nr, L.nc);
  L.rbs.resize(L.nr);
  L.cbs.resize(L.nc);
  if (!R.read(off, L.rbs.data(), 4 * (int64_t)L.nr, "row block sizes")) return false;
  off += 4 * (int64_t)L.nr;
  if (!R.read(off, L.cbs.data(), 4 * (int64_t)L.nc, "column block sizes")) return false;
  off += 4 * (int64_t)L.nc;
  for (int32_t i = 0; i < L.nr; ++i)
    if (L.rbs[i] < 0 || L.rbs[i] >= kMaxBlk)
      return R.fail("dbcsr: row block size %d of block row %d is outside 0..%d", L.rbs[i], i + 1,
                    kMaxBlk - 1);
  for (int32_t j = 0; j < L.nc; ++j)
    if (L.cbs[j] < 0 || L.cbs[j] >= kMaxBlk)
      return R.fail("dbcsr: column block size %d of block column %d is outside 0..%d",
                    L.cbs[j], j + 1, kMaxBlk - 1);
  const int64_t h1_size = off;
  L.nblks.resize(L.n_chunks);
  L.area.resize(L.n_chunks);
  for (int32_t c = 0; c < L.n_chunks; ++c) {
    int32_t h2[2];
    if (!R.read(h1_size + 8 * (int64_t)c, h2, sizeof h2, "sub-header2")) return false;
    if (h2[0] < 0 || h2[1] < 0)
      return R.fail("dbcsr: chunk %d has %d blocks and a data area of %d", c + 1, h2[0], h2[1]);
    L.nblks[c] = h2[0];
    L.area[c] = h2[1];
    L.n_blocks += h2[0];
  }
  L.off3.resize(L.n_chunks);
  L.offd.resize(L.n_chunks);
  int64_t o = h1_size + 8 * (int64_t)L.n_chunks;
  for (int32_t c = 0; c < L.n_chunks; ++c) {
    L.off3[c] = o;
    o += 4 * ((int64_t)L.nr + 1) + 8 * (int64_t)L.nblks[c];
  }
  for (int32_t c = 0; c < L.n_chunks; ++c) {
    L.offd[c] = o;
    o += 8 * (int64_t)L.area[c];
  }
  return true;
}

// One chunk's index: per block (file order) its 1-based row and column, the
// 0-based offset of its values in the chunk's data area and their count;
// used = values of the area the blocks reach.
struct ChunkIndex {
  std::vector<int32_t> row, col;
  std::vector<int64_t> lo, len;
  int64_t used = 0;
};

bool chunk_index(Reader& R, const Layout& L, int32_t c, ChunkIndex& X) {
  const int32_t nb = L.nblks[c];
  std::vector<int32_t> row_p(L.nr + 1), col_i(nb), blk_p(nb);
  int64_t off = L.off3[c];
  if (!R.read(off, row_p.data(), 4 * ((int64_t)L.nr + 1), "row_p")) return false;
  off += 4 * ((int64_t)L.nr + 1);
  if (!R.read(off, col_i.data(), 4 * (int64_t)nb, "col_i")) return false;
  off += 4 * (int64_t)nb;
  if (!R.read(off, blk_p.data(), 4 * (int64_t)nb, "blk_p")) return false;
  if (row_p[0] != 0 || row_p[L.nr] != nb)
    return R.fail("dbcsr: chunk %d row_p runs from %d to %d for %d blocks", c + 1, row_p[0],
                  row_p[L.nr], nb);
  X.row.resize(nb);
  X.col.resize(nb);
  X.lo.resize(nb);
  X.len.resize(nb);
  X.used = 0;
  for (int32_t r = 0; r < L.nr; ++r) {
    if (row_p[r + 1] < row_p[r])
      return R.fail("dbcsr: chunk %d row_p decreases at block row %d", c + 1, r + 1);
    for (int32_t b = row_p[r]; b < row_p[r + 1]; ++b) {
      const int32_t j = col_i[b];
      if (j < 1 || j > L.nc)
        return R.fail("dbcsr: chunk %d block %d has column %d outside 1..%d", c + 1, b + 1, j,
                      L.nc);
      const int64_t len = (int64_t)L.rbs[r] * L.cbs[j - 1];
      int64_t lo = 0;
      if (len > 0) {
        lo = (int64_t)blk_p[b] - 1;
        if (lo < 0 || lo + len > L.area[c])
          return R.fail("dbcsr: chunk %d block (%d,%d) extent [%lld, %lld] exceeds its data "
                        "area of %d",
                        c + 1, r + 1, j, (long long)(lo + 1), (long long)(lo + len), L.area[c]);
      }
      X.row[b] = r + 1;
      X.col[b] = j;
      X.lo[b] = lo;
      X.len[b] = len;
      X.used = std::max(X.used, lo + len);
    }
  }
  return true;
}

int32_t to_msg(const Reader& R, char* msg, int32_t msg_len) {
  if (msg && msg_len > 0) {
    strncpy(msg, R.err.c_str(), (size_t)msg_len - 1);
    msg[msg_len - 1] = 0;
  }
  return LAMB_HIP_FAILURE;
}

}  // namespace

extern "C" {

int32_t lamb_hip_dbcsr_read_info(const char* path, lamb_hip_dbcsr_info* info,
                                 int32_t* row_blk_size, int32_t* col_blk_size, char* msg,
                                 int32_t msg_len) {
  try {
    Reader R;
    Layout L;
    if (!info) {
      R.fail("dbcsr: no info structure");
      return to_msg(R, msg, msg_len);
    }
    if (!R.open(path) || !parse(R, L)) return to_msg(R, msg, msg_len);
    int64_t n_values = 0;
    for (int32_t c = 0; c < L.n_chunks; ++c) {
      ChunkIndex X;
      if (!chunk_index(R, L, c, X)) return to_msg(R, msg, msg_len);
      for (int64_t v : X.len) n_values += v;
    }
    memset(info, 0, sizeof *info);
    memcpy(info->version, L.version, kVersionLen);
    memcpy(info->name, L.name, kNameLen);
    info->matrix_type = L.type;
    info->n_chunks = L.n_chunks;
    info->data_type = L.data_type;
    info->n_block_rows = L.nr;
    info->n_block_cols = L.nc;
    info->n_blocks = L.n_blocks;
    info->n_values = n_values;
    if (row_blk_size) std::copy(L.rbs.begin(), L.rbs.end(), row_blk_size);
    if (col_blk_size) std::copy(L.cbs.begin(), L.cbs.end(), col_blk_size);
    return LAMB_HIP_OK;
  } catch (const std::bad_alloc&) {
    if (msg && msg_len > 0) snprintf(msg, (size_t)msg_len, "dbcsr: out of host memory");
    return LAMB_HIP_MEMORY;
  }
}

int32_t lamb_hip_dbcsr_read_blocks(const char* path, int64_t n_blocks, int32_t* rows,
                                   int32_t* cols, int64_t n_values, double* values, char* msg,
                                   int32_t msg_len) {
  try {
    Reader R;
    Layout L;
    if (!R.open(path) || !parse(R, L)) return to_msg(R, msg, msg_len);
    if (n_blocks != L.n_blocks || (n_blocks > 0 && (!rows || !cols))) {
      R.fail("dbcsr: the file holds %lld blocks, the arrays %lld", (long long)L.n_blocks,
             (long long)n_blocks);
      return to_msg(R, msg, msg_len);
    }
    int64_t b_out = 0, v_out = 0;
    std::vector<double> area;
    for (int32_t c = 0; c < L.n_chunks; ++c) {
      ChunkIndex X;
      if (!chunk_index(R, L, c, X)) return to_msg(R, msg, msg_len);
      area.resize(X.used);
      if (!R.read(L.offd[c], area.data(), 8 * X.used, "data area")) return to_msg(R, msg, msg_len);
      for (size_t b = 0; b < X.row.size(); ++b) {
        if (v_out + X.len[b] > n_values || (X.len[b] > 0 && !values)) {
          R.fail("dbcsr: the file holds more than the %lld values given", (long long)n_values);
          return to_msg(R, msg, msg_len);
        }
        rows[b_out] = X.row[b];
        cols[b_out] = X.col[b];
        ++b_out;
        std::copy(area.begin() + X.lo[b], area.begin() + X.lo[b] + X.len[b], values + v_out);
        v_out += X.len[b];
      }
    }
    if (v_out != n_values) {
      R.fail("dbcsr: the file holds %lld values, the array %lld", (long long)v_out,
             (long long)n_values);
      return to_msg(R, msg, msg_len);
    }
    return LAMB_HIP_OK;
  } catch (const std::bad_alloc&) {
    if (msg && msg_len > 0) snprintf(msg, (size_t)msg_len, "dbcsr: out of host memory");
    return LAMB_HIP_MEMORY;
  }
}

int32_t lamb_hip_dbcsr_load(lamb_hip_mat* m, const char* path) {
  if (!m) return LAMB_HIP_FAILURE;
  lamb_hip_dist* d = m->dist;
  lamb_hip_ctx* ctx = d->ctx;
  try {
    Reader R;
    Layout L;
    if (!R.open(path) || !parse(R, L))
      return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_EXTERNAL_ERROR, "%s", R.err.c_str());
    if (L.type != 'N')
      return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_UNIMPLEMENTED,
                       "dbcsr_load: matrix type '%c': only non-symmetric ('N') matrices are "
                       "supported",
                       L.type);
    if (L.nr != d->n || L.nc != d->n)
      return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                       "dbcsr_load: the file has %d x %d block rows / columns, the distribution "
                       "%d atoms",
                       L.nr, L.nc, d->n);
    for (int32_t i = 0; i < d->n; ++i)
      if (L.rbs[i] != d->blk[i] || L.cbs[i] != d->blk[i])
        return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                         "dbcsr_load: block %d is %d x %d in the file but %d in the "
                         "distribution",
                         i + 1, L.rbs[i], L.cbs[i], d->blk[i]);
    std::vector<double> area;
    std::vector<int32_t> subset_idx;
    std::vector<int16_t> local_idx;
    std::vector<int64_t> ext;
    for (int32_t c = 0; c < L.n_chunks; ++c) {
      ChunkIndex X;
      if (!chunk_index(R, L, c, X))
        return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_EXTERNAL_ERROR, "%s", R.err.c_str());
      area.resize(X.used);
      if (!R.read(L.offd[c], area.data(), 8 * X.used, "data area"))
        return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_EXTERNAL_ERROR, "%s", R.err.c_str());
      const size_t nb = X.row.size();
      subset_idx.resize(2 * nb);
      local_idx.resize(2 * nb);
      ext.resize(2 * nb);
      // the global -> (subset, local) step of lamb_dist_localize_index (:708-711)
      for (size_t b = 0; b < nb; ++b) {
        const int32_t gr = X.row[b] - 1, gc = X.col[b] - 1;
        subset_idx[2 * b] = d->subset_of[gr];
        subset_idx[2 * b + 1] = d->subset_of[gc];
        local_idx[2 * b] = (int16_t)d->local_of[gr];
        local_idx[2 * b + 1] = (int16_t)d->local_of[gc];
        ext[2 * b] = X.lo[b] + 1;
        ext[2 * b + 1] = X.lo[b] + X.len[b];
      }
      const int32_t rc = lamb_hip_distribute(m, (int64_t)nb, subset_idx.data(), local_idx.data(),
                                             ext.data(), area.data(), X.used);
      if (rc != LAMB_HIP_OK) return rc;
    }
    return lamb_hip_matrix_commit(m);
  } catch (const std::bad_alloc&) {
    return set_error(ctx, LAMB_HIP_MEMORY, LAMB_HIP_EXTERNAL_ERROR, "dbcsr_load: out of host memory");
  }
}

// lamb_multiply_test (src/lamb_multiplication.F90:85-114): load a DBCSR
// dump as A, copy it to B and C, and multiply C = A*B. The distribution is
// the file's atoms in file order, split into subsets of at most 32,767 atoms
// (the int16 local index); the reference takes it from a coordinate file.
int32_t lamb_hip_multiply_test(lamb_hip_ctx* ctx, const char* path, double eps_filter,
                               lamb_hip_stats* stats) {
  if (!ctx) return LAMB_HIP_FAILURE;
  if (!path)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS, "multiply_test: no path");
  std::vector<int8_t> blk;
  try {
    Reader R;
    Layout L;
    if (!R.open(path) || !parse(R, L))
      return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_EXTERNAL_ERROR, "%s", R.err.c_str());
    if (L.nr != L.nc || L.rbs != L.cbs)
      return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                       "multiply_test: %s is not square in atoms (%d x %d block rows / "
                       "columns or different sizes)",
                       path, L.nr, L.nc);
    blk.assign(L.rbs.begin(), L.rbs.end());
  } catch (const std::bad_alloc&) {
    return set_error(ctx, LAMB_HIP_MEMORY, LAMB_HIP_EXTERNAL_ERROR,
                     "multiply_test: out of host memory");
  }
  const int64_t n = (int64_t)blk.size();
  constexpr int64_t kSubsetMax = 32767;
  const int32_t n_subsets = (int32_t)std::max<int64_t>(1, (n + kSubsetMax - 1) / kSubsetMax);
  std::vector<int32_t> smap(n);
  std::vector<int16_t> omap(n);
  for (int64_t g = 0; g < n; ++g) {
    smap[g] = (int32_t)(g / kSubsetMax) + 1;
    omap[g] = (int16_t)(g % kSubsetMax + 1);
  }
  lamb_hip_dist* dist = nullptr;
  lamb_hip_mat* mats[3] = {nullptr, nullptr, nullptr};
  int32_t rc = lamb_hip_dist_create(ctx, n, blk.data(), n_subsets, smap.data(), omap.data(),
                                    nullptr, nullptr, &dist);
  for (int i = 0; i < 3 && rc == LAMB_HIP_OK; ++i) rc = lamb_hip_matrix_create(dist, &mats[i]);
  if (rc == LAMB_HIP_OK) rc = lamb_hip_dbcsr_load(mats[0], path);
  if (rc == LAMB_HIP_OK) rc = lamb_hip_matrix_copy(mats[0], mats[1]);
  if (rc == LAMB_HIP_OK) rc = lamb_hip_matrix_copy(mats[0], mats[2]);
  if (rc == LAMB_HIP_OK)
    rc = lamb_hip_multiply(ctx, mats[0], mats[1], mats[2], 1.0, 0.0, eps_filter, LAMB_HIP_FP64,
                           stats);
  for (int i = 2; i >= 0; --i)
    if (mats[i]) lamb_hip_matrix_destroy(mats[i]);
  if (dist) lamb_hip_dist_destroy(dist);
  return rc;
}

}  // extern "C"
