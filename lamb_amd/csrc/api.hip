// C-ABI layer: contexts, distributions, the device block-CSR pool, sector
// upload/download in the reference's layout, and the multiply driver.
#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "common.h"
#include "internal.h"

using namespace lamb;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
int32_t lamb::set_error(lamb_hip_ctx* ctx, int32_t level, int32_t type, const char* fmt, ...) {
  if (ctx) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    ctx->level = level;
    ctx->type = type;
    ctx->msg = buf;
  }
  return level;
}

int32_t lamb::hip_error(lamb_hip_ctx* ctx, hipError_t e, const char* what) {
  const int32_t level = (e == hipErrorOutOfMemory) ? LAMB_HIP_MEMORY : LAMB_HIP_FATAL;
  return set_error(ctx, level, LAMB_HIP_EXTERNAL_ERROR, "%s: %s", what, hipGetErrorString(e));
}

void lamb::DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

hipError_t lamb::DevBuf::ensure(size_t want) {
  if (want <= bytes && p) return hipSuccess;
  release();
  if (want == 0) want = 256;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    p = nullptr;
    bytes = 0;
    return e;
  }
  bytes = want;
  return hipSuccess;
}

hipError_t lamb::DevBuf::upload(hipStream_t s, const void* src, size_t n) {
  hipError_t e = ensure(n);
  if (e != hipSuccess) return e;
  if (n == 0) return hipSuccess;
  return hipMemcpyAsync(p, src, n, hipMemcpyHostToDevice, s);
}

#define HIPCHK(ctx, expr, what)                       \
  do {                                                \
    hipError_t e_ = (expr);                           \
    if (e_ != hipSuccess) return hip_error(ctx, e_, what); \
  } while (0)

extern "C" {

int32_t lamb_hip_init(int32_t device, lamb_hip_ctx** out) {
  if (!out) return LAMB_HIP_FAILURE;
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return LAMB_HIP_FATAL;
  if (device < 0 || device >= n) return LAMB_HIP_FAILURE;
  if (hipSetDevice(device) != hipSuccess) return LAMB_HIP_FATAL;
  auto* ctx = new lamb_hip_ctx();
  ctx->device = device;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return LAMB_HIP_FATAL;
  }
  for (auto& ev : ctx->ev)
    if (hipEventCreate(&ev) != hipSuccess) {
      delete ctx;
      return LAMB_HIP_FATAL;
    }
  if (install_guard(device, &ctx->guard) != hipSuccess || noalloc_setup(device) != hipSuccess) {
    delete ctx;
    return LAMB_HIP_FATAL;
  }
  *out = ctx;
  return LAMB_HIP_OK;
}

int32_t lamb_hip_finalize(lamb_hip_ctx* ctx) {
  if (!ctx) return LAMB_HIP_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  release_sym_state(ctx);
  ctx->d_kept.release();
  ctx->d_apack.release();
  ctx->d_bpack.release();
  ctx->d_aexp.release();
  ctx->d_bexp.release();
  ctx->d_aslice.release();
  ctx->d_bslice.release();
  ctx->d_slot_counter.release();
  for (auto& ev : ctx->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return LAMB_HIP_OK;
}

int32_t lamb_hip_last_error(const lamb_hip_ctx* ctx, int32_t* level, int32_t* type, char* msg,
                            int32_t msg_len) {
  if (!ctx) return LAMB_HIP_FAILURE;
  if (level) *level = ctx->level;
  if (type) *type = ctx->type;
  if (msg && msg_len > 0) {
    const size_t n = std::min<size_t>((size_t)msg_len - 1, ctx->msg.size());
    memcpy(msg, ctx->msg.data(), n);
    msg[n] = 0;
  }
  return LAMB_HIP_OK;
}

// ---------------------------------------------------------------------------
// distribution
// ---------------------------------------------------------------------------
int32_t lamb_hip_dist_create(lamb_hip_ctx* ctx, int64_t n_atoms, const int8_t* blk_size,
                             int32_t n_subsets, const int32_t* subset_map,
                             const int16_t* order_map, const float* coords, const float* box,
                             lamb_hip_dist** out) {
  if (!ctx || !out) return LAMB_HIP_FAILURE;
  *out = nullptr;
  if (n_atoms < 0 || n_atoms > INT32_MAX - 1 || n_subsets < 1)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "dist_create: n_atoms=%lld n_subsets=%d out of range", (long long)n_atoms,
                     n_subsets);
  if (n_atoms > 0 && (!blk_size || !subset_map || !order_map))
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS, "dist_create: NULL array");
  auto d = std::make_unique<lamb_hip_dist>();
  d->ctx = ctx;
  d->n = (int32_t)n_atoms;
  d->n_subsets = n_subsets;
  d->blk.resize(d->n);
  d->subset_of.resize(d->n);
  d->local_of.resize(d->n);
  d->subset_atoms.assign(n_subsets, {});
  std::vector<int32_t> count(n_subsets, 0);
  for (int32_t g = 0; g < d->n; ++g) {
    if (blk_size[g] < 0)
      return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                       "dist_create: negative block size for atom %d", g + 1);
    d->blk[g] = blk_size[g];
    d->max_blk = std::max(d->max_blk, d->blk[g]);
    if (d->blk[g] <= 127) d->size_mask[d->blk[g] >> 6] |= 1ull << (d->blk[g] & 63);
    const int32_t s = subset_map[g];
    if (s < 1 || s > n_subsets)
      return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                       "dist_create: atom %d has subset %d outside 1..%d", g + 1, s, n_subsets);
    d->subset_of[g] = s;
    d->local_of[g] = order_map[g];
    count[s - 1] += 1;
  }
  for (int32_t s = 0; s < n_subsets; ++s) d->subset_atoms[s].assign(count[s], -1);
  for (int32_t g = 0; g < d->n; ++g) {
    const int32_t s = d->subset_of[g] - 1;
    const int32_t l = d->local_of[g];
    if (l < 1 || l > count[s] || d->subset_atoms[s][l - 1] != -1)
      return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                       "dist_create: atom %d has local index %d invalid or repeated in subset %d",
                       g + 1, l, s + 1);
    d->subset_atoms[s][l - 1] = g;
  }
  // internal (tile) order
  d->iperm.resize(d->n);
  std::iota(d->iperm.begin(), d->iperm.end(), 0);
  if (coords && d->n > 1) spatial_order(d->n, coords, d->iperm);
  d->perm.resize(d->n);
  for (int32_t p = 0; p < d->n; ++p) d->perm[d->iperm[p]] = p;
  d->blk_int.resize(d->n);
  for (int32_t p = 0; p < d->n; ++p) d->blk_int[p] = d->blk[d->iperm[p]];
  (void)box;
  // bands
  d->band_of.resize(d->n);
  d->pos_in_band.resize(d->n);
  Band cur{};
  cur.first = 0;
  auto close_band = [&](int32_t next_first) {
    if (cur.natoms > 0) d->bands.push_back(cur);
    cur = Band{};
    cur.first = next_first;
  };
  for (int32_t p = 0; p < d->n; ++p) {
    const int b = d->blk_int[p];
    const int rows = cur.natoms ? cur.rs[cur.natoms] : 0;
    if (cur.natoms == kBandAtoms || (cur.natoms > 0 && rows + b > kBandRows)) close_band(p);
    const int r0 = cur.natoms ? cur.rs[cur.natoms] : 0;
    cur.rs[cur.natoms] = (int16_t)r0;
    d->band_of[p] = (int32_t)d->bands.size();
    d->pos_in_band[p] = cur.natoms;
    cur.natoms += 1;
    cur.rs[cur.natoms] = (int16_t)std::min(r0 + b, 32767);
  }
  close_band(d->n);
  (void)hipSetDevice(ctx->device);
  HIPCHK(ctx, d->d_blk.upload(ctx->stream, d->blk_int.data(), sizeof(int32_t) * d->n), "dist blk");
  HIPCHK(ctx, d->d_bands.upload(ctx->stream, d->bands.data(), sizeof(Band) * d->bands.size()),
         "dist bands");
  HIPCHK(ctx, d->d_band_of.upload(ctx->stream, d->band_of.data(), sizeof(int32_t) * d->n),
         "dist band_of");
  HIPCHK(ctx, d->d_pos_in_band.upload(ctx->stream, d->pos_in_band.data(), sizeof(int32_t) * d->n),
         "dist pos_in_band");
  HIPCHK(ctx, d->d_perm.upload(ctx->stream, d->perm.data(), sizeof(int32_t) * d->n), "dist perm");
  std::vector<int64_t> elem0(d->n + 1, 0);  // alive until the stream sync below
  for (int32_t p = 0; p < d->n; ++p) elem0[p + 1] = elem0[p] + d->blk_int[p];
  d->n_elem = elem0[d->n];
  HIPCHK(ctx, d->d_elem0.upload(ctx->stream, elem0.data(), sizeof(int64_t) * (d->n + 1)),
         "dist elem0");
  d->n_groups = 1;
  d->group_int.assign(d->n, 0);
  HIPCHK(ctx, d->d_group.upload(ctx->stream, d->group_int.data(), sizeof(int32_t) * d->n),
         "dist groups");
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "dist sync");
  *out = d.release();
  return LAMB_HIP_OK;
}

int32_t lamb_hip_dist_destroy(lamb_hip_dist* d) {
  if (!d) return LAMB_HIP_OK;
  (void)hipSetDevice(d->ctx->device);
  d->d_blk.release();
  d->d_bands.release();
  d->d_band_of.release();
  d->d_pos_in_band.release();
  d->d_perm.release();
  d->d_elem0.release();
  d->d_group.release();
  delete d;
  return LAMB_HIP_OK;
}

// ---------------------------------------------------------------------------
// matrices
// ---------------------------------------------------------------------------
int32_t lamb_hip_matrix_create(lamb_hip_dist* dist, lamb_hip_mat** out) {
  if (!dist || !out) return LAMB_HIP_FAILURE;
  auto* m = new lamb_hip_mat();
  m->dist = dist;
  m->row_p.assign(dist->n + 1, 0);
  *out = m;
  return LAMB_HIP_OK;
}

int32_t lamb_hip_matrix_destroy(lamb_hip_mat* m) {
  if (!m) return LAMB_HIP_OK;
  (void)hipSetDevice(m->dist->ctx->device);
  m->release_device();
  delete m;
  return LAMB_HIP_OK;
}

void lamb_hip_mat::release_device() {
  d_data.release();
  d_row_p.release();
  d_col.release();
  d_off.release();
  d_rows.release();
  d_nsq.release();
  release_spares();
}

void lamb_hip_mat::release_spares() {
  s_data.release();
  s_row_p.release();
  s_col.release();
  s_off.release();
  s_rows.release();
}

void lamb_hip_mat::invalidate_derived() {
  nsq_valid = false;
  host_data_valid = false;
  host_data.clear();
  host_data.shrink_to_fit();
}

int32_t lamb_hip_matrix_copy(const lamb_hip_mat* src, lamb_hip_mat* dst) {
  if (!src || !dst) return LAMB_HIP_FAILURE;
  lamb_hip_ctx* ctx = src->dist->ctx;
  if (src->dist != dst->dist)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "matrix_copy: matrices have different distributions");
  if (src == dst) return LAMB_HIP_OK;
  if (values_missing(src)) return refuse_missing_values(ctx, "matrix_copy");
  (void)hipSetDevice(ctx->device);
  hipStream_t s = ctx->stream;
  auto d2d = [&](DevBuf& to, const DevBuf& from, size_t bytes) -> hipError_t {
    hipError_t e = to.ensure(std::max<size_t>(bytes, 8));
    if (e != hipSuccess || bytes == 0) return e;
    return hipMemcpyAsync(to.p, from.p, bytes, hipMemcpyDeviceToDevice, s);
  };
  const size_t nb = (size_t)src->nblk;
  HIPCHK(ctx, d2d(dst->d_data, src->d_data, sizeof(double) * (size_t)src->ndata), "copy data");
  HIPCHK(ctx, d2d(dst->d_row_p, src->d_row_p, sizeof(int64_t) * (size_t)(src->dist->n + 1)),
         "copy row_p");
  HIPCHK(ctx, d2d(dst->d_col, src->d_col, sizeof(int32_t) * nb), "copy col");
  HIPCHK(ctx, d2d(dst->d_off, src->d_off, sizeof(int64_t) * nb), "copy off");
  HIPCHK(ctx, d2d(dst->d_rows, src->d_rows, sizeof(int32_t) * nb), "copy rows");
  dst->host_index_valid = src->host_index_valid;
  dst->row_p = src->row_p;
  dst->col = src->col;
  dst->off = src->off;
  dst->rows = src->rows;
  dst->nblk = src->nblk;
  dst->ndata = src->ndata;
  dst->nelems = src->nelems;
  dst->invalidate_derived();
  dst->staged_sectors.clear();
  dst->staged_lists.clear();
  HIPCHK(ctx, stream_sync(ctx), "copy sync");
  return LAMB_HIP_OK;
}

hipError_t lamb_hip_mat::download_index(hipStream_t s) {
  if (host_index_valid) return hipSuccess;
  const size_t nb = (size_t)nblk;
  row_p.resize((size_t)dist->n + 1);
  col.resize(nb);
  off.resize(nb);
  rows.resize(nb);
  hipError_t e;
  if ((e = hipMemcpyAsync(row_p.data(), d_row_p.p, sizeof(int64_t) * row_p.size(),
                          hipMemcpyDeviceToHost, s)) != hipSuccess)
    return e;
  if (nb) {
    if ((e = hipMemcpyAsync(col.data(), d_col.p, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, s)) !=
        hipSuccess)
      return e;
    if ((e = hipMemcpyAsync(off.data(), d_off.p, sizeof(int64_t) * nb, hipMemcpyDeviceToHost, s)) !=
        hipSuccess)
      return e;
    if ((e = hipMemcpyAsync(rows.data(), d_rows.p, sizeof(int32_t) * nb, hipMemcpyDeviceToHost,
                            s)) != hipSuccess)
      return e;
  }
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  host_index_valid = true;
  return hipSuccess;
}

hipError_t lamb_hip_mat::upload_index(hipStream_t s) {
  hipError_t e;
  if ((e = d_row_p.upload(s, row_p.data(), sizeof(int64_t) * row_p.size())) != hipSuccess) return e;
  if ((e = d_col.upload(s, col.data(), sizeof(int32_t) * col.size())) != hipSuccess) return e;
  if ((e = d_off.upload(s, off.data(), sizeof(int64_t) * off.size())) != hipSuccess) return e;
  if ((e = d_rows.upload(s, rows.data(), sizeof(int32_t) * rows.size())) != hipSuccess) return e;
  return hipSuccess;
}

int32_t lamb_hip_sector_upload(lamb_hip_mat* m, int32_t sr, int32_t sc, int64_t n_blks,
                               const int64_t* row_p, const int16_t* col_i, const int64_t* extents,
                               const double* data, int64_t data_len) {
  if (!m) return LAMB_HIP_FAILURE;
  lamb_hip_dist* d = m->dist;
  lamb_hip_ctx* ctx = d->ctx;
  if (sr < 1 || sr > d->n_subsets || sc < 1 || sc > d->n_subsets)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "sector_upload: subsets (%d,%d) outside 1..%d", sr, sc, d->n_subsets);
  const int64_t n_rows = (int64_t)d->subset_atoms[sr - 1].size();
  if (n_blks < 0 || data_len < 0 || (n_blks > 0 && (!col_i || !extents)) || !row_p ||
      (data_len > 0 && !data))
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS, "sector_upload: bad arguments");
  if (row_p[0] != 0 || row_p[n_rows] != n_blks)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "sector_upload: row_p[0]=%lld row_p[n_rows]=%lld, expected 0 and %lld",
                     (long long)row_p[0], (long long)row_p[n_rows], (long long)n_blks);
  // row r's blocks are row_p[r]+1 .. row_p[r+1] (src/lamb_index_methods.F90:91-92):
  // the offsets must be non-decreasing inside [0, n_blks], or the commit
  // would read col_i / extents past their ends
  for (int64_t r = 0; r < n_rows; ++r)
    if (row_p[r] < 0 || row_p[r] > row_p[r + 1] || row_p[r + 1] > n_blks)
      return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                       "sector_upload: row_p[%lld..%lld] = %lld, %lld is not a non-decreasing "
                       "pair inside [0, %lld]",
                       (long long)r, (long long)r + 1, (long long)row_p[r],
                       (long long)row_p[r + 1], (long long)n_blks);
  StagedSector st;
  st.sr = sr;
  st.sc = sc;
  st.row_p.assign(row_p, row_p + n_rows + 1);
  st.col_i.assign(col_i, col_i + n_blks);
  st.ext.assign(extents, extents + 2 * n_blks);
  st.data.assign(data, data + data_len);
  m->staged_sectors.push_back(std::move(st));
  return LAMB_HIP_OK;
}

int32_t lamb_hip_distribute(lamb_hip_mat* m, int64_t n, const int32_t* subset_idx,
                            const int16_t* local_idx, const int64_t* extents, const double* data,
                            int64_t data_len) {
  if (!m) return LAMB_HIP_FAILURE;
  lamb_hip_ctx* ctx = m->dist->ctx;
  if (n < 0 || data_len < 0 || (n > 0 && (!subset_idx || !local_idx || !extents)) ||
      (data_len > 0 && !data))
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS, "distribute: bad arguments");
  StagedList st;
  st.subset_idx.assign(subset_idx, subset_idx + 2 * n);
  st.local_idx.assign(local_idx, local_idx + 2 * n);
  st.ext.assign(extents, extents + 2 * n);
  st.data.assign(data, data + data_len);
  m->staged_lists.push_back(std::move(st));
  return LAMB_HIP_OK;
}

int32_t lamb_hip_matrix_commit(lamb_hip_mat* m) {
  if (!m) return LAMB_HIP_FAILURE;
  lamb_hip_dist* d = m->dist;
  lamb_hip_ctx* ctx = d->ctx;
  (void)hipSetDevice(ctx->device);
  struct Blk {
    int32_t p, q;
    int64_t off;
  };
  std::vector<Blk> blks;
  int64_t total = 0;
  std::vector<int64_t> base;
  // extents [lo, hi] per data store, to refuse two blocks sharing values
  struct Ext {
    int64_t lo, hi;
    int32_t gr, gc;
  };
  std::vector<Ext> store_ext;
  auto check_overlap = [&](int64_t store_no) -> int32_t {
    std::sort(store_ext.begin(), store_ext.end(),
              [](const Ext& a, const Ext& b) { return a.lo < b.lo; });
    for (size_t t = 1; t < store_ext.size(); ++t)
      if (store_ext[t].lo <= store_ext[t - 1].hi)
        return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                         "commit: blocks (%d,%d) and (%d,%d) overlap in data store %lld "
                         "(extents [%lld,%lld] and [%lld,%lld])",
                         store_ext[t - 1].gr + 1, store_ext[t - 1].gc + 1, store_ext[t].gr + 1,
                         store_ext[t].gc + 1, (long long)store_no + 1,
                         (long long)store_ext[t - 1].lo, (long long)store_ext[t - 1].hi,
                         (long long)store_ext[t].lo, (long long)store_ext[t].hi);
    store_ext.clear();
    return LAMB_HIP_OK;
  };
  auto atom_of = [&](int32_t s, int64_t local1, int32_t& g) -> bool {
    if (s < 1 || s > d->n_subsets) return false;
    const auto& v = d->subset_atoms[s - 1];
    if (local1 < 1 || local1 > (int64_t)v.size()) return false;
    g = v[local1 - 1];
    return true;
  };
  auto add = [&](int32_t gr, int32_t gc, int64_t lo, int64_t hi, int64_t store_len,
                 int64_t store_base) -> int32_t {
    const int64_t len = (int64_t)d->blk[gr] * d->blk[gc];
    if (hi - lo + 1 != len || lo < 1 || hi > store_len)
      return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                       "commit: block (%d,%d) extent [%lld,%lld] does not hold %lld values in a "
                       "store of %lld",
                       gr + 1, gc + 1, (long long)lo, (long long)hi, (long long)len,
                       (long long)store_len);
    blks.push_back({d->perm[gr], d->perm[gc], store_base + lo - 1});
    if (len > 0) store_ext.push_back({lo, hi, gr, gc});
    return LAMB_HIP_OK;
  };
  int64_t store_no = 0;
  for (const auto& st : m->staged_sectors) {
    base.push_back(total);
    const int64_t sb = total;
    const int64_t n_rows = (int64_t)st.row_p.size() - 1;
    for (int64_t r = 0; r < n_rows; ++r) {
      int32_t gr;
      if (!atom_of(st.sr, r + 1, gr))
        return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_INTERNAL_ERROR, "commit: bad row");
      for (int64_t b = st.row_p[r]; b < st.row_p[r + 1]; ++b) {
        int32_t gc;
        if (!atom_of(st.sc, st.col_i[b], gc))
          return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                           "commit: sector (%d,%d) block %lld has column %d outside the subset",
                           st.sr, st.sc, (long long)b + 1, (int)st.col_i[b]);
        int32_t rc = add(gr, gc, st.ext[2 * b], st.ext[2 * b + 1], (int64_t)st.data.size(), sb);
        if (rc) return rc;
      }
    }
    if (int32_t rc = check_overlap(store_no++)) return rc;
    total += (int64_t)st.data.size();
  }
  for (const auto& st : m->staged_lists) {
    base.push_back(total);
    const int64_t n = (int64_t)st.local_idx.size() / 2;
    for (int64_t b = 0; b < n; ++b) {
      int32_t gr, gc;
      if (!atom_of(st.subset_idx[2 * b], st.local_idx[2 * b], gr) ||
          !atom_of(st.subset_idx[2 * b + 1], st.local_idx[2 * b + 1], gc))
        return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                         "distribute: block %lld has an invalid subset/local index", (long long)b + 1);
      int32_t rc = add(gr, gc, st.ext[2 * b], st.ext[2 * b + 1], (int64_t)st.data.size(), total);
      if (rc) return rc;
    }
    if (int32_t rc = check_overlap(store_no++)) return rc;
    total += (int64_t)st.data.size();
  }
  std::sort(blks.begin(), blks.end(), [](const Blk& a, const Blk& b) {
    return a.p != b.p ? a.p < b.p : a.q < b.q;
  });
  for (size_t t = 1; t < blks.size(); ++t)
    if (blks[t].p == blks[t - 1].p && blks[t].q == blks[t - 1].q)
      return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                       "commit: block (%d,%d) given twice", d->iperm[blks[t].p] + 1,
                       d->iperm[blks[t].q] + 1);
  const int64_t nb = (int64_t)blks.size();
  m->row_p.assign(d->n + 1, 0);
  m->col.resize(nb);
  m->off.resize(nb);
  m->rows.resize(nb);
  m->nelems = 0;
  for (int64_t t = 0; t < nb; ++t) {
    m->row_p[blks[t].p + 1] += 1;
    m->col[t] = blks[t].q;
    m->off[t] = blks[t].off;
    m->rows[t] = blks[t].p;
    m->nelems += (int64_t)d->blk_int[blks[t].p] * d->blk_int[blks[t].q];
  }
  for (int32_t p = 0; p < d->n; ++p) m->row_p[p + 1] += m->row_p[p];
  m->nblk = nb;
  m->ndata = total;
  m->host_index_valid = true;
  m->invalidate_derived();
  HIPCHK(ctx, m->d_data.ensure(sizeof(double) * (size_t)std::max<int64_t>(total, 1)), "pool alloc");
  size_t si = 0;
  for (const auto& st : m->staged_sectors) {
    if (!st.data.empty())
      HIPCHK(ctx, hipMemcpyAsync((double*)m->d_data.p + base[si], st.data.data(),
                                 sizeof(double) * st.data.size(), hipMemcpyHostToDevice,
                                 ctx->stream), "pool upload");
    ++si;
  }
  for (const auto& st : m->staged_lists) {
    if (!st.data.empty())
      HIPCHK(ctx, hipMemcpyAsync((double*)m->d_data.p + base[si], st.data.data(),
                                 sizeof(double) * st.data.size(), hipMemcpyHostToDevice,
                                 ctx->stream), "pool upload");
    ++si;
  }
  HIPCHK(ctx, m->upload_index(ctx->stream), "index upload");
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "commit sync");
  m->staged_sectors.clear();
  m->staged_lists.clear();
  return LAMB_HIP_OK;
}

int32_t lamb_hip_matrix_info(const lamb_hip_mat* m, int64_t* n_blks, int64_t* n_elems) {
  if (!m) return LAMB_HIP_FAILURE;
  if (n_blks) *n_blks = m->nblk;
  if (n_elems) *n_elems = m->nelems;
  return LAMB_HIP_OK;
}

static int32_t ensure_host_data(lamb_hip_mat* m) {
  lamb_hip_ctx* ctx = m->dist->ctx;
  if (m->host_data_valid) return LAMB_HIP_OK;
  m->host_data.resize((size_t)m->ndata);
  if (m->ndata)
    HIPCHK(ctx, hipMemcpyAsync(m->host_data.data(), m->d_data.p, sizeof(double) * m->ndata,
                               hipMemcpyDeviceToHost, ctx->stream), "download");
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "download sync");
  m->host_data_valid = true;
  return LAMB_HIP_OK;
}

// Blocks of sector (sr, sc) as (local row, local col, internal block id), sorted.
static void sector_blocks(const lamb_hip_mat* m, int32_t sr, int32_t sc,
                          std::vector<std::pair<int32_t, int64_t>>& row_blocks,
                          std::vector<int64_t>& row_start) {
  const lamb_hip_dist* d = m->dist;
  const auto& atoms = d->subset_atoms[sr - 1];
  row_blocks.clear();
  row_start.assign(atoms.size() + 1, 0);
  for (size_t r = 0; r < atoms.size(); ++r) {
    const int32_t p = d->perm[atoms[r]];
    const size_t first = row_blocks.size();
    for (int64_t b = m->row_p[p]; b < m->row_p[p + 1]; ++b) {
      const int32_t g = d->iperm[m->col[b]];
      if (d->subset_of[g] != sc) continue;
      row_blocks.push_back({d->local_of[g], b});
    }
    std::sort(row_blocks.begin() + first, row_blocks.end());
    row_start[r + 1] = (int64_t)row_blocks.size();
  }
}

int32_t lamb_hip_sector_query(lamb_hip_mat* m, int32_t sr, int32_t sc, int64_t* n_blks,
                              int64_t* data_len) {
  if (!m) return LAMB_HIP_FAILURE;
  lamb_hip_dist* d = m->dist;
  if (sr < 1 || sr > d->n_subsets || sc < 1 || sc > d->n_subsets)
    return set_error(d->ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "sector_query: subsets (%d,%d) outside 1..%d", sr, sc, d->n_subsets);
  (void)hipSetDevice(d->ctx->device);
  HIPCHK(d->ctx, m->download_index(d->ctx->stream), "index download");
  std::vector<std::pair<int32_t, int64_t>> rb;
  std::vector<int64_t> rs;
  sector_blocks(m, sr, sc, rb, rs);
  int64_t len = 0;
  for (auto& x : rb) len += (int64_t)d->blk_int[m->rows[x.second]] * d->blk_int[m->col[x.second]];
  if (n_blks) *n_blks = (int64_t)rb.size();
  if (data_len) *data_len = len;
  return LAMB_HIP_OK;
}

int32_t lamb_hip_sector_download(lamb_hip_mat* m, int32_t sr, int32_t sc, int64_t* row_p,
                                 int16_t* col_i, int64_t* extents, double* data) {
  if (!m) return LAMB_HIP_FAILURE;
  lamb_hip_dist* d = m->dist;
  lamb_hip_ctx* ctx = d->ctx;
  if (sr < 1 || sr > d->n_subsets || sc < 1 || sc > d->n_subsets)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "sector_download: subsets (%d,%d) outside 1..%d", sr, sc, d->n_subsets);
  if (values_missing(m)) return refuse_missing_values(ctx, "sector_download");
  (void)hipSetDevice(ctx->device);
  HIPCHK(ctx, m->download_index(ctx->stream), "index download");
  int32_t rc = ensure_host_data(m);
  if (rc) return rc;
  std::vector<std::pair<int32_t, int64_t>> rb;
  std::vector<int64_t> rs;
  sector_blocks(m, sr, sc, rb, rs);
  if (row_p)
    for (size_t r = 0; r < rs.size(); ++r) row_p[r] = rs[r];
  int64_t pos = 0;
  for (size_t t = 0; t < rb.size(); ++t) {
    const int64_t b = rb[t].second;
    const int64_t len = (int64_t)d->blk_int[m->rows[b]] * d->blk_int[m->col[b]];
    if (col_i) col_i[t] = (int16_t)rb[t].first;
    if (extents) {
      extents[2 * t] = pos + 1;
      extents[2 * t + 1] = pos + len;
    }
    if (data && len) memcpy(data + pos, m->host_data.data() + m->off[b], sizeof(double) * len);
    pos += len;
  }
  return LAMB_HIP_OK;
}

int32_t lamb_hip_trace(const lamb_hip_mat* m, double* trace) {
  if (!m || !trace) return LAMB_HIP_FAILURE;
  lamb_hip_dist* d = m->dist;
  lamb_hip_ctx* ctx = d->ctx;
  if (values_missing(m)) return refuse_missing_values(ctx, "trace");
  (void)hipSetDevice(ctx->device);
  DevBuf part;
  HIPCHK(ctx, part.ensure(sizeof(double) * std::max(d->n, 1)), "trace alloc");
  HIPCHK(ctx, launch_trace(ctx->stream, d->n, (const int64_t*)m->d_row_p.p,
                           (const int32_t*)m->d_col.p, (const int64_t*)m->d_off.p,
                           (const int32_t*)d->d_blk.p, (const double*)m->d_data.p,
                           (double*)part.p), "trace kernel");
  std::vector<double> h(d->n);
  if (d->n)
    HIPCHK(ctx, hipMemcpyAsync(h.data(), part.p, sizeof(double) * d->n, hipMemcpyDeviceToHost,
                               ctx->stream), "trace download");
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "trace sync");
  part.release();
  // sum in global atom order
  double t = 0.0;
  for (int32_t g = 0; g < d->n; ++g) t += h[d->perm[g]];
  *trace = t;
  return LAMB_HIP_OK;
}

}  // extern "C"
