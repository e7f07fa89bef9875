// Adjoint of C = A*B for reverse-mode differentiation, and the matrix
// plumbing a differentiable caller needs:
//   dA = (G B^T) sampled on A's pattern,  dB = (A^T G) sampled on B's pattern
// (SDDMM: only blocks stored in the operand are computed). Each is one
// masked multiply by an explicit transpose, so it runs on the same MFMA tile
// kernels (FP64 or BF16 split) as the forward product. Transposes build their
// index and move the block data on the device (a stable sort by column, so
// rows of the transpose keep ascending columns). Values can be read and written in a
// canonical order (blocks in CSR order of the internal atom order, each
// column-major, back to back) through device pointers.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "common.h"
#include "device_util.h"
#include "internal.h"

using namespace lamb;

namespace lamb {

LAMB_GUARD_UNIT(adjoint)


// one thread per block of XT: XT(j, i) = X(i, j)^T; src_off[b] is the offset
// of X(i, j) in X's pool
__global__ __launch_bounds__(256) void transpose_blocks_kernel(
    int64_t nblk, const int32_t* __restrict__ t_rows, const int32_t* __restrict__ t_col,
    const int64_t* __restrict__ t_off, const int64_t* __restrict__ src_off,
    const int32_t* __restrict__ blk, const double* __restrict__ src, double* __restrict__ dst) {
  LAMB_GUARD;
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  const int32_t j = t_rows[b], i = t_col[b];
  const int m = blk[i], n = blk[j];  // X block is m x n
  const double* x = src + src_off[b];
  double* y = dst + t_off[b];
  for (int c = 0; c < n; ++c)
    for (int r = 0; r < m; ++r) y[(int64_t)r * n + c] = x[(int64_t)c * m + r];
}

// one thread per block of `like`: its values from src (same block found by
// binary search in src's row) or zero
__global__ __launch_bounds__(256) void project_blocks_kernel(
    int64_t nblk, const int32_t* __restrict__ l_rows, const int32_t* __restrict__ l_col,
    const int64_t* __restrict__ l_off, const int64_t* __restrict__ s_row_p,
    const int32_t* __restrict__ s_col, const int64_t* __restrict__ s_off,
    const int32_t* __restrict__ blk, const double* __restrict__ src, double* __restrict__ dst) {
  LAMB_GUARD;
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  const int32_t i = l_rows[b], j = l_col[b];
  const int64_t len = (int64_t)blk[i] * blk[j];
  int64_t lo = s_row_p[i], hi = s_row_p[i + 1] - 1, found = -1;
  while (lo <= hi) {
    const int64_t mid = (lo + hi) >> 1;
    const int32_t c = s_col[mid];
    if (c == j) {
      found = mid;
      break;
    }
    if (c < j) lo = mid + 1;
    else hi = mid - 1;
  }
  double* y = dst + l_off[b];
  if (found >= 0) {
    const double* x = src + s_off[found];
    for (int64_t e = 0; e < len; ++e) y[e] = x[e];
  } else {
    for (int64_t e = 0; e < len; ++e) y[e] = 0.0;
  }
}

__global__ void iota_i32_kernel(int64_t n, int32_t* __restrict__ out) {
  LAMB_GUARD;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = (int32_t)t;
}

// position p of XT in CSR order: X's block ids[p] = (i, j) becomes XT's
// (j, i); its source offset and element count
__global__ __launch_bounds__(256) void transpose_gather_kernel(
    int64_t nblk, const int32_t* __restrict__ ids, const int32_t* __restrict__ t_rows,
    const int32_t* __restrict__ x_rows, const int64_t* __restrict__ x_off,
    const int32_t* __restrict__ blk, int32_t* __restrict__ t_col, int64_t* __restrict__ src_off,
    int64_t* __restrict__ sz) {
  LAMB_GUARD;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nblk) return;
  const int32_t b = ids[p];
  const int32_t i = x_rows[b];
  t_col[p] = i;
  src_off[p] = x_off[b];
  sz[p] = (int64_t)blk[i] * blk[t_rows[p]];
}

// row_p[j] = first position whose row is >= j, over XT's sorted rows
__global__ void transpose_row_p_kernel(int32_t n, int64_t nblk, const int32_t* __restrict__ t_rows,
                                       int64_t* __restrict__ row_p) {
  LAMB_GUARD;
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j > n) return;
  int64_t lo = 0, hi = nblk;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (t_rows[mid] < j) lo = mid + 1;
    else hi = mid;
  }
  row_p[j] = lo;
}

__global__ void block_sizes_kernel(int64_t nblk, const int32_t* __restrict__ rows,
                                   const int32_t* __restrict__ col,
                                   const int32_t* __restrict__ blk, int64_t* __restrict__ sz) {
  LAMB_GUARD;
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nblk) sz[b] = (int64_t)blk[rows[b]] * blk[col[b]];
}

// one wave per block: canonical values <-> pool (to_pool selects the direction)
__global__ __launch_bounds__(256) void move_values_kernel(
    int64_t nblk, const int32_t* __restrict__ rows, const int32_t* __restrict__ col,
    const int64_t* __restrict__ off, const int64_t* __restrict__ voff,
    const int32_t* __restrict__ blk, int to_pool, double* __restrict__ pool,
    double* __restrict__ vals) {
  LAMB_GUARD;
  const int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= nblk) return;
  const int64_t len = (int64_t)blk[rows[b]] * blk[col[b]];
  double* p = pool + off[b];
  double* v = vals + voff[b];
  if (to_pool)
    for (int64_t e = lane; e < len; e += 64) p[e] = v[e];
  else
    for (int64_t e = lane; e < len; e += 64) v[e] = p[e];
}

}  // namespace lamb

#define HIPCHK(ctx, expr, what)                            \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return hip_error(ctx, e_, what); \
  } while (0)

namespace {

unsigned grid1(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

// XT = X^T (same distribution), index and values built on the device. X's
// blocks in CSR order have ascending rows, so a stable radix sort of them by
// column lists XT's blocks in CSR order with each row's columns ascending.
int32_t transpose_into(lamb_hip_mat* X, lamb_hip_mat* XT) {
  lamb_hip_dist* d = X->dist;
  lamb_hip_ctx* ctx = d->ctx;
  if (values_missing(X)) return refuse_missing_values(ctx, "transpose");
  hipStream_t s = ctx->stream;
  const int32_t n = d->n;
  const int64_t nb = X->nblk;
  if (nb > INT32_MAX)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_OVERFLOW_ERROR,
                     "transpose: %lld blocks exceed the sort's 32-bit count", (long long)nb);
  const size_t cap = (size_t)std::max<int64_t>(nb, 1);
  HIPCHK(ctx, XT->d_row_p.ensure(sizeof(int64_t) * (size_t)(n + 1)), "transpose: row_p");
  HIPCHK(ctx, XT->d_rows.ensure(sizeof(int32_t) * cap), "transpose: rows");
  HIPCHK(ctx, XT->d_col.ensure(sizeof(int32_t) * cap), "transpose: col");
  HIPCHK(ctx, XT->d_off.ensure(sizeof(int64_t) * (cap + 1)), "transpose: off");
  DevBuf ids, sorted_ids, src_off, sz, scr, temp;
  HIPCHK(ctx, ids.ensure(sizeof(int32_t) * cap), "transpose: ids");
  HIPCHK(ctx, sorted_ids.ensure(sizeof(int32_t) * cap), "transpose: ids");
  HIPCHK(ctx, src_off.ensure(sizeof(int64_t) * cap), "transpose: offsets");
  HIPCHK(ctx, sz.ensure(sizeof(int64_t) * cap), "transpose: sizes");
  HIPCHK(ctx, scr.ensure(sizeof(int64_t) * (size_t)scan_scratch_elems(nb)), "transpose: scan");
  int64_t total = 0;
  if (nb) {
    int end_bit = 1;
    while (end_bit < 31 && ((int64_t)1 << end_bit) < n) ++end_bit;
    hipLaunchKernelGGL(iota_i32_kernel, dim3(grid1(nb)), dim3(256), 0, s, nb, (int32_t*)ids.p);
    HIPCHK(ctx, hipGetLastError(), "transpose: ids kernel");
    size_t temp_bytes = 0;
    HIPCHK(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, (const int32_t*)X->d_col.p,
                                                   (int32_t*)XT->d_rows.p, (const int32_t*)ids.p,
                                                   (int32_t*)sorted_ids.p, (int)nb, 0, end_bit, s),
           "transpose: sort size");
    HIPCHK(ctx, temp.ensure(std::max<size_t>(temp_bytes, 8)), "transpose: sort scratch");
    HIPCHK(ctx, hipcub::DeviceRadixSort::SortPairs(temp.p, temp_bytes, (const int32_t*)X->d_col.p,
                                                   (int32_t*)XT->d_rows.p, (const int32_t*)ids.p,
                                                   (int32_t*)sorted_ids.p, (int)nb, 0, end_bit, s),
           "transpose: sort");
    hipLaunchKernelGGL(transpose_gather_kernel, dim3(grid1(nb)), dim3(256), 0, s, nb,
                       (const int32_t*)sorted_ids.p, (const int32_t*)XT->d_rows.p,
                       (const int32_t*)X->d_rows.p, (const int64_t*)X->d_off.p,
                       (const int32_t*)d->d_blk.p, (int32_t*)XT->d_col.p, (int64_t*)src_off.p,
                       (int64_t*)sz.p);
    HIPCHK(ctx, hipGetLastError(), "transpose: gather kernel");
    HIPCHK(ctx, exclusive_scan(s, (const int64_t*)sz.p, (int64_t*)XT->d_off.p, nb,
                               (int64_t*)scr.p), "transpose: offsets scan");
    HIPCHK(ctx, hipMemcpyAsync(&total, (const int64_t*)XT->d_off.p + nb, sizeof(int64_t),
                               hipMemcpyDeviceToHost, s), "transpose: size read");
  }
  hipLaunchKernelGGL(transpose_row_p_kernel, dim3(grid1(n + 1)), dim3(256), 0, s, n, nb,
                     (const int32_t*)XT->d_rows.p, (int64_t*)XT->d_row_p.p);
  HIPCHK(ctx, hipGetLastError(), "transpose: row_p kernel");
  HIPCHK(ctx, hipStreamSynchronize(s), "transpose: index sync");
  XT->nblk = nb;
  XT->ndata = total;
  XT->nelems = total;
  XT->host_index_valid = false;
  XT->invalidate_derived();
  XT->host_data_valid = false;
  XT->staged_sectors.clear();
  XT->staged_lists.clear();
  HIPCHK(ctx, XT->d_data.ensure(sizeof(double) * (size_t)std::max<int64_t>(total, 1)),
         "transpose: data");
  if (nb)
    hipLaunchKernelGGL(transpose_blocks_kernel, dim3(grid1(nb)), dim3(256), 0, s, nb,
                       (const int32_t*)XT->d_rows.p, (const int32_t*)XT->d_col.p,
                       (const int64_t*)XT->d_off.p, (const int64_t*)src_off.p,
                       (const int32_t*)d->d_blk.p, (const double*)X->d_data.p,
                       (double*)XT->d_data.p);
  HIPCHK(ctx, hipGetLastError(), "transpose kernel");
  HIPCHK(ctx, hipStreamSynchronize(s), "transpose sync");
  return LAMB_HIP_OK;
}

// dst = like's pattern with src's values (zero where src has no block);
// only like's index is read
int32_t project_into(const lamb_hip_mat* src, const lamb_hip_mat* like, lamb_hip_mat* dst) {
  lamb_hip_ctx* ctx = src->dist->ctx;
  if (values_missing(src)) return refuse_missing_values(ctx, "project");
  hipStream_t s = ctx->stream;
  auto d2d = [&](DevBuf& to, const DevBuf& from, size_t bytes) -> hipError_t {
    hipError_t e = to.ensure(std::max<size_t>(bytes, 8));
    if (e != hipSuccess || bytes == 0) return e;
    return hipMemcpyAsync(to.p, from.p, bytes, hipMemcpyDeviceToDevice, s);
  };
  const size_t nb = (size_t)like->nblk;
  HIPCHK(ctx, d2d(dst->d_row_p, like->d_row_p, sizeof(int64_t) * (size_t)(like->dist->n + 1)),
         "project: row_p");
  HIPCHK(ctx, d2d(dst->d_col, like->d_col, sizeof(int32_t) * nb), "project: col");
  HIPCHK(ctx, d2d(dst->d_off, like->d_off, sizeof(int64_t) * nb), "project: off");
  HIPCHK(ctx, d2d(dst->d_rows, like->d_rows, sizeof(int32_t) * nb), "project: rows");
  HIPCHK(ctx, dst->d_data.ensure(sizeof(double) * (size_t)std::max<int64_t>(like->ndata, 1)),
         "project: data");
  dst->host_index_valid = like->host_index_valid;
  dst->row_p = like->row_p;
  dst->col = like->col;
  dst->off = like->off;
  dst->rows = like->rows;
  dst->nblk = like->nblk;
  dst->ndata = like->ndata;
  dst->nelems = like->nelems;
  dst->staged_sectors.clear();
  dst->staged_lists.clear();
  if (dst->nblk)
    hipLaunchKernelGGL(project_blocks_kernel, dim3(grid1(dst->nblk)), dim3(256), 0, s, dst->nblk,
                       (const int32_t*)dst->d_rows.p, (const int32_t*)dst->d_col.p,
                       (const int64_t*)dst->d_off.p, (const int64_t*)src->d_row_p.p,
                       (const int32_t*)src->d_col.p, (const int64_t*)src->d_off.p,
                       (const int32_t*)src->dist->d_blk.p, (const double*)src->d_data.p,
                       (double*)dst->d_data.p);
  HIPCHK(ctx, hipGetLastError(), "project kernel");
  HIPCHK(ctx, hipStreamSynchronize(s), "project sync");
  dst->invalidate_derived();
  return LAMB_HIP_OK;
}

int32_t move_values(lamb_hip_mat* X, double* vals, int64_t n, bool to_pool) {
  lamb_hip_dist* d = X->dist;
  lamb_hip_ctx* ctx = d->ctx;
  if (n != X->nelems)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "values: %lld values given, the matrix stores %lld", (long long)n,
                     (long long)X->nelems);
  if (X->nblk == 0) return LAMB_HIP_OK;
  hipStream_t s = ctx->stream;
  if (to_pool)
    HIPCHK(ctx, X->d_data.ensure(sizeof(double) * (size_t)std::max<int64_t>(X->ndata, 1)),
           "values: pool");
  else if (!X->d_data.p)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_PRECONDITION_FAILED,
                     "values: the matrix's values were released");
  DevBuf sz, voff, scr;
  HIPCHK(ctx, sz.ensure(sizeof(int64_t) * (size_t)X->nblk), "values: sizes");
  HIPCHK(ctx, voff.ensure(sizeof(int64_t) * (size_t)(X->nblk + 1)), "values: offsets");
  HIPCHK(ctx, scr.ensure(sizeof(int64_t) * (size_t)scan_scratch_elems(X->nblk)), "values: scan");
  hipLaunchKernelGGL(block_sizes_kernel, dim3(grid1(X->nblk)), dim3(256), 0, s, X->nblk,
                     (const int32_t*)X->d_rows.p, (const int32_t*)X->d_col.p,
                     (const int32_t*)d->d_blk.p, (int64_t*)sz.p);
  HIPCHK(ctx, hipGetLastError(), "values: sizes kernel");
  HIPCHK(ctx, exclusive_scan(s, (const int64_t*)sz.p, (int64_t*)voff.p, X->nblk,
                             (int64_t*)scr.p), "values: scan");
  hipLaunchKernelGGL(move_values_kernel, dim3((unsigned)((X->nblk + 3) / 4)), dim3(256), 0, s,
                     X->nblk, (const int32_t*)X->d_rows.p, (const int32_t*)X->d_col.p,
                     (const int64_t*)X->d_off.p, (const int64_t*)voff.p,
                     (const int32_t*)d->d_blk.p, to_pool ? 1 : 0, (double*)X->d_data.p, vals);
  HIPCHK(ctx, hipGetLastError(), "values kernel");
  HIPCHK(ctx, hipStreamSynchronize(s), "values sync");
  sz.release();
  voff.release();
  scr.release();
  if (to_pool) {
    X->invalidate_derived();
    X->host_data_valid = false;
  }
  return LAMB_HIP_OK;
}

}  // namespace

extern "C" int32_t lamb_hip_matrix_transpose(const lamb_hip_mat* X, lamb_hip_mat* XT) {
  if (!X || !XT) return LAMB_HIP_FAILURE;
  if (X == XT || X->dist != XT->dist)
    return set_error(X->dist->ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "transpose: needs a distinct output on the same distribution");
  (void)hipSetDevice(X->dist->ctx->device);
  return transpose_into(const_cast<lamb_hip_mat*>(X), XT);
}

extern "C" int32_t lamb_hip_matrix_project(const lamb_hip_mat* src, const lamb_hip_mat* like,
                                           lamb_hip_mat* dst) {
  if (!src || !like || !dst) return LAMB_HIP_FAILURE;
  if (src->dist != like->dist || dst->dist != src->dist || dst == src)
    return set_error(src->dist->ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "project: one distribution, and dst distinct from src");
  (void)hipSetDevice(src->dist->ctx->device);
  if (dst == like) {  // in place on like's pattern: go through a copy of src
    lamb_hip_mat tmp;
    tmp.dist = src->dist;
    int32_t rc = lamb_hip_matrix_copy(src, &tmp);
    if (!rc) rc = project_into(&tmp, like, dst);
    tmp.release_device();
    return rc;
  }
  return project_into(src, like, dst);
}

extern "C" int32_t lamb_hip_matrix_release_values(lamb_hip_mat* X) {
  if (!X) return LAMB_HIP_FAILURE;
  (void)hipSetDevice(X->dist->ctx->device);
  X->d_data.release();
  X->invalidate_derived();
  X->host_data_valid = false;
  return LAMB_HIP_OK;
}

extern "C" int32_t lamb_hip_matrix_get_values(const lamb_hip_mat* X, double* dev_values,
                                              int64_t n) {
  if (!X || (n > 0 && !dev_values)) return LAMB_HIP_FAILURE;
  (void)hipSetDevice(X->dist->ctx->device);
  return move_values(const_cast<lamb_hip_mat*>(X), dev_values, n, false);
}

extern "C" int32_t lamb_hip_matrix_set_values(lamb_hip_mat* X, const double* dev_values,
                                              int64_t n) {
  if (!X || (n > 0 && !dev_values)) return LAMB_HIP_FAILURE;
  (void)hipSetDevice(X->dist->ctx->device);
  return move_values(X, const_cast<double*>(dev_values), n, true);
}

extern "C" int32_t lamb_hip_matrix_block_atoms(lamb_hip_mat* X, int32_t* rows, int32_t* cols) {
  if (!X || (X->nblk > 0 && (!rows || !cols))) return LAMB_HIP_FAILURE;
  lamb_hip_ctx* ctx = X->dist->ctx;
  (void)hipSetDevice(ctx->device);
  HIPCHK(ctx, X->download_index(ctx->stream), "block_atoms: index download");
  for (int64_t b = 0; b < X->nblk; ++b) {
    rows[b] = X->dist->iperm[X->rows[b]];
    cols[b] = X->dist->iperm[X->col[b]];
  }
  return LAMB_HIP_OK;
}

extern "C" int32_t lamb_hip_multiply_adjoint(lamb_hip_ctx* ctx, const lamb_hip_mat* A,
                                             const lamb_hip_mat* B, const lamb_hip_mat* G,
                                             lamb_hip_mat* dA, lamb_hip_mat* dB, int32_t flags,
                                             lamb_hip_stats* stats) {
  if (!ctx || !A || !B || !G) return LAMB_HIP_FAILURE;
  lamb_hip_dist* d = A->dist;
  if (B->dist != d || G->dist != d || (dA && dA->dist != d) || (dB && dB->dist != d))
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "multiply_adjoint: all matrices must share one distribution");
  const lamb_hip_mat* outs[2] = {dA, dB};
  const lamb_hip_mat* ins[3] = {A, B, G};
  for (auto o : outs)
    for (auto in : ins)
      if (o && o == in)
        return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                         "multiply_adjoint: dA / dB may not alias A, B or G");
  if (dA && dA == dB)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "multiply_adjoint: dA and dB must differ");
  if (values_missing(A) || values_missing(B) || values_missing(G))
    return refuse_missing_values(ctx, "multiply_adjoint");
  auto t0 = std::chrono::steady_clock::now();
  (void)hipSetDevice(ctx->device);
  lamb_hip_stats acc, s;
  memset(&acc, 0, sizeof(acc));
  auto add = [&]() {
    acc.n_triples_kept += s.n_triples_kept;
    acc.n_c_blocks_struct += s.n_c_blocks_struct;
    acc.n_c_blocks_kept += s.n_c_blocks_kept;
    acc.n_c_elems_kept += s.n_c_elems_kept;
    acc.n_tiles += s.n_tiles;
    acc.n_tile_entries += s.n_tile_entries;
    acc.flops += s.flops;
    acc.ms_symbolic += s.ms_symbolic;
    acc.ms_numeric += s.ms_numeric;
    acc.ms_index += s.ms_index;
  };
  lamb_hip_mat T, P;
  T.dist = d;
  P.dist = d;
  int32_t rc = LAMB_HIP_OK;
  if (dA) {  // dA = (G B^T) on A's pattern
    rc = transpose_into(const_cast<lamb_hip_mat*>(B), &T);
    if (!rc) rc = multiply_masked(ctx, G, &T, &P, 1.0, 0.0, 0.0, flags, &s, A);
    if (!rc) {
      add();
      rc = project_into(&P, A, dA);
    }
  }
  if (!rc && dB) {  // dB = (A^T G) on B's pattern
    rc = transpose_into(const_cast<lamb_hip_mat*>(A), &T);
    if (!rc) rc = multiply_masked(ctx, &T, G, &P, 1.0, 0.0, 0.0, flags, &s, B);
    if (!rc) {
      add();
      rc = project_into(&P, B, dB);
    }
  }
  T.release_device();
  P.release_device();
  if (stats) {
    acc.ms_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *stats = acc;
  }
  return rc;
}
