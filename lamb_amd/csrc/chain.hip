// Matrix functions of a McWeeny purification chain, per group of atoms (a
// group is one independent system of a batch; matrices are block diagonal
// over groups): trace, Frobenius norm, ||X - Y||_F, Gershgorin bounds,
// X <- a X + b I, and the chain P <- 3 P^2 - 2 P^3 itself.
//
// Device kernels write one partial per internal atom; the host adds the
// partials of each group in internal atom order, so every result is
// deterministic.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "common.h"
#include "device_util.h"
#include "internal.h"

using namespace lamb;

namespace lamb {

LAMB_GUARD_UNIT(chain)


__device__ __forceinline__ double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// one wave per atom row: sum of squares of the row's stored values
__global__ __launch_bounds__(256) void row_sumsq_kernel(int32_t n, const int64_t* __restrict__ row_p,
                                                        const int32_t* __restrict__ col,
                                                        const int64_t* __restrict__ off,
                                                        const int32_t* __restrict__ blk,
                                                        const double* __restrict__ data,
                                                        double* __restrict__ out) {
  LAMB_GUARD;
  const int32_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  double s = 0.0;
  const int m = blk[i];
  for (int64_t b = row_p[i]; b < row_p[i + 1]; ++b) {
    const int64_t len = (int64_t)m * blk[col[b]];
    const double* x = data + off[b];
    for (int64_t e = lane; e < len; e += 64) s = __builtin_fma(x[e], x[e], s);
  }
  s = wsum(s);
  if (lane == 0) out[i] = s;
}

// one wave per atom row: sum of squares of X - Y over the union of the
// rows' patterns (columns ascending in both)
__global__ __launch_bounds__(256) void row_diff_sumsq_kernel(
    int32_t n, const int64_t* __restrict__ xr, const int32_t* __restrict__ xc,
    const int64_t* __restrict__ xo, const double* __restrict__ xd, const int64_t* __restrict__ yr,
    const int32_t* __restrict__ yc, const int64_t* __restrict__ yo, const double* __restrict__ yd,
    const int32_t* __restrict__ blk, double* __restrict__ out) {
  LAMB_GUARD;
  const int32_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  double s = 0.0;
  const int m = blk[i];
  int64_t p = xr[i], q = yr[i];
  const int64_t pe = xr[i + 1], qe = yr[i + 1];
  while (p < pe || q < qe) {
    const int32_t cx = p < pe ? xc[p] : INT32_MAX, cy = q < qe ? yc[q] : INT32_MAX;
    const int32_t c = min(cx, cy);
    const int64_t len = (int64_t)m * blk[c];
    const double* x = cx == c ? xd + xo[p] : nullptr;
    const double* y = cy == c ? yd + yo[q] : nullptr;
    for (int64_t e = lane; e < len; e += 64) {
      const double d = (x ? x[e] : 0.0) - (y ? y[e] : 0.0);
      s = __builtin_fma(d, d, s);
    }
    if (cx == c) ++p;
    if (cy == c) ++q;
  }
  s = wsum(s);
  if (lane == 0) out[i] = s;
}

// one thread per atom: bounds of the Gershgorin discs of its matrix rows
// (center H_rr, radius sum of |H_rc| over c != r); missing counts atoms
// with rows but no stored diagonal block
__global__ __launch_bounds__(256) void row_gershgorin_kernel(
    int32_t n, const int64_t* __restrict__ row_p, const int32_t* __restrict__ col,
    const int64_t* __restrict__ off, const int32_t* __restrict__ blk,
    const double* __restrict__ data, double* __restrict__ lo, double* __restrict__ hi,
    unsigned long long* __restrict__ missing) {
  LAMB_GUARD;
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int m = blk[i];
  double l = INFINITY, h = -INFINITY;
  bool diag = false;
  for (int r = 0; r < m; ++r) {
    double center = 0.0, radius = 0.0;
    for (int64_t b = row_p[i]; b < row_p[i + 1]; ++b) {
      const int32_t j = col[b];
      const int nj = blk[j];
      const double* x = data + off[b];
      for (int c = 0; c < nj; ++c) {
        const double v = x[(int64_t)c * m + r];
        if (j == i && c == r) {
          center = v;
          diag = true;
        } else {
          radius += fabs(v);
        }
      }
    }
    l = fmin(l, center - radius);
    h = fmax(h, center + radius);
  }
  lo[i] = l;
  hi[i] = h;
  if (m > 0 && !diag) atomicAdd(missing, 1ull);
}

// one thread per block: X <- a[g] X + b[g] I, g the group of the row atom
__global__ __launch_bounds__(256) void scale_shift_kernel(
    int64_t nblk, const int32_t* __restrict__ rows, const int32_t* __restrict__ col,
    const int64_t* __restrict__ off, const int32_t* __restrict__ blk,
    const int32_t* __restrict__ group, const double* __restrict__ a,
    const double* __restrict__ b, double* __restrict__ data) {
  LAMB_GUARD;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nblk) return;
  const int32_t i = rows[k], j = col[k];
  const int m = blk[i], nj = blk[j];
  const int32_t g = group[i];
  const double s = a[g], t = b[g];
  double* x = data + off[k];
  for (int c = 0; c < nj; ++c)
    for (int r = 0; r < m; ++r) {
      double v = s * x[(int64_t)c * m + r];
      if (i == j && r == c) v += t;
      x[(int64_t)c * m + r] = v;
    }
}

// per-atom partials of ||X - Y||_F^2 (Y == nullptr: ||X||_F^2) into out[n]
hipError_t launch_diff_partials(hipStream_t s, const lamb_hip_mat* X, const lamb_hip_mat* Y,
                                double* out) {
  const lamb_hip_dist* d = X->dist;
  if (d->n == 0) return hipSuccess;
  const unsigned grid = (unsigned)((d->n + 3) / 4);
  if (Y)
    hipLaunchKernelGGL(row_diff_sumsq_kernel, dim3(grid), dim3(256), 0, s, d->n,
                       (const int64_t*)X->d_row_p.p, (const int32_t*)X->d_col.p,
                       (const int64_t*)X->d_off.p, (const double*)X->d_data.p,
                       (const int64_t*)Y->d_row_p.p, (const int32_t*)Y->d_col.p,
                       (const int64_t*)Y->d_off.p, (const double*)Y->d_data.p,
                       (const int32_t*)d->d_blk.p, out);
  else
    hipLaunchKernelGGL(row_sumsq_kernel, dim3(grid), dim3(256), 0, s, d->n,
                       (const int64_t*)X->d_row_p.p, (const int32_t*)X->d_col.p,
                       (const int64_t*)X->d_off.p, (const int32_t*)d->d_blk.p,
                       (const double*)X->d_data.p, out);
  return hipGetLastError();
}

hipError_t launch_trace_partials(hipStream_t s, const lamb_hip_mat* X, double* out) {
  const lamb_hip_dist* d = X->dist;
  return launch_trace(s, d->n, (const int64_t*)X->d_row_p.p, (const int32_t*)X->d_col.p,
                      (const int64_t*)X->d_off.p, (const int32_t*)d->d_blk.p,
                      (const double*)X->d_data.p, out);
}

}  // namespace lamb

#define HIPCHK(ctx, expr, what)                            \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return hip_error(ctx, e_, what); \
  } while (0)

namespace {

unsigned grid_waves(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 3) / 4); }
unsigned grid_threads(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

// per-group sums of per-atom partials (internal atom order)
void group_sums(const lamb_hip_dist* d, const std::vector<double>& part, double* out) {
  std::fill(out, out + d->n_groups, 0.0);
  for (int32_t p = 0; p < d->n; ++p) out[d->group_int[p]] += part[p];
}

int32_t atom_partials(lamb_hip_ctx* ctx, std::vector<double>& host, DevBuf& dev, int32_t n) {
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "partials sync");
  host.resize(n);
  if (n)
    HIPCHK(ctx, hipMemcpy(host.data(), dev.p, sizeof(double) * n, hipMemcpyDeviceToHost),
           "partials download");
  return LAMB_HIP_OK;
}

int32_t group_trace(lamb_hip_mat* A, double* out) {
  lamb_hip_dist* d = A->dist;
  lamb_hip_ctx* ctx = d->ctx;
  if (values_missing(A)) return refuse_missing_values(ctx, "trace");
  DevBuf part;
  HIPCHK(ctx, part.ensure(sizeof(double) * std::max(d->n, 1)), "trace partials");
  HIPCHK(ctx, launch_trace(ctx->stream, d->n, (const int64_t*)A->d_row_p.p,
                           (const int32_t*)A->d_col.p, (const int64_t*)A->d_off.p,
                           (const int32_t*)d->d_blk.p, (const double*)A->d_data.p,
                           (double*)part.p), "trace kernel");
  std::vector<double> h;
  int32_t rc = atom_partials(ctx, h, part, d->n);
  part.release();
  if (rc) return rc;
  group_sums(d, h, out);
  return LAMB_HIP_OK;
}

int32_t group_diff_sumsq(lamb_hip_mat* X, lamb_hip_mat* Y, double* out) {
  lamb_hip_dist* d = X->dist;
  lamb_hip_ctx* ctx = d->ctx;
  if (values_missing(X) || (Y && values_missing(Y))) return refuse_missing_values(ctx, "norm");
  DevBuf part;
  HIPCHK(ctx, part.ensure(sizeof(double) * std::max(d->n, 1)), "norm partials");
  if (d->n) {
    if (Y)
      hipLaunchKernelGGL(row_diff_sumsq_kernel, dim3(grid_waves(d->n)), dim3(256), 0, ctx->stream,
                         d->n, (const int64_t*)X->d_row_p.p, (const int32_t*)X->d_col.p,
                         (const int64_t*)X->d_off.p, (const double*)X->d_data.p,
                         (const int64_t*)Y->d_row_p.p, (const int32_t*)Y->d_col.p,
                         (const int64_t*)Y->d_off.p, (const double*)Y->d_data.p,
                         (const int32_t*)d->d_blk.p, (double*)part.p);
    else
      hipLaunchKernelGGL(row_sumsq_kernel, dim3(grid_waves(d->n)), dim3(256), 0, ctx->stream, d->n,
                         (const int64_t*)X->d_row_p.p, (const int32_t*)X->d_col.p,
                         (const int64_t*)X->d_off.p, (const int32_t*)d->d_blk.p,
                         (const double*)X->d_data.p, (double*)part.p);
    HIPCHK(ctx, hipGetLastError(), "norm kernel");
  }
  std::vector<double> h;
  int32_t rc = atom_partials(ctx, h, part, d->n);
  part.release();
  if (rc) return rc;
  group_sums(d, h, out);
  return LAMB_HIP_OK;
}

}  // namespace

extern "C" int32_t lamb_hip_dist_set_groups(lamb_hip_dist* d, int32_t n_groups,
                                            const int32_t* group_of_atom) {
  if (!d) return LAMB_HIP_FAILURE;
  lamb_hip_ctx* ctx = d->ctx;
  if (n_groups < 1 || (d->n > 0 && !group_of_atom))
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS, "set_groups: bad arguments");
  std::vector<int32_t> g(d->n);
  for (int32_t a = 0; a < d->n; ++a) {
    const int32_t v = group_of_atom[a];
    if (v < 0 || v >= n_groups)
      return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                       "set_groups: atom %d has group %d outside 0..%d", a, v, n_groups - 1);
    g[d->perm[a]] = v;
  }
  (void)hipSetDevice(ctx->device);
  d->group_int = g;
  d->n_groups = n_groups;
  HIPCHK(ctx, d->d_group.upload(ctx->stream, d->group_int.data(), sizeof(int32_t) * d->n),
         "groups upload");
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "groups sync");
  return LAMB_HIP_OK;
}

extern "C" int32_t lamb_hip_group_trace(const lamb_hip_mat* A, double* out) {
  if (!A || !out) return LAMB_HIP_FAILURE;
  (void)hipSetDevice(A->dist->ctx->device);
  return group_trace(const_cast<lamb_hip_mat*>(A), out);
}

extern "C" int32_t lamb_hip_group_norm(const lamb_hip_mat* A, double* out) {
  if (!A || !out) return LAMB_HIP_FAILURE;
  (void)hipSetDevice(A->dist->ctx->device);
  int32_t rc = group_diff_sumsq(const_cast<lamb_hip_mat*>(A), nullptr, out);
  if (rc) return rc;
  for (int32_t g = 0; g < A->dist->n_groups; ++g) out[g] = std::sqrt(out[g]);
  return LAMB_HIP_OK;
}

extern "C" int32_t lamb_hip_group_diff_norm(const lamb_hip_mat* X, const lamb_hip_mat* Y,
                                            double* out) {
  if (!X || !Y || !out) return LAMB_HIP_FAILURE;
  if (X->dist != Y->dist)
    return set_error(X->dist->ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "diff_norm: X and Y must share one distribution");
  (void)hipSetDevice(X->dist->ctx->device);
  int32_t rc = group_diff_sumsq(const_cast<lamb_hip_mat*>(X), const_cast<lamb_hip_mat*>(Y), out);
  if (rc) return rc;
  for (int32_t g = 0; g < X->dist->n_groups; ++g) out[g] = std::sqrt(out[g]);
  return LAMB_HIP_OK;
}

extern "C" int32_t lamb_hip_group_gershgorin(const lamb_hip_mat* A, double* lo, double* hi) {
  if (!A || !lo || !hi) return LAMB_HIP_FAILURE;
  lamb_hip_dist* d = A->dist;
  lamb_hip_ctx* ctx = d->ctx;
  if (values_missing(A)) return refuse_missing_values(ctx, "gershgorin");
  (void)hipSetDevice(ctx->device);
  DevBuf plo, phi, miss;
  HIPCHK(ctx, plo.ensure(sizeof(double) * std::max(d->n, 1)), "gershgorin partials");
  HIPCHK(ctx, phi.ensure(sizeof(double) * std::max(d->n, 1)), "gershgorin partials");
  HIPCHK(ctx, miss.ensure(sizeof(int64_t)), "gershgorin flag");
  hipLaunchKernelGGL(fill_i64, dim3(1), dim3(64), 0, ctx->stream, (int64_t*)miss.p, (int64_t)1,
                     (int64_t)0);
  if (d->n)
    hipLaunchKernelGGL(row_gershgorin_kernel, dim3(grid_threads(d->n)), dim3(256), 0, ctx->stream,
                       d->n, (const int64_t*)A->d_row_p.p, (const int32_t*)A->d_col.p,
                       (const int64_t*)A->d_off.p, (const int32_t*)d->d_blk.p,
                       (const double*)A->d_data.p, (double*)plo.p, (double*)phi.p,
                       (unsigned long long*)miss.p);
  HIPCHK(ctx, hipGetLastError(), "gershgorin kernel");
  std::vector<double> l, h;
  int64_t missing = 0;
  int32_t rc = atom_partials(ctx, l, plo, d->n);
  if (!rc) rc = atom_partials(ctx, h, phi, d->n);
  if (!rc && hipMemcpy(&missing, miss.p, sizeof(missing), hipMemcpyDeviceToHost) != hipSuccess)
    rc = set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_EXTERNAL_ERROR, "gershgorin: flag download");
  plo.release();
  phi.release();
  miss.release();
  if (rc) return rc;
  for (int32_t g = 0; g < d->n_groups; ++g) {
    lo[g] = INFINITY;
    hi[g] = -INFINITY;
  }
  for (int32_t p = 0; p < d->n; ++p) {
    if (d->blk_int[p] == 0) continue;
    const int32_t g = d->group_int[p];
    lo[g] = std::min(lo[g], l[p]);
    hi[g] = std::max(hi[g], h[p]);
  }
  if (missing)
    return set_error(ctx, LAMB_HIP_WARNING, LAMB_HIP_PRECONDITION_FAILED,
                     "gershgorin: %lld atoms have no stored diagonal block", (long long)missing);
  return LAMB_HIP_OK;
}

extern "C" int32_t lamb_hip_group_scale_shift(lamb_hip_mat* A, const double* a, const double* b) {
  if (!A || !a || !b) return LAMB_HIP_FAILURE;
  lamb_hip_dist* d = A->dist;
  lamb_hip_ctx* ctx = d->ctx;
  if (values_missing(A)) return refuse_missing_values(ctx, "scale_shift");
  (void)hipSetDevice(ctx->device);
  // the shift needs every diagonal block of an atom with rows
  bool has_shift = false;
  for (int32_t g = 0; g < d->n_groups; ++g) has_shift |= b[g] != 0.0;
  if (has_shift) {
    DevBuf plo, phi, miss;
    HIPCHK(ctx, plo.ensure(sizeof(double) * std::max(d->n, 1)), "diag check");
    HIPCHK(ctx, phi.ensure(sizeof(double) * std::max(d->n, 1)), "diag check");
    HIPCHK(ctx, miss.ensure(sizeof(int64_t)), "diag check");
    hipLaunchKernelGGL(fill_i64, dim3(1), dim3(64), 0, ctx->stream, (int64_t*)miss.p,
                       (int64_t)1, (int64_t)0);
    if (d->n)
      hipLaunchKernelGGL(row_gershgorin_kernel, dim3(grid_threads(d->n)), dim3(256), 0,
                         ctx->stream, d->n, (const int64_t*)A->d_row_p.p,
                         (const int32_t*)A->d_col.p, (const int64_t*)A->d_off.p,
                         (const int32_t*)d->d_blk.p, (const double*)A->d_data.p,
                         (double*)plo.p, (double*)phi.p, (unsigned long long*)miss.p);
    int64_t missing = 0;
    HIPCHK(ctx, hipGetLastError(), "diag check kernel");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "diag check sync");
    HIPCHK(ctx, hipMemcpy(&missing, miss.p, sizeof(missing), hipMemcpyDeviceToHost), "flag");
    plo.release();
    phi.release();
    miss.release();
    if (missing)
      return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_PRECONDITION_FAILED,
                       "scale_shift: %lld atoms have no stored diagonal block",
                       (long long)missing);
  }
  DevBuf da, db;
  HIPCHK(ctx, da.upload(ctx->stream, a, sizeof(double) * d->n_groups), "coefficients");
  HIPCHK(ctx, db.upload(ctx->stream, b, sizeof(double) * d->n_groups), "coefficients");
  if (A->nblk)
    hipLaunchKernelGGL(scale_shift_kernel, dim3(grid_threads(A->nblk)), dim3(256), 0, ctx->stream,
                       A->nblk, (const int32_t*)A->d_rows.p, (const int32_t*)A->d_col.p,
                       (const int64_t*)A->d_off.p, (const int32_t*)d->d_blk.p,
                       (const int32_t*)d->d_group.p, (const double*)da.p, (const double*)db.p,
                       (double*)A->d_data.p);
  HIPCHK(ctx, hipGetLastError(), "scale_shift kernel");
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "scale_shift sync");
  da.release();
  db.release();
  A->invalidate_derived();
  A->host_data_valid = false;
  return LAMB_HIP_OK;
}

// P <- 3 P^2 - 2 P^3, steps times, in place. Per step and group:
// idem[k*G + g] = ||P^2 - P||_F of the step's input P, traces[k*G + g] =
// tr of its output. Each step is two filtered multiplies,
//   P2 = P * P;  T = copy(P2);  T = -2 * P2 * P + 3 * T,
// so P3 is never formed and the update is the union pattern, post-filtered.
extern "C" int32_t lamb_hip_mcweeny(lamb_hip_ctx* ctx, lamb_hip_mat* P, int32_t steps,
                                    double eps_filter, int32_t flags, double* traces,
                                    double* idem, lamb_hip_stats* stats) {
  if (!ctx || !P || steps < 0) return LAMB_HIP_FAILURE;
  lamb_hip_dist* d = P->dist;
  const int32_t G = d->n_groups;
  auto t0 = std::chrono::steady_clock::now();
  lamb_hip_mat P2, T;
  P2.dist = d;
  T.dist = d;
  lamb_hip_stats acc;
  memset(&acc, 0, sizeof(acc));
  int32_t rc = LAMB_HIP_OK;
  auto add = [&](const lamb_hip_stats& s) {
    acc.n_triples_kept += s.n_triples_kept;
    acc.n_triples_filtered += s.n_triples_filtered;
    acc.n_c_blocks_struct += s.n_c_blocks_struct;
    acc.n_tiles += s.n_tiles;
    acc.n_tile_entries += s.n_tile_entries;
    acc.flops += s.flops;
    acc.ms_symbolic += s.ms_symbolic;
    acc.ms_numeric += s.ms_numeric;
    acc.ms_index += s.ms_index;
  };
  std::vector<double> tmp(G);
  for (int32_t k = 0; k < steps && rc == LAMB_HIP_OK; ++k) {
    lamb_hip_stats s;
    rc = lamb_hip_multiply(ctx, P, P, &P2, 1.0, 0.0, eps_filter, flags, &s);
    if (rc) break;
    add(s);
    if (idem) {
      rc = lamb_hip_group_diff_norm(&P2, P, tmp.data());
      if (rc) break;
      std::copy(tmp.begin(), tmp.end(), idem + (int64_t)k * G);
    }
    rc = lamb_hip_matrix_copy(&P2, &T);
    if (rc) break;
    rc = lamb_hip_multiply(ctx, &P2, P, &T, -2.0, 3.0, eps_filter, flags, &s);
    if (rc) break;
    add(s);
    std::swap(*P, T);
    P->invalidate_derived();
    P->host_data_valid = false;
    if (traces) {
      rc = group_trace(P, tmp.data());
      if (rc) break;
      std::copy(tmp.begin(), tmp.end(), traces + (int64_t)k * G);
    }
  }
  P2.release_device();
  T.release_device();
  if (stats) {
    acc.n_c_blocks_kept = P->nblk;
    acc.n_c_elems_kept = P->nelems;
    acc.ms_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *stats = acc;
  }
  return rc;
}
