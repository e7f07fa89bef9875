// Atom -> subset partition on the device: the loop of
// lamb_set_partition_update (src/lamb_set_partition_methods.F90:98-120)
// for atoms whose coordinates change (a moving-atom step re-partitions
// without a host pass over the atoms).
//
//   subset(g) = 1 + cx + cy*nx + cz*nx*ny, c_d = mod(int(x_d * mult_d), n_d)
//     (lamb_partition_which_cube1 and lamb_partition_cube_to_cube_i,
//     src/lamb_partition_methods.F90:85-107), mult_d = real(n_d) / h_dd in
//     FP32, the reference's space_k;
//   order(g)  = arrival order of g among the atoms of its subset (:104-105).
#include "common.h"
#include "internal.h"

namespace lamb {
namespace {

constexpr int kMaxSubsets = 4096;  // subset counters held in LDS

// one thread per atom. The FP32 product and the truncating conversion are
// Fortran's int(x * mult); C's % keeps the dividend's sign like Fortran's mod.
__global__ __launch_bounds__(256) void which_cube_kernel(int64_t n, const float* __restrict__ xyz,
                                                         float mx, float my, float mz, int32_t nx,
                                                         int32_t ny, int32_t nz,
                                                         int32_t* __restrict__ subset) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const int32_t cx = (int32_t)(xyz[3 * g] * mx) % nx;
  const int32_t cy = (int32_t)(xyz[3 * g + 1] * my) % ny;
  const int32_t cz = (int32_t)(xyz[3 * g + 2] * mz) % nz;
  subset[g] = 1 + cx + cy * nx + cz * nx * ny;
}

// one wave walks the atoms 64 at a time in global order; the lanes of one
// subset in a batch take consecutive ranks (ballot + popcount below lane).
// bad[0] counts atoms outside 1..n_subsets, bad[1] ranks beyond int16.
__global__ __launch_bounds__(64) void arrival_order_kernel(int64_t n, int32_t n_subsets,
                                                           const int32_t* __restrict__ subset,
                                                           int16_t* __restrict__ order,
                                                           int32_t* __restrict__ bad) {
  __shared__ int32_t count[kMaxSubsets];
  const int lane = threadIdx.x;
  for (int s = lane; s < n_subsets; s += 64) count[s] = 0;
  __syncthreads();
  const uint64_t below = (1ull << lane) - 1ull;
  int32_t n_bad = 0, n_over = 0;
  for (int64_t base = 0; base < n; base += 64) {
    const int64_t g = base + lane;
    const int32_t s = g < n ? subset[g] : 0;
    const bool ok = g < n && s >= 1 && s <= n_subsets;
    if (g < n && !ok) ++n_bad;
    uint64_t todo = __ballot(ok);
    int32_t rank = 0;
    while (todo) {
      const int leader = __builtin_ctzll(todo);
      const int32_t sl = __shfl(s, leader, 64);
      const uint64_t m = __ballot(ok && s == sl);
      const int32_t c0 = count[sl - 1];
      if (ok && s == sl) rank = c0 + __builtin_popcountll(m & below) + 1;
      __syncthreads();  // every lane has read c0 before it moves on
      if (lane == leader) count[sl - 1] = c0 + __builtin_popcountll(m);
      __syncthreads();
      todo &= ~m;
    }
    if (ok && rank > 32767) ++n_over;
    if (g < n) order[g] = (ok && rank <= 32767) ? (int16_t)rank : (int16_t)0;
  }
  if (n_bad) atomicAdd(&bad[0], n_bad);
  if (n_over) atomicAdd(&bad[1], n_over);
}

}  // namespace
}  // namespace lamb

using lamb::set_error;

extern "C" int32_t lamb_hip_partition_atoms(lamb_hip_ctx* ctx, int64_t n_atoms, const float* coords,
                                            const float* box, const int32_t* dims,
                                            int32_t* subset_map, int16_t* order_map) {
  if (!ctx) return LAMB_HIP_FAILURE;
  if (n_atoms < 0 || !box || !dims || (n_atoms > 0 && (!coords || !subset_map || !order_map)))
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS, "partition_atoms: bad arguments");
  int64_t n_subsets = 1;
  float mult[3];
  for (int d = 0; d < 3; ++d) {
    if (dims[d] < 1 || !(box[d] > 0.0f))
      return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                       "partition_atoms: dims %d and box %g along axis %d", dims[d],
                       (double)box[d], d + 1);
    n_subsets *= dims[d];
    mult[d] = (float)dims[d] / box[d];  // real(ncubes, space_k) / h(d,d)
  }
  if (n_subsets > lamb::kMaxSubsets)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_UNIMPLEMENTED,
                     "partition_atoms: %lld subsets (at most %d)", (long long)n_subsets,
                     lamb::kMaxSubsets);
  if (n_atoms == 0) return LAMB_HIP_OK;
  (void)hipSetDevice(ctx->device);
  hipStream_t s = ctx->stream;
  lamb::DevBuf d_xyz, d_subset, d_order, d_bad;
  auto done = [&](int32_t rc) {
    d_xyz.release();
    d_subset.release();
    d_order.release();
    d_bad.release();
    return rc;
  };
  hipError_t e;
  if ((e = d_xyz.ensure(sizeof(float) * 3 * (size_t)n_atoms)) != hipSuccess ||
      (e = d_subset.ensure(sizeof(int32_t) * (size_t)n_atoms)) != hipSuccess ||
      (e = d_order.ensure(sizeof(int16_t) * (size_t)n_atoms)) != hipSuccess ||
      (e = d_bad.ensure(2 * sizeof(int32_t))) != hipSuccess)
    return done(lamb::hip_error(ctx, e, "partition_atoms alloc"));
  const int32_t zero[2] = {0, 0};
  int32_t bad[2] = {0, 0};
  if ((e = hipMemcpyAsync(d_xyz.p, coords, sizeof(float) * 3 * (size_t)n_atoms,
                          hipMemcpyHostToDevice, s)) != hipSuccess ||
      (e = hipMemcpyAsync(d_bad.p, zero, sizeof zero, hipMemcpyHostToDevice, s)) != hipSuccess)
    return done(lamb::hip_error(ctx, e, "partition_atoms upload"));
  hipLaunchKernelGGL(lamb::which_cube_kernel, dim3((unsigned)((n_atoms + 255) / 256)), dim3(256),
                     0, s, n_atoms, (const float*)d_xyz.p, mult[0], mult[1], mult[2], dims[0],
                     dims[1], dims[2], (int32_t*)d_subset.p);
  hipLaunchKernelGGL(lamb::arrival_order_kernel, dim3(1), dim3(64), 0, s, n_atoms,
                     (int32_t)n_subsets, (const int32_t*)d_subset.p, (int16_t*)d_order.p,
                     (int32_t*)d_bad.p);
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = hipMemcpyAsync(subset_map, d_subset.p, sizeof(int32_t) * (size_t)n_atoms,
                          hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (e = hipMemcpyAsync(order_map, d_order.p, sizeof(int16_t) * (size_t)n_atoms,
                          hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (e = hipMemcpyAsync(bad, d_bad.p, sizeof bad, hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (e = hipStreamSynchronize(s)) != hipSuccess)
    return done(lamb::hip_error(ctx, e, "partition_atoms"));
  if (bad[0])
    return done(set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                          "partition_atoms: %d atoms fall outside the cubes (coordinates must be "
                          "wrapped into the box)",
                          bad[0]));
  if (bad[1])
    return done(set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_OVERFLOW_ERROR,
                          "partition_atoms: a subset holds more than 32767 atoms (the int16 "
                          "local index)"));
  return done(LAMB_HIP_OK);
}
