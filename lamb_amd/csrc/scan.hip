// Exclusive prefix sums of int64 arrays on the device (three-phase:
// per-block scan, scan of the block totals, add back).
#include "common.h"
#include "device_util.h"

namespace lamb {

LAMB_GUARD_UNIT(scan)


__global__ void fill_i64(int64_t* __restrict__ p, int64_t n, int64_t v) {
  LAMB_GUARD;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void copy_i64(int64_t* __restrict__ dst, const int64_t* __restrict__ src, int64_t n) {
  LAMB_GUARD;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

namespace {
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;  // elements per thread
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ int64_t block_exclusive_scan(int64_t v, int64_t* total) {
  __shared__ int64_t warp_tot[kScanBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) warp_tot[wave] = x;
  __syncthreads();
  int64_t base = 0, all = 0;
  for (int w = 0; w < kScanBlock / 64; ++w) {
    if (w < wave) base += warp_tot[w];
    all += warp_tot[w];
  }
  __syncthreads();
  *total = all;
  return base + x - v;
}

// out[i] = sum of in[0..i) within each tile; tile totals to sums[blockIdx].
// in and out may be the same array (no __restrict__ on them): every thread
// reads its elements before the block scan's barriers and writes after.
__global__ __launch_bounds__(kScanBlock) void scan_tiles(const int64_t* in, int64_t* out,
                                                         int64_t n,
                                                         int64_t* __restrict__ sums) {
  LAMB_GUARD;
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t v[kScanItems];
  int64_t t = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = base + i < n ? in[base + i] : 0;
    t += v[i];
  }
  int64_t total;
  int64_t run = block_exclusive_scan(t, &total);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void add_tile_base(int64_t* __restrict__ out, int64_t n,
                                                            const int64_t* __restrict__ bases) {
  LAMB_GUARD;
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  const int64_t add = bases[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) out[base + i] += add;
}
// Arrays of at most kSmallScan values: one workgroup walks the tiles in
// order with a running carry and writes out[n] = total, one launch in place
// of the three phases' five. in and out may be the same array, as above.
constexpr int64_t kSmallScan = 8 * (int64_t)kScanTile;
__global__ __launch_bounds__(kScanBlock) void scan_small(const int64_t* in, int64_t* out,
                                                         int64_t n) {
  LAMB_GUARD;
  int64_t carry = 0;
  for (int64_t t0 = 0; t0 < n; t0 += kScanTile) {
    const int64_t base = t0 + (int64_t)threadIdx.x * kScanItems;
    int64_t v[kScanItems];
    int64_t t = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      v[i] = base + i < n ? in[base + i] : 0;
      t += v[i];
    }
    int64_t total;
    int64_t run = carry + block_exclusive_scan(t, &total);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      if (base + i < n) out[base + i] = run;
      run += v[i];
    }
    carry += total;
  }
  if (threadIdx.x == 0) out[n] = carry;
}
}  // namespace

// Exclusive scan of in[0..n) into out[0..n]; out[n] = total. in and out may
// alias. scratch needs scan_scratch_elems(n) int64s.
int64_t scan_scratch_elems(int64_t n) {
  int64_t tiles = (n + kScanTile - 1) / kScanTile;
  int64_t tot = 0;
  while (tiles > 1) {
    tot += tiles + 1;
    tiles = (tiles + kScanTile - 1) / kScanTile;
  }
  return tot + 2;
}

hipError_t exclusive_scan(hipStream_t s, const int64_t* in, int64_t* out, int64_t n,
                          int64_t* scratch) {
  if (n <= 0) {
    hipLaunchKernelGGL(fill_i64, dim3(1), dim3(64), 0, s, out, (int64_t)1, (int64_t)0);
    return hipGetLastError();
  }
  if (n <= kSmallScan) {
    hipLaunchKernelGGL(scan_small, dim3(1), dim3(kScanBlock), 0, s, in, out, n);
    return hipGetLastError();
  }
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  int64_t* sums = scratch;
  hipLaunchKernelGGL(scan_tiles, dim3((unsigned)tiles), dim3(kScanBlock), 0, s, in, out, n, sums);
  if (tiles > 1) {
    // scan the tile totals (recursively) in place, then add back
    hipError_t e = exclusive_scan(s, sums, sums, tiles, scratch + tiles + 1);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(add_tile_base, dim3((unsigned)tiles), dim3(kScanBlock), 0, s, out, n, sums);
    // total = sums[tiles]
    hipLaunchKernelGGL(copy_i64, dim3(1), dim3(64), 0, s, out + n, sums + tiles, (int64_t)1);
  } else {
    hipLaunchKernelGGL(copy_i64, dim3(1), dim3(64), 0, s, out + n, sums, (int64_t)1);
  }
  return hipGetLastError();
}

}  // namespace lamb
