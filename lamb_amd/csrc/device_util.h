// Small device helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lamb {

__global__ void fill_i64(int64_t* __restrict__ p, int64_t n, int64_t v);
__global__ void copy_i64(int64_t* __restrict__ dst, const int64_t* __restrict__ src, int64_t n);

// wave-wide helpers (wave64)
__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// exclusive prefix within the wave
__device__ __forceinline__ int64_t wave_excl_i64(int64_t v) {
  const int lane = threadIdx.x & 63;
  int64_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x - v;
}

// exclusive scan of in[0..n) into out[0..n], out[n] = total (scan.hip)
int64_t scan_scratch_elems(int64_t n);
hipError_t exclusive_scan(hipStream_t s, const int64_t* in, int64_t* out, int64_t n,
                          int64_t* scratch);

}  // namespace lamb
