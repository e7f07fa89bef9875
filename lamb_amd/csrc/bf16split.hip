// BF16-split numeric phase for gfx950: exponent-aligned BF16 slices and
// v_mfma_f32_16x16x32_bf16, FP64 across accumulation chains.
//
// Method (error-free splitting of Ozaki, Ogita, Oishi and Rump, "Error-free
// transformations of matrix multiplication by using fast routines of matrix
// multiplication and its applications", Numer. Algorithms 59, 2012), with
// one exponent per matrix row of A and per matrix column of B:
//
//   E_i = exponent with |A_ik| < 2^E_i for every element of row i,
//   unit_t = 2^(E_i - 8t), t = 1..S,
//   m_1 = rint(x / unit_1), r_1 = x - m_1 unit_1, m_t = rint(r_{t-1} / unit_t),
//   A_ik = sum_t m_t unit_t + r_S,  |m_1| <= 256, |m_t| <= 128 (t >= 2),
//
// so every slice m_t unit_t is exact in BF16 (8 significant bits) and the
// split is exact (FP64 remainders). B is split the same way per column with
// exponents F_j. A product of slices t and u is an integer below 2^16 times
// 2^(E_i + F_j - 8(t+u)); within one level L = t+u all products of an (i,j)
// share that unit, so their FP32 sum is exact while it stays below 2^24
// units: at most 256 K columns per accumulation chain (the two pairs of
// level 3, each below 2^15 units, also reach 2^24 at 256 columns; the three
// pairs of level 4 stay below 1.25 * 2^23). Levels 2, 3 and 4 get their own
// exact FP32 accumulators; levels 5 .. S+1 share one, whose FP32 rounding
// is 2^-24 of a running sum of terms 2^-25 below the leading level. (The
// rounding of a shared accumulator is relative to the sum of |terms|, which
// cancellation can make ~sqrt(K) times larger than the result: a shared
// accumulator from level 4 on measured 1.7e-10 against the FP64 product.)
// Only pairs with t + u <= S + 1 are formed: S(S+1)/2 MFMA products. After
// at most 256 K columns every accumulator is added into FP64 and cleared,
// so the cross-chain (cross-k) accumulation into C is FP64. Norms, filters
// and the epilogue are FP64 and shared with the FP64 kernel.
#include <algorithm>

#include "common.h"
#include "tile_common.h"

namespace lamb {

LAMB_GUARD_UNIT(bf16split)


typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int32_t kNoExp = -30000;       // row / column without a nonzero
constexpr int kChunk = 8;                // K columns per slice chunk
constexpr int kTableChunks = 2048;       // chunks of the LDS table per window
constexpr int kChainSteps = 8;           // live K steps (32 columns) per chain

__device__ __forceinline__ int32_t exp_above(double x) {  // |x| < 2^e
  if (x == 0.0) return kNoExp;
  int e;
  (void)frexp(x, &e);
  return e;
}

__global__ void fill_exp_kernel(int32_t* __restrict__ ex, int64_t n) {
  LAMB_GUARD;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ex[i] = kNoExp;
}

// Sixteen lanes per stored block, one per row (row_side: A, rows of the
// matrix) or column (B) of it (and rows / columns l + 16, l + 32, ... of
// larger blocks): the largest exponent of that line into ex by atomicMax.
// For A the sixteen lanes read consecutive elements of a block column.
__global__ __launch_bounds__(256) void slice_exp_kernel(
    int64_t nblk, const int32_t* __restrict__ rows, const int32_t* __restrict__ cols,
    const int32_t* __restrict__ blk, const int64_t* __restrict__ elem0,
    const int64_t* __restrict__ off, const double* __restrict__ data, int row_side,
    int32_t* __restrict__ ex) {
  LAMB_GUARD;
  const int64_t b = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
  const int l = threadIdx.x & 15;
  if (b >= nblk) return;
  const int32_t ra = rows[b], ca = cols[b];
  const int m = blk[ra], n = blk[ca];
  const double* x = data + off[b];
  if (row_side) {
    for (int i = l; i < m; i += 16) {
      int32_t e = kNoExp;
      for (int j = 0; j < n; ++j) e = max(e, exp_above(x[(int64_t)j * m + i]));
      if (e != kNoExp) atomicMax(&ex[elem0[ra] + i], e);
    }
  } else {
    for (int j = l; j < n; j += 16) {
      int32_t e = kNoExp;
      for (int i = 0; i < m; ++i) e = max(e, exp_above(x[(int64_t)j * m + i]));
      if (e != kNoExp) atomicMax(&ex[elem0[ca] + j], e);
    }
  }
}

// BF16 bits of a double that is exactly representable in BF16
__device__ __forceinline__ uint32_t bf16_bits(double v) {
  return __float_as_uint((float)v) >> 16;
}

// Slice packing: strip s (panel offset pk_off[s] doubles, a multiple of 512)
// owns chunks pk_off[s]/512 .. + ceil(kdim/8); slice t of chunk c is the
// 64 x 8 BF16 array out[(t * stride + c) * 64 + lane] (16 bytes per lane:
// the 8 K values of band row / column `lane`). Chunk n_chunks of every slice
// is zero. One wave per strip.
template <int S>
__global__ __launch_bounds__(256) void pack_slices_kernel(
    int64_t n_strips, const Strip* __restrict__ strips, const int64_t* __restrict__ pk_off,
    const Band* __restrict__ bands, const int32_t* __restrict__ blk,
    const int64_t* __restrict__ elem0, const int32_t* __restrict__ ex,
    const double* __restrict__ data, int row_side, int64_t n_chunks, uint4* __restrict__ out) {
  LAMB_GUARD;
  const int lane = threadIdx.x & 63;
  const int64_t stride = n_chunks + 1;
  if (blockIdx.x == 0 && threadIdx.x < 64) {
#pragma unroll
    for (int t = 0; t < S; ++t) out[((int64_t)t * stride + n_chunks) * 64 + lane] = make_uint4(0, 0, 0, 0);
  }
  const int64_t s = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (s >= n_strips) return;
  const int32_t band = strips[s].band;
  const int32_t k = strips[s].k;
  const uint32_t mask = strips[s].mask;
  const int kdim = blk[k];
  const int natoms = bands[band].natoms;
  int pos = -1, lo = 0, hi = 0;
  for (int a = 0; a < natoms; ++a) {
    const int r0 = bands[band].rs[a], r1 = bands[band].rs[a + 1];
    if (lane >= r0 && lane < r1) {
      pos = a;
      lo = r0;
      hi = r1;
    }
  }
  const bool have = pos >= 0 && ((mask >> pos) & 1u);
  const int64_t o = have ? strips[s].off[pos] : 0;
  const int in = lane - lo, ext = hi - lo;
  const int32_t E = pos >= 0 ? ex[elem0[bands[band].first + pos] + in] : kNoExp;
  const int64_t c0 = pk_off[s] / (64 * kChunk);
  const int nch = (kdim + kChunk - 1) / kChunk;
  for (int c = 0; c < nch; ++c) {
    double r[kChunk];
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
      const int kk = kChunk * c + j;
      double v = 0.0;
      if (have && kk < kdim)
        v = row_side ? data[o + (int64_t)kk * ext + in] : data[o + (int64_t)in * kdim + kk];
      r[j] = v;
    }
#pragma unroll
    for (int t = 0; t < S; ++t) {
      const int ue = E - 8 * (t + 1);  // slice t+1 is an integer times 2^ue
      const bool live = E != kNoExp && ue >= -120;
      const double unit = live ? ldexp(1.0, ue) : 0.0;
      const double inv = live ? ldexp(1.0, -ue) : 0.0;
      uint32_t w[kChunk / 2];
#pragma unroll
      for (int p = 0; p < kChunk / 2; ++p) {
        uint32_t h2[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int j = 2 * p + q;
          const double sl = live ? rint(r[j] * inv) * unit : 0.0;
          r[j] -= sl;
          h2[q] = bf16_bits(sl);
        }
        w[p] = h2[0] | (h2[1] << 16);
      }
      out[((int64_t)t * stride + c0 + c) * 64 + lane] = make_uint4(w[0], w[1], w[2], w[3]);
    }
  }
}

// ---------------------------------------------------------------------------
// Tile kernel. Same tile / wave geometry as numeric_f64_kernel (4 waves,
// 32 x 32 quadrant each, four 16 x 16 blocks). The LDS table holds chunks of
// 8 K columns (A chunk, B chunk, atom masks); a K step is 4 chunks, one per
// lane group (lane >> 4), the K = 32 of one v_mfma_f32_16x16x32_bf16. Each
// entry contributes ceil(kdim / 8) chunks, the last zero-padded. Waves walk
// their live steps without barriers (ballot over 64 steps, s_ff1 +
// readlane); masked lanes load the zero chunk so every step issues a fixed
// number of loads.
// ---------------------------------------------------------------------------
template <int S, bool kLoop>
__global__ __launch_bounds__(256, 2) void numeric_bf16_kernel(
    int32_t n_tiles, int32_t* __restrict__ slot_counter, const int32_t* __restrict__ xrange,
    const TileDesc* __restrict__ tiles,
    const int32_t* __restrict__ perm,
    const Entry* __restrict__ entries, const Band* __restrict__ bands,
    const uint4* __restrict__ a_sl, const uint4* __restrict__ b_sl, uint32_t a_zero,
    uint32_t b_zero, double alpha, double beta, double eps2,
    const int64_t* __restrict__ old_row_p, const int32_t* __restrict__ old_col,
    const int64_t* __restrict__ old_off, const double* __restrict__ old_data,
    double* __restrict__ c_data, uint16_t* __restrict__ kept_out) {
  LAMB_GUARD;
  // Without kLoop the body runs once, on slot blockIdx.x, as a plain one-tile
  // kernel with that kernel's registers. With kLoop (a persistent grid) a
  // workgroup of residue x = blockIdx.x mod 8 runs its turns j = blockIdx.x
  // / 8, then the values of counter x, one claimed during each tile's
  // epilogue, on the slots persistent_slot gives: interleaved (each XCD walks
  // the tiles it walks with one workgroup per tile) or the XCD's own range.
  __shared__ int32_t next_slot;
  int32_t slot = (int32_t)blockIdx.x;
  if constexpr (kLoop) {
    slot = persistent_slot(xrange, n_tiles, (int32_t)(blockIdx.x >> 3));
    if (slot < 0) return;
  }
  for (;;) {
    __shared__ double lds[kBandRows * kPitch];
    uint2* tbl_ab = reinterpret_cast<uint2*>(lds);  // A, B chunk
    uint32_t* tbl_m = reinterpret_cast<uint32_t*>(tbl_ab + kTableChunks);  // rmask | cmask << 16
    uint32_t* step_g = tbl_m + kTableChunks;  // row groups | col groups << 16
    static_assert(sizeof(uint32_t) * (3 * kTableChunks + kTableChunks / 4) <=
                      sizeof(double) * kBandRows * kPitch, "chunk table exceeds LDS buffer");
    __shared__ TileSmem sm;
    __shared__ int32_t scan_w[4];
    // FP64 sums of the finished chains, one element per lane of one wave (the
    // wave quadrants are disjoint), read by the epilogue
    __shared__ double ctile[kBandRows * kPitch];

    const int tid = tile_tid<kLoop>();
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = wave >> 1, wc = wave & 1;
    const TileInfo T = tile_setup(tiles, perm, bands, sm, slot, tid);
    uint32_t rbit[2], cbit[2];
    lane_atom_bits(T, sm, wr, wc, lane, rbit, cbit);
    const uint32_t gsh_r = 8 * wr, gsh_c = 8 * wc;
    const int64_t a_stride = (int64_t)a_zero + 1, b_stride = (int64_t)b_zero + 1;
    const uint4* a_lane = a_sl + 32 * wr + (lane & 15);
    const uint4* b_lane = b_sl + 32 * wc + (lane & 15);

    // FP32 level accumulators [level 2, 3, 4, 5+][block 2h+q] (levels past
    // S+1 unused). v_mfma_f32_16x16x32_bf16 result: D[i][j] in lane
    // 16*(i/4) + j, register i%4; flush() adds them into ctile and clears them.
    constexpr int NA = S < 4 ? S : 4;
    f32x4 acc[NA][4];
  #pragma unroll
    for (int b = 0; b < 4; ++b)
  #pragma unroll
      for (int L = 0; L < NA; ++L) acc[L][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int t = tid; t < kBandRows * kPitch; t += 256) ctile[t] = 0.0;
    __syncthreads();  // zeros visible before any wave flushes into its quadrant
    auto flush = [&]() {
  #pragma unroll
      for (int h = 0; h < 2; ++h)
  #pragma unroll
        for (int q = 0; q < 2; ++q)
  #pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = 32 * wr + 16 * h + 4 * (lane >> 4) + r;
            const int col = 32 * wc + 16 * q + (lane & 15);
            double v = ctile[row * kPitch + col];
  #pragma unroll
            for (int L = 0; L < NA; ++L) v += (double)acc[L][2 * h + q][r];
            ctile[row * kPitch + col] = v;
          }
  #pragma unroll
      for (int b = 0; b < 4; ++b)
  #pragma unroll
        for (int L = 0; L < NA; ++L) acc[L][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    };
    int chain = 0;  // live steps in the current chain (wave-uniform)

    for (int E0 = 0; E0 < T.ne; E0 += 256) {
      // ---- chunk offsets of this window's entries (block scan) ----
      const int e = E0 + tid;
      int nch = 0;
      Entry en;
      int ks = 0;
      if (e < T.ne) {
        en = entries[T.e0 + e];
        ks = en.kstart;
        const int ke = (e + 1 < T.ne) ? entries[T.e0 + e + 1].kstart : T.ktot;
        nch = (ke - ks + kChunk - 1) / kChunk;
      }
      int incl = nch;
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
      }
      __syncthreads();  // previous window's walk is done with the table
      if (lane == 63) scan_w[wave] = incl;
      __syncthreads();
      int base = 0;
      for (int w = 0; w < wave; ++w) base += scan_w[w];
      const int nchunks = scan_w[0] + scan_w[1] + scan_w[2] + scan_w[3];
      const int nsteps = (nchunks + 3) >> 2;
      for (int t = nchunks + tid; t < 4 * nsteps; t += 256) {
        tbl_ab[t] = make_uint2(a_zero, b_zero);
        tbl_m[t] = 0u;
      }
      for (int t = tid; t < nsteps; t += 256) step_g[t] = 0;
      __syncthreads();
      if (nch) {
        const int c0 = base + incl - nch;
        const uint32_t m = (uint32_t)en.rmask | ((uint32_t)en.cmask << 16);
        const uint32_t g = entry_groups(sm, en.rmask, en.cmask);
        const uint32_t ac = (en.a_col & ~kEntryPartial) / kChunk, bc = en.b_col / kChunk;
        for (int c = 0; c < nch; ++c) {
          tbl_ab[c0 + c] = make_uint2(ac + c, bc + c);
          tbl_m[c0 + c] = m;
        }
        for (int st = c0 >> 2; st <= (c0 + nch - 1) >> 2; ++st) atomicOr(&step_g[st], g);
      }
      __syncthreads();

      // ---- per-wave walk over the window's live steps ----
      int cbase = -64;
      uint64_t live = 0;
      uint32_t chunk_bm = 0;
      for (;;) {
        while (live == 0 && cbase + 64 < nsteps) {
          cbase += 64;
          const int sl = cbase + lane;
          chunk_bm = quadrant_blocks(sl < nsteps ? step_g[sl] : 0u, gsh_r, gsh_c);
          live = __ballot(chunk_bm != 0u);
        }
        if (live == 0) break;
        const int l = __builtin_ctzll(live);
        live &= live - 1;
        const int st = cbase + l;
        const uint32_t bm = __builtin_amdgcn_readlane(chunk_bm, l);
        const int Qc = 4 * st + (lane >> 4);
        const uint2 ab = tbl_ab[Qc];
        const uint32_t cm = tbl_m[Qc];
        bf16x8 ra[S][2], rb[S][2];
  #pragma unroll
        for (int h = 0; h < 2; ++h) {
          const uint32_t need = (bm & (3u << (2 * h))) ? rbit[h] : 0u;
          const int64_t c = (cm & need) ? ab.x : a_zero;
  #pragma unroll
          for (int t = 0; t < S; ++t) {
            const uint4 v = a_lane[(t * a_stride + c) * 64 + 16 * h];
            ra[t][h] = __builtin_bit_cast(bf16x8, v);
          }
        }
  #pragma unroll
        for (int q = 0; q < 2; ++q) {
          const uint32_t need = (bm & (5u << q)) ? cbit[q] : 0u;
          const int64_t c = (cm & need) ? ab.y : b_zero;
  #pragma unroll
          for (int u = 0; u < S; ++u) {
            const uint4 v = b_lane[(u * b_stride + c) * 64 + 16 * q];
            rb[u][q] = __builtin_bit_cast(bf16x8, v);
          }
        }
  #pragma unroll
        for (int h = 0; h < 2; ++h)
  #pragma unroll
          for (int q = 0; q < 2; ++q)
            if ((bm >> (2 * h + q)) & 1u) {
  #pragma unroll
              for (int t = 0; t < S; ++t)
  #pragma unroll
                for (int u = 0; u < S - t; ++u) {
                  const int L = t + u < NA - 1 ? t + u : NA - 1;
                  acc[L][2 * h + q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                      ra[t][h], rb[u][q], acc[L][2 * h + q], 0, 0, 0);
                }
            }
        if (++chain == kChainSteps) {
          chain = 0;
          flush();
        }
      }
    }
    flush();
    __syncthreads();
    tile_epilogue<256>(ctile, T, sm, alpha, beta, eps2, old_row_p, old_col, old_off, old_data,
                       c_data, kept_out, tid,
                       kLoop ? slot_counter + (blockIdx.x & 7) : nullptr, &next_slot);
    if constexpr (!kLoop) break;
    __syncthreads();  // next_slot is visible; the next set-up rewrites the LDS
    slot = persistent_slot(xrange, n_tiles, __builtin_amdgcn_readfirstlane(next_slot));
    if (slot < 0) break;
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int64_t grid1(int64_t n) { return (n + 255) / 256; }

hipError_t launch_slice_exponents(hipStream_t s, int64_t nblk, const int32_t* rows,
                                  const int32_t* cols, const int32_t* blk, const int64_t* elem0,
                                  const int64_t* off, const double* data, bool row_side,
                                  int64_t n_elem, int32_t* ex) {
  if (n_elem > 0)
    hipLaunchKernelGGL(fill_exp_kernel, dim3((unsigned)grid1(n_elem)), dim3(256), 0, s, ex, n_elem);
  if (nblk > 0)
    hipLaunchKernelGGL(slice_exp_kernel, dim3((unsigned)grid1(nblk * 16)), dim3(256), 0, s, nblk, rows,
                       cols, blk, elem0, off, data, row_side ? 1 : 0, ex);
  return hipGetLastError();
}

hipError_t launch_pack_slices(hipStream_t s, int slices, int64_t n_strips, const Strip* strips,
                              const int64_t* pk_off, const Band* bands, const int32_t* blk,
                              const int64_t* elem0, const int32_t* ex, const double* data,
                              bool row_side, int64_t n_chunks, void* out) {
  if (n_strips == 0) return hipSuccess;
  const dim3 grid((unsigned)((n_strips + 3) / 4)), block(256);
  const int rs = row_side ? 1 : 0;
  uint4* o = static_cast<uint4*>(out);
  switch (slices) {
    case 3: hipLaunchKernelGGL(pack_slices_kernel<3>, grid, block, 0, s, n_strips, strips, pk_off, bands, blk, elem0, ex, data, rs, n_chunks, o); break;
    case 4: hipLaunchKernelGGL(pack_slices_kernel<4>, grid, block, 0, s, n_strips, strips, pk_off, bands, blk, elem0, ex, data, rs, n_chunks, o); break;
    case 5: hipLaunchKernelGGL(pack_slices_kernel<5>, grid, block, 0, s, n_strips, strips, pk_off, bands, blk, elem0, ex, data, rs, n_chunks, o); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_numeric_bf16(hipStream_t s, int slices, int32_t n_tiles, const TileDesc* tiles,
                               const int32_t* perm, const Entry* entries, const Band* bands,
                               const void* a_sl, const void* b_sl, uint32_t a_zero,
                               uint32_t b_zero, double alpha, double beta, double eps2,
                               const int64_t* old_row_p, const int32_t* old_col,
                               const int64_t* old_off, const double* old_data, double* c_data,
                               uint16_t* kept, int form, int32_t* slot_counter,
                               int32_t* grid_used) {
  if (grid_used) *grid_used = 0;
  if (n_tiles == 0) return hipSuccess;
  const uint4* a = static_cast<const uint4*>(a_sl);
  const uint4* b = static_cast<const uint4*>(b_sl);
  auto go = [&](auto kernel) {
    int32_t grid = n_tiles;
    const int32_t* xrange = nullptr;
    if (form != kFormGrid) {
      const int32_t r = resident_workgroups(reinterpret_cast<const void*>(kernel));
      if (r > 0) grid = std::min(grid, r);
      if (form == kFormXcdRanges) {
        // every residue needs a workgroup, whatever the size of its range
        grid = std::max(grid, 8);
        xrange = slot_counter + kXcdRangeWords;
      }
      launch_slot_counter_reset(s, slot_counter, grid);
    }
    if (grid_used) *grid_used = grid;
    hipLaunchKernelGGL(kernel, dim3((unsigned)grid), dim3(256), 0, s, n_tiles, slot_counter, xrange,
                       tiles, perm, entries, bands, a, b, a_zero, b_zero, alpha, beta, eps2,
                       old_row_p, old_col, old_off, old_data, c_data, kept);
  };
  // a persistent grid runs the looping instantiation
  switch (slices) {
    case 3: form != kFormGrid ? go(numeric_bf16_kernel<3, true>) : go(numeric_bf16_kernel<3, false>);
      break;
    case 4: form != kFormGrid ? go(numeric_bf16_kernel<4, true>) : go(numeric_bf16_kernel<4, false>);
      break;
    case 5: form != kFormGrid ? go(numeric_bf16_kernel<5, true>) : go(numeric_bf16_kernel<5, false>);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace lamb
