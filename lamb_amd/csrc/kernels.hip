// HIP kernels for gfx950: block norms, FP64 numeric phase, trace.
#include "common.h"
#include "tile_common.h"

namespace lamb {

// ---------------------------------------------------------------------------
// Block norms: nsq(X) = sum of squares in column-major order, one FMA per
// step, starting from +0. One thread owns one block so the summation order
// is exactly the sequential one.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void block_nsq_kernel(int64_t nblk, const int32_t* __restrict__ rows,
                                                        const int32_t* __restrict__ cols,
                                                        const int32_t* __restrict__ blk,
                                                        const int64_t* __restrict__ off,
                                                        const double* __restrict__ data,
                                                        double* __restrict__ nsq) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  const int64_t len = (int64_t)blk[rows[b]] * blk[cols[b]];
  const double* x = data + off[b];
  double s = 0.0;
  for (int64_t e = 0; e < len; ++e) s = __builtin_fma(x[e], x[e], s);
  nsq[b] = s;
}

hipError_t launch_block_nsq(hipStream_t s, int64_t nblk, const int32_t* d_rows,
                            const int32_t* d_cols, const int32_t* d_blk, const int64_t* d_off,
                            const double* d_data, double* d_nsq) {
  if (nblk == 0) return hipSuccess;
  const int64_t grid = (nblk + 255) / 256;
  hipLaunchKernelGGL(block_nsq_kernel, dim3((unsigned)grid), dim3(256), 0, s, nblk, d_rows,
                     d_cols, d_blk, d_off, d_data, d_nsq);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Strip packing: each strip (the blocks of one band against one atom k)
// becomes a dense [kdim][64] panel, zero where a band atom has no block.
// pk_off[n_strips] (the panel total) is followed by one zero column.
// A strips (row_side): panel[kk][r] = A_{i(r),k}(r_in, kk).
// B strips:            panel[kk][c] = B_{k,j(c)}(kk, c_in).
// One wave per strip; lane = band row / column.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pack_strips_kernel(int64_t n_strips,
                                                          const Strip* __restrict__ strips,
                                                          const int64_t* __restrict__ pk_off,
                                                          const Band* __restrict__ bands,
                                                          const int32_t* __restrict__ blk,
                                                          const double* __restrict__ data,
                                                          int row_side, double* __restrict__ packed) {
  const int lane = threadIdx.x & 63;
  // the panel column after the last strip is all zeros (the numeric kernel
  // points masked-off operand lanes at it)
  if (blockIdx.x == 0 && threadIdx.x < 64) packed[pk_off[n_strips] + lane] = 0.0;
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
  double* out = packed + pk_off[s];
  for (int kk = 0; kk < kdim; ++kk) {
    double v = 0.0;
    if (have) v = row_side ? data[o + (int64_t)kk * ext + in] : data[o + (int64_t)in * kdim + kk];
    out[kk * 64 + lane] = v;
  }
}

hipError_t launch_pack_strips(hipStream_t s, int64_t n_strips, const Strip* strips,
                              const int64_t* pk_off, const Band* bands, const int32_t* d_blk,
                              const double* data, bool row_side, double* packed) {
  if (n_strips == 0) return hipSuccess;
  const int64_t grid = (n_strips + 3) / 4;
  hipLaunchKernelGGL(pack_strips_kernel, dim3((unsigned)grid), dim3(256), 0, s, n_strips, strips,
                     pk_off, bands, d_blk, data, row_side ? 1 : 0, packed);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP64 numeric phase.
//
// One workgroup (4 waves) owns one C tile = row band I x col band J, at most
// 64 x 64 values; wave (wr, wc) owns the 32 x 32 quadrant rows 32wr..,
// cols 32wc.. as four 16 x 16 v_mfma_f64_16x16x4 accumulators.
//
// The tile's entries (strip pairs, one K atom each) are staged in LDS as a
// table of K columns (A panel column, B panel column, row / column atom
// masks); four consecutive columns form a K step. Every wave then walks, on
// its own and without barriers, the steps that touch its quadrant: 64 steps
// at a time, one lane per step computes the step's 4-bit mask of live 16 x 16
// blocks, a ballot collects the live steps, and the wave takes them in order
// (s_ff1 + readlane). MFMA operands are loaded straight from the packed
// [k][64] panels into a ring of register slots, kPrefetch steps ahead of the
// MFMAs that consume them. Every step issues exactly four loads: a lane whose
// operand is zero (atom outside the entry's masks, block not live, step past
// the end) loads from the zero column after the panels instead, so the load
// count is static and the MFMAs wait only for their own slot.
//
// v_mfma_f64_16x16x4 lane layout (lane l): A[i][k] at l = i + 16k,
// B[k][j] at l = j + 16k, D[i][j] in lane 16*(i%4) + j, register i/4.
// ---------------------------------------------------------------------------
typedef double dbl4 __attribute__((ext_vector_type(4)));

// K columns of the per-tile step table held in LDS at a time
constexpr int kTableCols = 2048;
#ifndef LAMB_NUMERIC_PREFETCH
#define LAMB_NUMERIC_PREFETCH 4
#endif
#ifndef LAMB_NUMERIC_WAVES
#define LAMB_NUMERIC_WAVES 4
#endif
constexpr int kPrefetch = LAMB_NUMERIC_PREFETCH;  // K steps in flight per wave

__global__ __launch_bounds__(256, LAMB_NUMERIC_WAVES) void numeric_f64_kernel(
    const TileDesc* __restrict__ tiles, const int32_t* __restrict__ perm,
    const Entry* __restrict__ entries, const Band* __restrict__ bands,
    const double* __restrict__ a_pack, const double* __restrict__ b_pack, uint32_t a_zero,
    uint32_t b_zero, double alpha, double beta, double eps2,
    const int64_t* __restrict__ old_row_p, const int32_t* __restrict__ old_col,
    const int64_t* __restrict__ old_off, const double* __restrict__ old_data,
    double* __restrict__ c_data, uint16_t* __restrict__ kept_out) {
  // main loop: the step table of a window of K columns; epilogue: C tile
  __shared__ double lds[kBandRows * kPitch];
  uint2* col_ab = reinterpret_cast<uint2*>(lds);  // A, B panel column
  uint32_t* col_m = reinterpret_cast<uint32_t*>(col_ab + kTableCols);  // rmask | cmask << 16
  uint32_t* step_g = col_m + kTableCols;  // row groups | col groups << 16
  static_assert(sizeof(uint32_t) * (3 * kTableCols + kTableCols / 4) <=
                    sizeof(double) * kBandRows * kPitch, "step table exceeds LDS buffer");
  __shared__ TileSmem sm;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wr = wave >> 1, wc = wave & 1;
  const TileInfo T = tile_setup(tiles, perm, bands, sm);
  const int e0 = T.e0, ne = T.ne, ktot = T.ktot;
  uint32_t rbit[2], cbit[2];
  lane_atom_bits(T, sm, wr, wc, lane, rbit, cbit);
  const int kc = lane >> 4;  // K column of this lane within a step
  // the 16-bit row / column group masks (4 rows or columns per group) of
  // this wave's halves and quads: rows 32wr+16h.., cols 32wc+16q..
  const uint32_t gsh_r = 8 * wr, gsh_c = 8 * wc;
  const double* a_lane = a_pack + 32 * wr + (lane & 15);
  const double* b_lane = b_pack + 32 * wc + (lane & 15);

  dbl4 acc[4];
#pragma unroll
  for (int x = 0; x < 4; ++x) acc[x] = dbl4{0.0, 0.0, 0.0, 0.0};

  for (int Q0 = 0; Q0 < ktot; Q0 += kTableCols) {
    const int ncols = min(kTableCols, ktot - Q0);
    const int nsteps = (ncols + 3) >> 2;
    __syncthreads();  // previous window done
    for (int t = tid; t < 4 * nsteps; t += 256) {
      col_ab[t] = make_uint2(a_zero, b_zero);
      col_m[t] = 0u;
    }
    for (int t = tid; t < nsteps; t += 256) step_g[t] = 0;
    __syncthreads();
    // one thread per entry writes its columns that fall in this window
    for (int e = tid; e < ne; e += 256) {
      const Entry en = entries[e0 + e];
      const int ks = en.kstart;
      const int ke = (e + 1 < ne) ? entries[e0 + e + 1].kstart : ktot;
      const int lo = max(ks, Q0), hi = min(ke, Q0 + ncols);
      if (lo >= hi) continue;
      const uint32_t m = (uint32_t)en.rmask | ((uint32_t)en.cmask << 16);
      const uint32_t g = entry_groups(sm, en.rmask, en.cmask);
      for (int Q = lo; Q < hi; ++Q) {
        const int q = Q - Q0;
        col_ab[q] = make_uint2(en.a_col + (uint32_t)(Q - ks), en.b_col + (uint32_t)(Q - ks));
        col_m[q] = m;
      }
      for (int st = (lo - Q0) >> 2; st <= (hi - 1 - Q0) >> 2; ++st) atomicOr(&step_g[st], g);
    }
    __syncthreads();

    // ---- per-wave walk over the window's live steps, kPrefetch ahead ----
    int cbase = -64;        // first step of the current chunk
    uint64_t live = 0;      // live steps of the chunk not yet taken
    uint32_t chunk_bm = 0;  // lane l: block mask of step cbase + l
    double ra[8][2], rb[8][2];  // slots 0..kPrefetch-1 used
    uint32_t blkmask[8];
#define LAMB_FETCH(S)                                                              \
    {                                                                              \
      while (live == 0 && cbase + 64 < nsteps) {                                   \
        cbase += 64;                                                               \
        const int sl = cbase + lane;                                               \
        chunk_bm = quadrant_blocks(sl < nsteps ? step_g[sl] : 0u, gsh_r, gsh_c);   \
        live = __ballot(chunk_bm != 0u);                                           \
      }                                                                            \
      uint32_t bm = 0;                                                             \
      int st = 0;                                                                  \
      if (live) {                                                                  \
        const int l = __builtin_ctzll(live);                                       \
        live &= live - 1;                                                          \
        st = cbase + l;                                                            \
        bm = __builtin_amdgcn_readlane(chunk_bm, l);                               \
      }                                                                            \
      blkmask[S] = bm;                                                             \
      const int Q = 4 * st + kc;                                                   \
      const uint2 ab = col_ab[Q];                                                  \
      const uint32_t cm = col_m[Q];                                                \
      _Pragma("unroll") for (int h = 0; h < 2; ++h) {                              \
        const uint32_t need = (bm & (3u << (2 * h))) ? rbit[h] : 0u;               \
        const uint32_t c = (cm & need) ? ab.x : a_zero;                            \
        ra[S][h] = a_lane[(int64_t)c * 64 + 16 * h];                               \
      }                                                                            \
      _Pragma("unroll") for (int q = 0; q < 2; ++q) {                              \
        const uint32_t need = (bm & (5u << q)) ? cbit[q] : 0u;                     \
        const uint32_t c = (cm & need) ? ab.y : b_zero;                            \
        rb[S][q] = b_lane[(int64_t)c * 64 + 16 * q];                               \
      }                                                                            \
    }
#define LAMB_MFMA(S)                                                               \
    {                                                                              \
      const uint32_t bm = blkmask[S];                                              \
      _Pragma("unroll") for (int h = 0; h < 2; ++h)                                \
        _Pragma("unroll") for (int q = 0; q < 2; ++q)                              \
          if ((bm >> (2 * h + q)) & 1u)                                            \
            acc[2 * h + q] = __builtin_amdgcn_mfma_f64_16x16x4f64(ra[S][h], rb[S][q], acc[2 * h + q], 0, 0, 0); \
    }
    // slots spelled out so that every register array index is a constant
#define LAMB_SLOTS(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7)
    static_assert(kPrefetch >= 1 && kPrefetch <= 8, "prefetch depth 1..8");
#define LAMB_PRIME(S) if (S < kPrefetch) LAMB_FETCH(S)
#define LAMB_ANY(S) if (S < kPrefetch) any |= blkmask[S];
#define LAMB_STEP(S) if (S < kPrefetch) { LAMB_MFMA(S); LAMB_FETCH(S); }
    LAMB_SLOTS(LAMB_PRIME)
    for (;;) {
      uint32_t any = 0;
      LAMB_SLOTS(LAMB_ANY)
      if (!any) break;
      LAMB_SLOTS(LAMB_STEP)
    }
#undef LAMB_STEP
#undef LAMB_ANY
#undef LAMB_PRIME
#undef LAMB_SLOTS
#undef LAMB_FETCH
#undef LAMB_MFMA
  }
  __syncthreads();

  // ---- epilogue: C tile to LDS ----
  // v_mfma_f64_16x16x4 result: D[i][j] in lane 16*(i%4) + j, register i/4.
  double* Ct = lds;
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 32 * wr + 16 * h + 4 * r + (lane >> 4);
        const int col = 32 * wc + 16 * q + (lane & 15);
        Ct[row * kPitch + col] = acc[2 * h + q][r];
      }
  __syncthreads();
  tile_epilogue(Ct, T, sm, alpha, beta, eps2, old_row_p, old_col, old_off, old_data, c_data,
                kept_out);
}

hipError_t launch_numeric_f64(hipStream_t s, int32_t n_tiles, const TileDesc* tiles,
                              const int32_t* perm, const Entry* entries, const Band* bands,
                              const double* a_pack, const double* b_pack, uint32_t a_zero,
                              uint32_t b_zero, double alpha,
                              double beta, double eps2, const int64_t* old_row_p,
                              const int32_t* old_col, const int64_t* old_off,
                              const double* old_data, double* c_data, uint16_t* kept) {
  if (n_tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(numeric_f64_kernel, dim3((unsigned)n_tiles), dim3(256), 0, s, tiles, perm,
                     entries, bands, a_pack, b_pack, a_zero, b_zero, alpha, beta, eps2, old_row_p, old_col,
                     old_off, old_data, c_data, kept);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Trace: per-atom partial sums of the diagonal of diagonal blocks.
// ---------------------------------------------------------------------------
__global__ void trace_kernel(int32_t n_atoms, const int64_t* __restrict__ row_p,
                             const int32_t* __restrict__ col, const int64_t* __restrict__ off,
                             const int32_t* __restrict__ blk, const double* __restrict__ data,
                             double* __restrict__ out) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_atoms) return;
  double t = 0.0;
  int64_t lo = row_p[i], hi = row_p[i + 1] - 1;
  while (lo <= hi) {
    const int64_t mid = (lo + hi) >> 1;
    const int32_t c = col[mid];
    if (c == i) {
      const int m = blk[i];
      const double* x = data + off[mid];
      for (int r = 0; r < m; ++r) t += x[r + r * m];
      break;
    }
    if (c < i) lo = mid + 1;
    else hi = mid - 1;
  }
  out[i] = t;
}

hipError_t launch_trace(hipStream_t s, int32_t n_atoms, const int64_t* row_p,
                        const int32_t* col, const int64_t* off, const int32_t* blk,
                        const double* data, double* out) {
  if (n_atoms == 0) return hipSuccess;
  hipLaunchKernelGGL(trace_kernel, dim3((n_atoms + 255) / 256), dim3(256), 0, s, n_atoms,
                     row_p, col, off, blk, data, out);
  return hipGetLastError();
}

}  // namespace lamb
