// HIP kernels for gfx950: block norms, FP64 numeric phase, trace.
#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>
#include <utility>

#include "common.h"
#include "tile_common.h"

namespace lamb {

LAMB_GUARD_UNIT(kernels)


// ---------------------------------------------------------------------------
// Block norms: nsq(X) = sum of squares in column-major order, one FMA per
// step, starting from +0: the sequential order is part of the contract. An
// 8-lane group per block loads the block through LDS in coalesced chunks of
// kNsqChunk values; lane 0 of the group then runs the chain over them, so a
// wave runs eight chains at once. (One thread streaming its own block, 64
// blocks per wave, re-fetched each line from beyond the L2 about six times:
// 3.8 GB for a 630 MB matrix, 0.75 ms; 16-lane groups with 256-value chunks,
// four chains per wave at five waves per SIMD: 0.55 ms; this form 0.26 ms,
// and 4-lane groups with 32-value chunks the same.)
// ---------------------------------------------------------------------------
constexpr int kNsqGroup = 8;   // lanes per block
constexpr int kNsqChunk = 64;  // values staged per block and pass
constexpr int kNsqBlocks = 256 / kNsqGroup;

__global__ __launch_bounds__(256) void block_nsq_kernel(int64_t nblk, const int32_t* __restrict__ rows,
                                                        const int32_t* __restrict__ cols,
                                                        const int32_t* __restrict__ blk,
                                                        const int64_t* __restrict__ off,
                                                        const double* __restrict__ data,
                                                        double* __restrict__ nsq) {
  LAMB_GUARD;
  __shared__ double buf[kNsqBlocks][kNsqChunk];
  const int grp = threadIdx.x / kNsqGroup, gl = threadIdx.x % kNsqGroup;
  const int64_t b = (int64_t)blockIdx.x * kNsqBlocks + grp;
  const bool have = b < nblk;
  const int64_t len = have ? (int64_t)blk[rows[b]] * blk[cols[b]] : 0;
  const double* x = data + (have ? off[b] : 0);
  double s = 0.0;
  for (int64_t c0 = 0; c0 < len; c0 += kNsqChunk) {
    const int n = (int)(len - c0 < kNsqChunk ? len - c0 : kNsqChunk);
    for (int e = gl; e < n; e += kNsqGroup) buf[grp][e] = x[c0 + e];
    // the group's lanes are one wave's: its LDS writes land before its reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (gl == 0)
      for (int e = 0; e < n; ++e) s = __builtin_fma(buf[grp][e], buf[grp][e], s);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (have && gl == 0) nsq[b] = s;
}

hipError_t launch_block_nsq(hipStream_t s, int64_t nblk, const int32_t* d_rows,
                            const int32_t* d_cols, const int32_t* d_blk, const int64_t* d_off,
                            const double* d_data, double* d_nsq) {
  if (nblk == 0) return hipSuccess;
  const int64_t grid = (nblk + kNsqBlocks - 1) / kNsqBlocks;
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
  LAMB_GUARD;
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
  // 16 K columns per pass (every 5- and 13-row atom in one pass): the loads
  // of a pass are independent and in flight together before the stores
  constexpr int kPass = 16;
  for (int k0 = 0; k0 < kdim; k0 += kPass) {
    double v[kPass];
#pragma unroll
    for (int j = 0; j < kPass; ++j) {
      const int kk = k0 + j;
      v[j] = 0.0;
      if (have && kk < kdim)
        v[j] = row_side ? data[o + (int64_t)kk * ext + in] : data[o + (int64_t)in * kdim + kk];
    }
#pragma unroll
    for (int j = 0; j < kPass; ++j)
      if (k0 + j < kdim) out[(k0 + j) * 64 + lane] = v[j];
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
#ifndef LAMB_F64_STEP_COLS
#define LAMB_F64_STEP_COLS 4
#endif
// K columns per step of the quadrant kernel: one or two groups of the
// 4 columns a v_mfma_f64_16x16x4 consumes
constexpr int kStepCols = LAMB_F64_STEP_COLS;
constexpr int kGroups = kStepCols / 4;

// The step table of the K window [Q0, Q0 + kTableCols) of a tile: per K
// column its A and B panel columns (the zero columns past the end) and atom
// masks, per step of KS columns the live row / column 4-groups. One thread
// per entry writes its columns; *partial becomes 1 when an entry of the
// window is partial (its atoms cut below its strips'). Returns the window's
// step count; the table and *partial are visible to all threads on return.
// tid = threadIdx.x.
template <int KS, int NCOLS = kTableCols>
__device__ __forceinline__ int build_step_table(const TileInfo& T, const TileSmem& sm,
                                                const Entry* __restrict__ entries, int Q0,
                                                uint2* col_ab, uint32_t* col_m, uint32_t* step_g,
                                                uint32_t a_zero, uint32_t b_zero, int* partial,
                                                int tid) {
  static_assert(KS == 4 || KS == 8, "K columns per step");
  const int ncols = min(NCOLS, T.ktot - Q0);
  const int nsteps = (ncols + KS - 1) / KS;
  __syncthreads();  // the previous window is done with the table
  if (tid == 0) *partial = 0;
  for (int t = tid; t < KS * nsteps; t += 256) {
    col_ab[t] = make_uint2(a_zero, b_zero);
    col_m[t] = 0u;
  }
  for (int t = tid; t < nsteps; t += 256) step_g[t] = 0;
  __syncthreads();
  for (int e = tid; e < T.ne; e += 256) {
    const Entry en = entries[T.e0 + e];
    const int ks = en.kstart;
    const int ke = (e + 1 < T.ne) ? entries[T.e0 + e + 1].kstart : T.ktot;
    const int lo = max(ks, Q0), hi = min(ke, Q0 + ncols);
    if (lo >= hi) continue;
    if (en.a_col & kEntryPartial) *partial = 1;
    const uint32_t ac = en.a_col & ~kEntryPartial;
    const uint32_t m = (uint32_t)en.rmask | ((uint32_t)en.cmask << 16);
    const uint32_t g = entry_groups(sm, en.rmask, en.cmask);
    for (int Q = lo; Q < hi; ++Q) {
      const int q = Q - Q0;
      col_ab[q] = make_uint2(ac + (uint32_t)(Q - ks), en.b_col + (uint32_t)(Q - ks));
      col_m[q] = m;
    }
    for (int st = (lo - Q0) / KS; st <= (hi - 1 - Q0) / KS; ++st) atomicOr(&step_g[st], g);
  }
  __syncthreads();
  return nsteps;
}

template <bool kLoop>
__global__ __launch_bounds__(256, LAMB_NUMERIC_WAVES) void numeric_f64_kernel(
    int32_t n_tiles, int32_t* __restrict__ slot_counter, const int32_t* __restrict__ xrange,
    const TileDesc* __restrict__ tiles,
    const int32_t* __restrict__ perm,
    const Entry* __restrict__ entries, const Band* __restrict__ bands,
    const double* __restrict__ a_pack, const double* __restrict__ b_pack, uint32_t a_zero,
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
    // main loop: the step table of a window of K columns; epilogue: C tile
    __shared__ double lds[kBandRows * kPitch];
    uint2* col_ab = reinterpret_cast<uint2*>(lds);  // A, B panel column
    uint32_t* col_m = reinterpret_cast<uint32_t*>(col_ab + kTableCols);  // rmask | cmask << 16
    uint32_t* step_g = col_m + kTableCols;  // row groups | col groups << 16
    static_assert(sizeof(uint32_t) * (3 * kTableCols + kTableCols / 4) <=
                      sizeof(double) * kBandRows * kPitch, "step table exceeds LDS buffer");
    __shared__ TileSmem sm;

    const int tid = tile_tid<kLoop>();
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = wave >> 1, wc = wave & 1;
    const TileInfo T = tile_setup(tiles, perm, bands, sm, slot, tid);
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

    __shared__ int win_partial;
    // the walk over one window's live steps, kPrefetch ahead; masked: lanes
    // test their atom against each K column's entry masks (windows holding a
    // partial entry), else whole panel columns are loaded (the panels are zero
    // outside the strips' atoms, which then equal the entries' atoms)
    auto walk = [&](int nsteps, auto masked_tag) {
      constexpr bool kMasked = decltype(masked_tag)::value;
      int cbase = -64;        // first step of the current chunk
      uint64_t live = 0;      // live steps of the chunk not yet taken
      uint32_t chunk_bm = 0;  // lane l: block mask of step cbase + l
      // slot S holds, per group g of 4 K columns, A halves [2g], [2g+1] and
      // B quads [2g], [2g+1] (slots 0..kPrefetch-1 used)
      double ra[8][2 * kGroups], rb[8][2 * kGroups];
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
        _Pragma("unroll") for (int g = 0; g < kGroups; ++g) {                        \
          const int Q = kStepCols * st + 4 * g + kc;                                 \
          const uint2 ab = col_ab[Q];                                                \
          if constexpr (kMasked) {                                                   \
            const uint32_t cm = col_m[Q];                                            \
            _Pragma("unroll") for (int h = 0; h < 2; ++h) {                          \
              const uint32_t need = (bm & (3u << (2 * h))) ? rbit[h] : 0u;           \
              const uint32_t c = (cm & need) ? ab.x : a_zero;                        \
              ra[S][2 * g + h] = a_lane[(int64_t)c * 64 + 16 * h];                   \
            }                                                                        \
            _Pragma("unroll") for (int q = 0; q < 2; ++q) {                          \
              const uint32_t need = (bm & (5u << q)) ? cbit[q] : 0u;                 \
              const uint32_t c = (cm & need) ? ab.y : b_zero;                        \
              rb[S][2 * g + q] = b_lane[(int64_t)c * 64 + 16 * q];                   \
            }                                                                        \
          } else {                                                                   \
            /* past the last live step the loads read step 0 and feed no MFMA */     \
            const double* pa = a_lane + (int64_t)ab.x * 64;                          \
            const double* pb = b_lane + (int64_t)ab.y * 64;                          \
            ra[S][2 * g] = pa[0];                                                    \
            ra[S][2 * g + 1] = pa[16];                                               \
            rb[S][2 * g] = pb[0];                                                    \
            rb[S][2 * g + 1] = pb[16];                                               \
          }                                                                          \
        }                                                                            \
      }
      // the groups' MFMAs interleave over the four accumulators
  #define LAMB_MFMA(S)                                                               \
      {                                                                              \
        const uint32_t bm = blkmask[S];                                              \
        _Pragma("unroll") for (int g = 0; g < kGroups; ++g)                          \
          _Pragma("unroll") for (int h = 0; h < 2; ++h)                              \
            _Pragma("unroll") for (int q = 0; q < 2; ++q)                            \
              if ((bm >> (2 * h + q)) & 1u)                                          \
                acc[2 * h + q] = __builtin_amdgcn_mfma_f64_16x16x4f64(               \
                    ra[S][2 * g + h], rb[S][2 * g + q], acc[2 * h + q], 0, 0, 0);    \
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
    };

    for (int Q0 = 0; Q0 < ktot; Q0 += kTableCols) {
      const int nsteps = build_step_table<kStepCols>(T, sm, entries, Q0, col_ab, col_m, step_g,
                                                     a_zero, b_zero,
                                          &win_partial, tid);
      if (win_partial)
        walk(nsteps, std::true_type{});
      else
        walk(nsteps, std::false_type{});
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
    tile_epilogue<256>(Ct, T, sm, alpha, beta, eps2, old_row_p, old_col, old_off, old_data,
                       c_data, kept_out, tid,
                       kLoop ? slot_counter + (blockIdx.x & 7) : nullptr, &next_slot);
    if constexpr (!kLoop) break;
    __syncthreads();  // next_slot is visible; the next set-up rewrites the LDS
    slot = persistent_slot(xrange, n_tiles, __builtin_amdgcn_readfirstlane(next_slot));
    if (slot < 0) break;
  }
}

// ---------------------------------------------------------------------------
// FP64 numeric phase, K-split variant. One workgroup (4 waves) per C tile as
// above, but every wave owns the whole 64 x 64 tile (sixteen 16 x 16
// accumulators) and the waves split the tile's live K steps: wave w takes
// the live steps whose rank is w mod 4. A step then feeds up to 16 MFMAs from
// 8 loads (the quadrant kernel: 4 MFMAs from 4 loads, each operand loaded by
// two waves). The four partial tiles are summed in LDS in wave order, so the
// result is deterministic. Occupancy: two workgroups per CU.
// ---------------------------------------------------------------------------
#ifndef LAMB_KSPLIT_PREFETCH
#define LAMB_KSPLIT_PREFETCH 2
#endif
constexpr int kKsPrefetch = LAMB_KSPLIT_PREFETCH;

// live 16 x 16 blocks of the whole tile in a step with row / column groups g:
// bit 4h + q = row block h and column block q both hold a live 4-group
__device__ __forceinline__ uint32_t tile_blocks(uint32_t g) {
  const uint32_t rg = g & 0xFFFFu, cg = g >> 16;
  uint32_t rh = 0, cq = 0;
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    rh |= ((rg >> (4 * h)) & 0xFu) ? 1u << h : 0u;
    cq |= ((cg >> (4 * h)) & 0xFu) ? 1u << h : 0u;
  }
  uint32_t bm = 0;
#pragma unroll
  for (int h = 0; h < 4; ++h) bm |= ((rh >> h) & 1u) ? cq << (4 * h) : 0u;
  return bm;
}

// kWhole: every entry of the product covers its strips' atoms (no
// kEntryPartial), so lanes load whole panel columns without masks.
template <bool kWhole, bool kLoop>
__global__ __launch_bounds__(256, 2) void numeric_f64_ksplit_kernel(
    int32_t n_tiles, int32_t* __restrict__ slot_counter, const int32_t* __restrict__ xrange,
    const TileDesc* __restrict__ tiles,
    const int32_t* __restrict__ perm,
    const Entry* __restrict__ entries, const Band* __restrict__ bands,
    const double* __restrict__ a_pack, const double* __restrict__ b_pack, uint32_t a_zero,
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
    uint2* col_ab = reinterpret_cast<uint2*>(lds);
    uint32_t* col_m = reinterpret_cast<uint32_t*>(col_ab + kTableCols);
    uint32_t* step_g = col_m + kTableCols;
    __shared__ TileSmem sm;
    __shared__ int win_partial;

    const int tid = tile_tid<kLoop>();
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const TileInfo T = tile_setup(tiles, perm, bands, sm, slot, tid);
    // atom mask bit of this lane's row in row block h / column in column block q
    uint32_t rbit[4] = {0u, 0u, 0u, 0u}, cbit[4] = {0u, 0u, 0u, 0u};
    for (int a = 0; a < T.natI; ++a)
  #pragma unroll
      for (int h = 0; h < 4; ++h) {
        const int r = 16 * h + (lane & 15);
        if (r >= sm.rs_r[a] && r < sm.rs_r[a + 1]) rbit[h] = 1u << a;
      }
    for (int a = 0; a < T.natJ; ++a)
  #pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int c = 16 * q + (lane & 15);
        if (c >= sm.rs_c[a] && c < sm.rs_c[a + 1]) cbit[q] = 1u << (16 + a);
      }
    const int kc = lane >> 4;
    const double* a_lane = a_pack + (lane & 15);
    const double* b_lane = b_pack + (lane & 15);

    dbl4 acc[16];
  #pragma unroll
    for (int x = 0; x < 16; ++x) acc[x] = dbl4{0.0, 0.0, 0.0, 0.0};
    uint32_t seen = 0;  // live steps met so far (wave-uniform)

    for (int Q0 = 0; Q0 < T.ktot; Q0 += kTableCols) {
      const int nsteps = build_step_table<4>(T, sm, entries, Q0, col_ab, col_m, step_g, a_zero, b_zero,
                                          &win_partial, tid);
      int cbase = -64;
      uint64_t live = 0;
      uint32_t chunk_bm = 0;
      double ra[8][4], rb[8][4];
      uint32_t blkmask[8];
  #define KS_FETCH(S)                                                                \
      {                                                                              \
        uint32_t bm = 0;                                                             \
        int st = 0;                                                                  \
        while (live == 0 && cbase + 64 < nsteps) {                                   \
          cbase += 64;                                                               \
          const int sl = cbase + lane;                                               \
          chunk_bm = tile_blocks(sl < nsteps ? step_g[sl] : 0u);                     \
          /* this wave's live steps: those whose rank is wave mod 4 */               \
          const uint64_t any = __ballot(chunk_bm != 0u);                             \
          const uint32_t rank = seen + __builtin_amdgcn_mbcnt_hi(                    \
              (uint32_t)(any >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)any, 0u));  \
          live = __ballot(chunk_bm != 0u && (rank & 3u) == (uint32_t)wave);          \
          seen += (uint32_t)__builtin_popcountll(any);                               \
        }                                                                            \
        if (live) {                                                                  \
          const int l = __builtin_ctzll(live);                                       \
          live &= live - 1;                                                          \
          st = cbase + l;                                                            \
          bm = __builtin_amdgcn_readlane(chunk_bm, l);                               \
        }                                                                            \
        blkmask[S] = bm;                                                             \
        const int Q = 4 * st + kc;                                                   \
        const uint2 ab = col_ab[Q];                                                  \
        if constexpr (kWhole) {                                                      \
          /* without a step of this wave the loads read step 0, feeding no MFMA */   \
          const double* pa = a_lane + (int64_t)ab.x * 64;                            \
          const double* pb = b_lane + (int64_t)ab.y * 64;                            \
          _Pragma("unroll") for (int h = 0; h < 4; ++h) ra[S][h] = pa[16 * h];       \
          _Pragma("unroll") for (int q = 0; q < 4; ++q) rb[S][q] = pb[16 * q];       \
          /* keeps the scheduler from hoisting these loads over the MFMAs, */        \
          /* which spilled 100+ VGPRs */                                             \
          __builtin_amdgcn_sched_barrier(0);                                         \
        } else {                                                                     \
          const uint32_t cm = col_m[Q];                                              \
          _Pragma("unroll") for (int h = 0; h < 4; ++h) {                            \
            const uint32_t need = (bm & (0xFu << (4 * h))) ? rbit[h] : 0u;           \
            const uint32_t c = (cm & need) ? ab.x : a_zero;                          \
            ra[S][h] = a_lane[(int64_t)c * 64 + 16 * h];                             \
          }                                                                          \
          _Pragma("unroll") for (int q = 0; q < 4; ++q) {                            \
            const uint32_t need = (bm & (0x1111u << q)) ? cbit[q] : 0u;              \
            const uint32_t c = (cm & need) ? ab.y : b_zero;                          \
            rb[S][q] = b_lane[(int64_t)c * 64 + 16 * q];                             \
          }                                                                          \
        }                                                                            \
      }
  #define KS_MFMA(S)                                                                 \
      {                                                                              \
        const uint32_t bm = blkmask[S];                                              \
        _Pragma("unroll") for (int h = 0; h < 4; ++h)                                \
          _Pragma("unroll") for (int q = 0; q < 4; ++q)                              \
            if ((bm >> (4 * h + q)) & 1u)                                            \
              acc[4 * h + q] = __builtin_amdgcn_mfma_f64_16x16x4f64(ra[S][h], rb[S][q], acc[4 * h + q], 0, 0, 0); \
      }
  #define KS_SLOTS(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7)
      static_assert(kKsPrefetch >= 1 && kKsPrefetch <= 8, "prefetch depth 1..8");
  #define KS_PRIME(S) if (S < kKsPrefetch) KS_FETCH(S)
  #define KS_ANY(S) if (S < kKsPrefetch) any |= blkmask[S];
  #define KS_STEP(S) if (S < kKsPrefetch) { KS_MFMA(S); KS_FETCH(S); }
      KS_SLOTS(KS_PRIME)
      for (;;) {
        uint32_t any = 0;
        KS_SLOTS(KS_ANY)
        if (!any) break;
        KS_SLOTS(KS_STEP)
      }
  #undef KS_STEP
  #undef KS_ANY
  #undef KS_PRIME
  #undef KS_SLOTS
  #undef KS_FETCH
  #undef KS_MFMA
    }
    __syncthreads();

    // ---- epilogue: the four partial tiles summed into LDS in wave order ----
    // v_mfma_f64_16x16x4 result: D[i][j] in lane 16*(i%4) + j, register i/4.
    double* Ct = lds;
    for (int w = 0; w < 4; ++w) {
      if (wave == w) {
  #pragma unroll
        for (int h = 0; h < 4; ++h)
  #pragma unroll
          for (int q = 0; q < 4; ++q)
  #pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int idx = (16 * h + 4 * r + (lane >> 4)) * kPitch + 16 * q + (lane & 15);
              const double v = acc[4 * h + q][r];
              Ct[idx] = w == 0 ? v : Ct[idx] + v;
            }
      }
      __syncthreads();
    }
    tile_epilogue<256>(Ct, T, sm, alpha, beta, eps2, old_row_p, old_col, old_off, old_data,
                       c_data, kept_out, tid,
                       kLoop ? slot_counter + (blockIdx.x & 7) : nullptr, &next_slot);
    if constexpr (!kLoop) break;
    __syncthreads();  // next_slot is visible; the next set-up rewrites the LDS
    slot = persistent_slot(xrange, n_tiles, __builtin_amdgcn_readfirstlane(next_slot));
    if (slot < 0) break;
  }
}

// ---------------------------------------------------------------------------
// FP64 numeric phase, half-tile variant. One workgroup (4 waves) per C tile;
// wave w owns the 64 x 32 column half c = w & 1 (eight 16 x 16 accumulators)
// and takes the live steps of that half whose rank has parity p = w >> 1. A
// step feeds up to 8 MFMAs from 6 loads, and each live step of a half is
// walked by two waves (the quadrant kernel: up to four, with 4 MFMAs from 4
// loads). The two partial halves are added in LDS: a two-term sum, so the
// result does not depend on which wave writes first.
// ---------------------------------------------------------------------------
// The fine form's live groups of a step for the column half wc of a tile:
// bits 0-3 the 16-row groups holding a live 4-row group, bits 4-11 the
// half's live 4-column groups; 0 when either is empty.
__device__ __forceinline__ uint32_t fine_groups(uint32_t g, int wc) {
  const uint32_t rg = g & 0xFFFFu;
  const uint32_t cg = (g >> (16 + 8 * wc)) & 0xFFu;
  uint32_t rh = 0;
#pragma unroll
  for (int h = 0; h < 4; ++h) rh |= ((rg >> (4 * h)) & 0xFu) ? 1u << h : 0u;
  return (rh && cg) ? (rh | (cg << 4)) : 0u;
}

// B operand of v_mfma_f64_4x4x4_4b with all four blocks on the columns
// 4c..4c+3 of a 16-column block, from that block's v_mfma_f64_16x16x4 B
// operand (lane 16k + j holds B[k][j]): lane 16k + 4b + i takes lane
// 16k + 4c + i, a ds_swizzle in bit mode (and 0x13, or 4c) on each half.
__device__ __forceinline__ double replicate_cols(double v, int c) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  uint32_t lo, hi;
  switch (c) {  // the pattern is an immediate
    case 0:
      lo = (uint32_t)__builtin_amdgcn_ds_swizzle((int)(uint32_t)u, 0x13);
      hi = (uint32_t)__builtin_amdgcn_ds_swizzle((int)(uint32_t)(u >> 32), 0x13);
      break;
    case 1:
      lo = (uint32_t)__builtin_amdgcn_ds_swizzle((int)(uint32_t)u, 0x13 | (4 << 5));
      hi = (uint32_t)__builtin_amdgcn_ds_swizzle((int)(uint32_t)(u >> 32), 0x13 | (4 << 5));
      break;
    case 2:
      lo = (uint32_t)__builtin_amdgcn_ds_swizzle((int)(uint32_t)u, 0x13 | (8 << 5));
      hi = (uint32_t)__builtin_amdgcn_ds_swizzle((int)(uint32_t)(u >> 32), 0x13 | (8 << 5));
      break;
    default:
      lo = (uint32_t)__builtin_amdgcn_ds_swizzle((int)(uint32_t)u, 0x13 | (12 << 5));
      hi = (uint32_t)__builtin_amdgcn_ds_swizzle((int)(uint32_t)(u >> 32), 0x13 | (12 << 5));
      break;
  }
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

#ifndef LAMB_HALF_PREFETCH
#define LAMB_HALF_PREFETCH 3
#endif
constexpr int kHalfPrefetch = LAMB_HALF_PREFETCH;

template <bool kWhole, bool kLoop, bool kRuns = false, bool kFine = false>
__global__ __launch_bounds__(256, 4) void numeric_f64_half_kernel(
    int32_t n_tiles, int32_t* __restrict__ slot_counter, const int32_t* __restrict__ xrange,
    const TileDesc* __restrict__ tiles,
    const int32_t* __restrict__ perm,
    const Entry* __restrict__ entries, const Band* __restrict__ bands,
    const double* __restrict__ a_pack, const double* __restrict__ b_pack, uint32_t a_zero,
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
    uint2* col_ab = reinterpret_cast<uint2*>(lds);
    uint32_t* col_m = reinterpret_cast<uint32_t*>(col_ab + kTableCols);
    uint32_t* step_g = col_m + kTableCols;
    __shared__ TileSmem sm;
    __shared__ int win_partial;

    const int tid = tile_tid<kLoop>();
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wc = wave & 1;
    const uint32_t par = (uint32_t)(wave >> 1);
    const TileInfo T = tile_setup(tiles, perm, bands, sm, slot, tid);
    // atom mask bit of this lane's row in row block h / column in column block q
    uint32_t rbit[4] = {0u, 0u, 0u, 0u}, cbit[2] = {0u, 0u};
    for (int a = 0; a < T.natI; ++a)
  #pragma unroll
      for (int h = 0; h < 4; ++h) {
        const int r = 16 * h + (lane & 15);
        if (r >= sm.rs_r[a] && r < sm.rs_r[a + 1]) rbit[h] = 1u << a;
      }
    for (int a = 0; a < T.natJ; ++a)
  #pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int c = 32 * wc + 16 * q + (lane & 15);
        if (c >= sm.rs_c[a] && c < sm.rs_c[a + 1]) cbit[q] = 1u << (16 + a);
      }
    const int kc = lane >> 4;
    const double* a_lane = a_pack + (lane & 15);
    const double* b_lane = b_pack + 32 * wc + (lane & 15);
    // byte offsets of this lane's value in a panel column and in a run of
    // four columns (kRuns)
    const char* a_bytes = reinterpret_cast<const char*>(a_pack);
    const char* b_bytes = reinterpret_cast<const char*>(b_pack);
    const uint32_t row_a = 8u * (lane & 15), row_b = 8u * (32 * wc + (lane & 15));
    const uint32_t run_a = row_a + 512u * kc, run_b = row_b + 512u * kc;
    // bits 4h + 2wc + q of a tile block mask: this wave's half
    const uint32_t half_bits = 0x3333u << (2 * wc);

    static_assert(!kFine || (kWhole && kRuns), "the fine form serves whole products");
    dbl4 acc[8];
    // kFine: row group h x 16-column block q x 4-column group c, one
    // v_mfma_f64_4x4x4_4b accumulator each (16 rows x 4 columns)
    double accf[4][2][4];
  #pragma unroll
    for (int x = 0; x < 8; ++x) acc[x] = dbl4{0.0, 0.0, 0.0, 0.0};
  #pragma unroll
    for (int h = 0; h < 4; ++h)
  #pragma unroll
      for (int q = 0; q < 2; ++q)
  #pragma unroll
        for (int c = 0; c < 4; ++c) accf[h][q][c] = 0.0;
    uint32_t seen = 0;  // live steps of this half met so far (wave-uniform)

    for (int Q0 = 0; Q0 < T.ktot; Q0 += kTableCols) {
      const int nsteps = build_step_table<4>(T, sm, entries, Q0, col_ab, col_m, step_g, a_zero,
                                             b_zero, &win_partial, tid);
      int cbase = -64;
      uint64_t live = 0;
      uint32_t chunk_bm = 0;
      // with kRuns, lane l of a chunk also holds step cbase + l's first A and
      // B panel columns, and bit 16 of its block mask says whether the step's
      // 4 columns are consecutive in both panels (inside one entry they are):
      // such a step loads from a wave-uniform column and fixed lane offsets,
      // with no table read between finding the step and loading it
      uint32_t chunk_a = 0, chunk_b = 0;
      double ra[8][4], rb[8][2];
      uint32_t blkmask[8];
  #define HF_FETCH(S)                                                                \
      {                                                                              \
        uint32_t bm = 0;                                                             \
        int st = 0;                                                                  \
        while (live == 0 && cbase + 64 < nsteps) {                                   \
          cbase += 64;                                                               \
          const int sl = cbase + lane;                                               \
          if constexpr (kFine)                                                       \
            chunk_bm = fine_groups(sl < nsteps ? step_g[sl] : 0u, wc);               \
          else                                                                       \
            chunk_bm = tile_blocks(sl < nsteps ? step_g[sl] : 0u) & half_bits;       \
          if constexpr (kRuns) {                                                     \
            uint32_t c0a = 0, c0b = 0;                                               \
            if (chunk_bm) {                                                          \
              const uint2 c0 = col_ab[4 * sl];                                       \
              c0a = c0.x;                                                            \
              c0b = c0.y;                                                            \
              bool run = true;                                                       \
              _Pragma("unroll") for (int q = 1; q < 4; ++q) {                        \
                const uint2 c = col_ab[4 * sl + q];                                  \
                run = run && c.x == c0a + q && c.y == c0b + q;                       \
              }                                                                      \
              if (run) chunk_bm |= 1u << 16;                                         \
            }                                                                        \
            chunk_a = c0a;                                                           \
            chunk_b = c0b;                                                           \
          }                                                                          \
          /* this wave's live steps: those of rank parity par in the half */         \
          const uint64_t any = __ballot(chunk_bm != 0u);                             \
          const uint32_t rank = seen + __builtin_amdgcn_mbcnt_hi(                    \
              (uint32_t)(any >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)any, 0u));  \
          live = __ballot(chunk_bm != 0u && (rank & 1u) == par);                     \
          seen += (uint32_t)__builtin_popcountll(any);                               \
        }                                                                            \
        uint32_t a0 = 0, b0 = 0;                                                     \
        if (live) {                                                                  \
          const int l = __builtin_ctzll(live);                                       \
          live &= live - 1;                                                          \
          st = cbase + l;                                                            \
          bm = __builtin_amdgcn_readlane(chunk_bm, l);                               \
          if constexpr (kRuns) {                                                     \
            a0 = __builtin_amdgcn_readlane(chunk_a, l);                              \
            b0 = __builtin_amdgcn_readlane(chunk_b, l);                              \
          }                                                                          \
        }                                                                            \
        blkmask[S] = bm & 0xFFFFu;                                                   \
        if constexpr (kRuns) {                                                       \
          /* byte offsets (panels below 4 GB); without a step: step 0 */             \
          uint32_t va, vb;                                                           \
          if (bm & (1u << 16)) {                                                     \
            va = a0 * 512u + run_a;                                                  \
            vb = b0 * 512u + run_b;                                                  \
          } else {                                                                   \
            const uint2 ab = col_ab[4 * st + kc];                                    \
            va = ab.x * 512u + row_a;                                                \
            vb = ab.y * 512u + row_b;                                                \
          }                                                                          \
          _Pragma("unroll") for (int h = 0; h < 4; ++h)                              \
            ra[S][h] = *reinterpret_cast<const double*>(a_bytes + va + 128 * h);     \
          _Pragma("unroll") for (int q = 0; q < 2; ++q)                              \
            rb[S][q] = *reinterpret_cast<const double*>(b_bytes + vb + 128 * q);     \
          __builtin_amdgcn_sched_barrier(0);                                         \
        } else {                                                                     \
        const int Q = 4 * st + kc;                                                   \
        const uint2 ab = col_ab[Q];                                                  \
        if constexpr (kWhole) {                                                      \
          /* without a step of this wave the loads read step 0, feeding no MFMA */   \
          const double* pa = a_lane + (int64_t)ab.x * 64;                            \
          const double* pb = b_lane + (int64_t)ab.y * 64;                            \
          _Pragma("unroll") for (int h = 0; h < 4; ++h) ra[S][h] = pa[16 * h];       \
          _Pragma("unroll") for (int q = 0; q < 2; ++q) rb[S][q] = pb[16 * q];       \
          __builtin_amdgcn_sched_barrier(0);                                         \
        } else {                                                                     \
          const uint32_t cm = col_m[Q];                                              \
          _Pragma("unroll") for (int h = 0; h < 4; ++h) {                            \
            const uint32_t need = (bm & (0xFu << (4 * h))) ? rbit[h] : 0u;           \
            const uint32_t c = (cm & need) ? ab.x : a_zero;                          \
            ra[S][h] = a_lane[(int64_t)c * 64 + 16 * h];                             \
          }                                                                          \
          _Pragma("unroll") for (int q = 0; q < 2; ++q) {                            \
            const uint32_t need = (bm & (0x1111u << (2 * wc + q))) ? cbit[q] : 0u;   \
            const uint32_t c = (cm & need) ? ab.y : b_zero;                          \
            rb[S][q] = b_lane[(int64_t)c * 64 + 16 * q];                             \
          }                                                                          \
        }                                                                            \
        }                                                                            \
      }
  #define HF_MFMA(S)                                                                 \
      if constexpr (kFine) {                                                         \
        /* bits 0-3: live 16-row groups; bits 4 + 4q + c: live 4-column groups */    \
        const uint32_t bm = blkmask[S];                                              \
        _Pragma("unroll") for (int q = 0; q < 2; ++q)                                \
          _Pragma("unroll") for (int c = 0; c < 4; ++c)                              \
            if (__builtin_expect((bm >> (4 + 4 * q + c)) & 1u, 1)) {                 \
              const double bc = replicate_cols(rb[S][q], c);                         \
              _Pragma("unroll") for (int h = 0; h < 4; ++h)                          \
                if (__builtin_expect((bm >> h) & 1u, 1))                             \
                  accf[h][q][c] = __builtin_amdgcn_mfma_f64_4x4x4f64(ra[S][h], bc, accf[h][q][c], 0, 0, 0); \
            }                                                                        \
      } else {                                                                       \
        const uint32_t bm = blkmask[S] >> (2 * wc);                                  \
        _Pragma("unroll") for (int h = 0; h < 4; ++h)                                \
          _Pragma("unroll") for (int q = 0; q < 2; ++q)                              \
            if ((bm >> (4 * h + q)) & 1u)                                            \
              acc[2 * h + q] = __builtin_amdgcn_mfma_f64_16x16x4f64(ra[S][h], rb[S][q], acc[2 * h + q], 0, 0, 0); \
      }
  #define HF_SLOTS(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7)
      // the masked walk holds per-lane atom masks besides the ring: two steps
      // in flight keep it within the registers of four waves per SIMD
      constexpr int kDepth = kWhole ? kHalfPrefetch : (kHalfPrefetch < 2 ? kHalfPrefetch : 2);
      static_assert(kHalfPrefetch >= 1 && kHalfPrefetch <= 8, "prefetch depth 1..8");
  #define HF_PRIME(S) if (S < kDepth) HF_FETCH(S)
  #define HF_ANY(S) if (S < kDepth) any |= blkmask[S];
  #define HF_STEP(S) if (S < kDepth) { HF_MFMA(S); HF_FETCH(S); }
      HF_SLOTS(HF_PRIME)
      for (;;) {
        uint32_t any = 0;
        HF_SLOTS(HF_ANY)
        if (!any) break;
        HF_SLOTS(HF_STEP)
      }
  #undef HF_STEP
  #undef HF_ANY
  #undef HF_PRIME
  #undef HF_SLOTS
  #undef HF_FETCH
  #undef HF_MFMA
    }
    __syncthreads();

    // ---- epilogue: the two partial halves added in LDS ----
    // v_mfma_f64_16x16x4 result: D[i][j] in lane 16*(i%4) + j, register i/4.
    double* Ct = lds;
    for (int p = 0; p < 2; ++p) {
      if ((wave >> 1) == p) {
        if constexpr (kFine) {
          // v_mfma_f64_4x4x4_4b result: element (i, j) of block b in lane
          // 16 i + 4 b + j, block b = rows 4b.. of the 16-row group
          const int fr = 4 * ((lane >> 2) & 3) + (lane >> 4), fc = lane & 3;
  #pragma unroll
          for (int h = 0; h < 4; ++h)
  #pragma unroll
            for (int q = 0; q < 2; ++q)
  #pragma unroll
              for (int c = 0; c < 4; ++c) {
                const int idx = (16 * h + fr) * kPitch + 32 * wc + 16 * q + 4 * c + fc;
                const double v = accf[h][q][c];
                Ct[idx] = p == 0 ? v : Ct[idx] + v;
              }
        } else {
  #pragma unroll
          for (int h = 0; h < 4; ++h)
  #pragma unroll
            for (int q = 0; q < 2; ++q)
  #pragma unroll
              for (int r = 0; r < 4; ++r) {
                const int idx =
                    (16 * h + 4 * r + (lane >> 4)) * kPitch + 32 * wc + 16 * q + (lane & 15);
                const double v = acc[2 * h + q][r];
                Ct[idx] = p == 0 ? v : Ct[idx] + v;
              }
        }
      }
      __syncthreads();
    }
    tile_epilogue<256>(Ct, T, sm, alpha, beta, eps2, old_row_p, old_col, old_off, old_data,
                       c_data, kept_out, tid,
                       kLoop ? slot_counter + (blockIdx.x & 7) : nullptr, &next_slot);
    if constexpr (!kLoop) break;
    __syncthreads();  // next_slot is visible; the next set-up rewrites the LDS
    slot = persistent_slot(xrange, n_tiles, __builtin_amdgcn_readfirstlane(next_slot));
    if (slot < 0) break;
  }
}

// ---------------------------------------------------------------------------
// FP64 numeric phase on the vector ALU, for whole products. One workgroup (4
// waves) per C tile; wave g owns the 16-row group g of the tile and walks the
// tile twice, once per 32-column half. Lane (kc, r) = (lane >> 4, lane & 15)
// holds row 16g + r and accumulates, for the 32 columns of the half, the K
// columns of each step of 4 whose position in the step is kc. Per step the
// wave loads the A panel column values of its 16 rows (one per lane, the
// v_mfma_f64_16x16x4 A layout) and the B values of the half's 32 columns
// (two loads, the MFMA B layout), then runs up to 32 v_fmac_f64 whose B
// operand is broadcast from lane (kc, j) to the lane group kc by DPP
// row_newbcast:j. v_fmac_f64 issues at the full FP64 rate, against about
// 100 cycles per v_mfma_f64_16x16x4, and the live-block granularity is 16
// rows x 4 columns instead of 16 x 16: a step skips a wave whose row group
// is dead, and a 4-column group whose columns are dead. After a half's walk
// the four K partials of each value are added across the lane groups (xor
// 32, then xor 16: the same two-level sum in every lane), and lane (kc, r)
// keeps the columns 4j + kc. Panels are zero outside their strips' atoms and
// every entry is whole, so dead rows and columns inside a live group add
// exact zeros.
// ---------------------------------------------------------------------------
// K columns of the lanes kernel's step table, and where its first half
// waits (in doubles from the start of the tile's LDS buffer)
constexpr int kLanesCols = 1024;
constexpr int kLanesHalf0 = 1664;
#ifndef LAMB_LANES_DEPTH
#define LAMB_LANES_DEPTH 2
#endif
constexpr int kLanesDepth = LAMB_LANES_DEPTH;  // steps in flight per wave (1..3)
static_assert(kLanesDepth >= 1 && kLanesDepth <= 3, "lanes kernel depth");

#define LAMB_FMAC_BCAST(ACC, B, A, J)                                                        \
  asm("v_fmac_f64_dpp %0, %1, %2 row_newbcast:" #J " row_mask:0xf bank_mask:0xf"          \
      : "+v"(ACC)                                                                          \
      : "v"(B), "v"(A))

template <bool kLoop>
__global__ __launch_bounds__(256, 4) void numeric_f64_lanes_kernel(
    int32_t n_tiles, int32_t* __restrict__ slot_counter, const int32_t* __restrict__ xrange,
    const TileDesc* __restrict__ tiles,
    const int32_t* __restrict__ perm,
    const Entry* __restrict__ entries, const Band* __restrict__ bands,
    const double* __restrict__ a_pack, const double* __restrict__ b_pack, uint32_t a_zero,
    uint32_t b_zero, double alpha, double beta, double eps2,
    const int64_t* __restrict__ old_row_p, const int32_t* __restrict__ old_col,
    const int64_t* __restrict__ old_off, const double* __restrict__ old_data,
    double* __restrict__ c_data, uint16_t* __restrict__ kept_out) {
  LAMB_GUARD;
  __shared__ int32_t next_slot;
  int32_t slot = (int32_t)blockIdx.x;
  if constexpr (kLoop) {
    slot = persistent_slot(xrange, n_tiles, (int32_t)(blockIdx.x >> 3));
    if (slot < 0) return;
  }
  for (;;) {
    // main loop: the step table of a window of kLanesCols K columns, then
    // the reduced first half (64 x 32 values); epilogue: C tile
    __shared__ double lds[kBandRows * kPitch];
    uint2* col_ab = reinterpret_cast<uint2*>(lds);
    uint32_t* col_m = reinterpret_cast<uint32_t*>(col_ab + kLanesCols);
    uint32_t* step_g = col_m + kLanesCols;
    double* half0 = lds + kLanesHalf0;
    static_assert(sizeof(uint32_t) * (3 * kLanesCols + kLanesCols / 4) <=
                      sizeof(double) * kLanesHalf0 &&
                  kLanesHalf0 + kBandRows * 32 <= kBandRows * kPitch,
                  "step table and first half exceed the LDS buffer");
    __shared__ TileSmem sm;
    __shared__ int win_partial;

    const int tid = tile_tid<kLoop>();
    const int lane = tid & 63;
    const int g = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int kc = lane >> 4, r = lane & 15;
    const TileInfo T = tile_setup(tiles, perm, bands, sm, slot, tid);
    const char* a_bytes = reinterpret_cast<const char*>(a_pack);
    const char* b_bytes = reinterpret_cast<const char*>(b_pack);
    const uint32_t row_a = (uint32_t)(8 * (16 * g + r));  // byte offset in a panel column
    const uint32_t off_a = row_a + 512u * kc;             // ... and in a run of columns
    // lane (kc, r): C(16g + r, 32 + 4j + kc) after the second half
    double keep1[8];

#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
      const uint32_t row_b = (uint32_t)(8 * (32 * h + r));
      const uint32_t off_b = row_b + 512u * kc;
      double acc[32];
#pragma unroll
      for (int x = 0; x < 32; ++x) acc[x] = 0.0;
      for (int Q0 = 0; Q0 < T.ktot; Q0 += kLanesCols) {
        // one window's table serves both halves
        const int nsteps =
            (h == 0 || T.ktot > kLanesCols)
                ? build_step_table<8, kLanesCols>(T, sm, entries, Q0, col_ab, col_m, step_g,
                                                  a_zero, b_zero, &win_partial, tid)
                : (min(kLanesCols, T.ktot - Q0) + 7) / 8;
        // Steps of 8 K columns: lane (kc, r) takes columns kc and 4 + kc. Per
        // chunk of 64 steps, lane l reads step cbase + l's first panel
        // columns and whether its 8 columns are consecutive in both panels
        // (bit 8 of the mask; inside one entry they are): such a step's loads
        // take a wave-uniform base and fixed lane offsets, with no table read
        // between finding the step and loading its operands.
        int cbase = -64;
        uint64_t live = 0;
        uint32_t chunk_m = 0, chunk_a = 0, chunk_b = 0;
        double ra[3][2], rb[3][4];
        uint32_t cm[3];
#define LN_FETCH(S)                                                                   \
        {                                                                             \
          while (live == 0 && cbase + 64 < nsteps) {                                  \
            cbase += 64;                                                              \
            const int sl = cbase + lane;                                              \
            uint32_t m = 0, c0a = 0, c0b = 0;                                         \
            if (sl < nsteps) {                                                        \
              const uint32_t sg = step_g[sl];                                         \
              m = ((sg >> (4 * g)) & 0xFu) ? ((sg >> (16 + 8 * h)) & 0xFFu) : 0u;     \
              if (m) {                                                                \
                const uint2 c0 = col_ab[8 * sl];                                      \
                c0a = c0.x;                                                           \
                c0b = c0.y;                                                           \
                bool run = true;                                                      \
                _Pragma("unroll") for (int q = 1; q < 8; ++q) {                       \
                  const uint2 c = col_ab[8 * sl + q];                                 \
                  run = run && c.x == c0a + q && c.y == c0b + q;                      \
                }                                                                     \
                if (run) m |= 0x100u;                                                 \
              }                                                                       \
            }                                                                         \
            chunk_m = m;                                                              \
            chunk_a = c0a;                                                            \
            chunk_b = c0b;                                                            \
            live = __ballot(m != 0u);                                                 \
          }                                                                           \
          uint32_t m = 0, a0 = 0, b0 = 0;                                             \
          int st = 0;                                                                 \
          if (live) {                                                                 \
            const int l = __builtin_ctzll(live);                                      \
            live &= live - 1;                                                         \
            st = cbase + l;                                                           \
            m = __builtin_amdgcn_readlane(chunk_m, l);                                \
            a0 = __builtin_amdgcn_readlane(chunk_a, l);                               \
            b0 = __builtin_amdgcn_readlane(chunk_b, l);                               \
          }                                                                           \
          cm[S] = m & 0xFFu;                                                          \
          /* byte offsets into the panels (below 4 GB each, which the variant  */   \
          /* rule checks): a run of 8 columns from the step's first columns,    */   \
          /* else each lane's own columns from the table                        */   \
          uint32_t va0, va1, vb0, vb1;                                                \
          if (m & 0x100u) {                                                           \
            va0 = a0 * 512u + off_a;                                                  \
            va1 = va0 + 2048u;                                                        \
            vb0 = b0 * 512u + off_b;                                                  \
            vb1 = vb0 + 2048u;                                                        \
          } else {                                                                    \
            /* without a step the loads read step 0 and feed no FMA */               \
            const uint2 c0 = col_ab[8 * st + kc], c1 = col_ab[8 * st + 4 + kc];       \
            va0 = c0.x * 512u + row_a;                                                \
            va1 = c1.x * 512u + row_a;                                                \
            vb0 = c0.y * 512u + row_b;                                                \
            vb1 = c1.y * 512u + row_b;                                                \
          }                                                                           \
          ra[S][0] = *reinterpret_cast<const double*>(a_bytes + va0);                 \
          ra[S][1] = *reinterpret_cast<const double*>(a_bytes + va1);                 \
          rb[S][0] = *reinterpret_cast<const double*>(b_bytes + vb0);                 \
          rb[S][1] = *reinterpret_cast<const double*>(b_bytes + vb0 + 128);           \
          rb[S][2] = *reinterpret_cast<const double*>(b_bytes + vb1);                 \
          rb[S][3] = *reinterpret_cast<const double*>(b_bytes + vb1 + 128);           \
          __builtin_amdgcn_sched_barrier(0);                                          \
        }
#define LN_GROUP(S, C, X, J0, J1, J2, J3)                                             \
        if (__builtin_expect((cm[S] >> (X)) & 1u, 1)) {                               \
          LAMB_FMAC_BCAST(acc[4 * (X) + 0], rb[S][C], ra[S][0], J0);                  \
          LAMB_FMAC_BCAST(acc[4 * (X) + 1], rb[S][C], ra[S][0], J1);                  \
          LAMB_FMAC_BCAST(acc[4 * (X) + 2], rb[S][C], ra[S][0], J2);                  \
          LAMB_FMAC_BCAST(acc[4 * (X) + 3], rb[S][C], ra[S][0], J3);                  \
          LAMB_FMAC_BCAST(acc[4 * (X) + 0], rb[S][C + 2], ra[S][1], J0);              \
          LAMB_FMAC_BCAST(acc[4 * (X) + 1], rb[S][C + 2], ra[S][1], J1);              \
          LAMB_FMAC_BCAST(acc[4 * (X) + 2], rb[S][C + 2], ra[S][1], J2);              \
          LAMB_FMAC_BCAST(acc[4 * (X) + 3], rb[S][C + 2], ra[S][1], J3);              \
        }
#define LN_FMA(S)                                                                     \
        {                                                                             \
          LN_GROUP(S, 0, 0, 0, 1, 2, 3)                                               \
          LN_GROUP(S, 0, 1, 4, 5, 6, 7)                                               \
          LN_GROUP(S, 0, 2, 8, 9, 10, 11)                                             \
          LN_GROUP(S, 0, 3, 12, 13, 14, 15)                                           \
          LN_GROUP(S, 1, 4, 0, 1, 2, 3)                                               \
          LN_GROUP(S, 1, 5, 4, 5, 6, 7)                                               \
          LN_GROUP(S, 1, 6, 8, 9, 10, 11)                                             \
          LN_GROUP(S, 1, 7, 12, 13, 14, 15)                                           \
        }
#define LN_SLOTS(X) X(0) X(1) X(2)
#define LN_PRIME(S) if (S < kLanesDepth) LN_FETCH(S)
#define LN_ANY(S) if (S < kLanesDepth) any |= cm[S];
#define LN_STEP(S) if (S < kLanesDepth) { LN_FMA(S); LN_FETCH(S); }
        LN_SLOTS(LN_PRIME)
        for (;;) {
          uint32_t any = 0;
          LN_SLOTS(LN_ANY)
          if (!any) break;
          LN_SLOTS(LN_STEP)
        }
#undef LN_STEP
#undef LN_ANY
#undef LN_PRIME
#undef LN_SLOTS
#undef LN_FMA
#undef LN_GROUP
#undef LN_FETCH
      }
      // the four K partials of column c = 4j + kc, summed into lane group kc
      // as (p_kc + p_kc^2) + (p_kc^1 + p_kc^3): across xor 32 each lane group
      // keeps the column pair whose bit 1 is its own, across xor 16 the
      // column of its own parity
      const bool up = (kc & 2) != 0, odd = (kc & 1) != 0;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        double k0 = up ? acc[4 * j + 2] : acc[4 * j];
        double k1 = up ? acc[4 * j + 3] : acc[4 * j + 1];
        const double s0 = up ? acc[4 * j] : acc[4 * j + 2];
        const double s1 = up ? acc[4 * j + 1] : acc[4 * j + 3];
        k0 += __shfl_xor(s0, 32);
        k1 += __shfl_xor(s1, 32);
        double v = odd ? k1 : k0;
        v += __shfl_xor(odd ? k0 : k1, 16);
        // the first half waits in LDS beside the table (rows of 32 values)
        if (h == 0) half0[(16 * g + r) * 32 + 4 * j + kc] = v;
        else keep1[j] = v;
      }
    }
    double keep0[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) keep0[j] = half0[(16 * g + r) * 32 + 4 * j + kc];
    __syncthreads();  // every wave is done with the table and the first half

    double* Ct = lds;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      Ct[(16 * g + r) * kPitch + 4 * j + kc] = keep0[j];
      Ct[(16 * g + r) * kPitch + 32 + 4 * j + kc] = keep1[j];
    }
    __syncthreads();
    tile_epilogue<256>(Ct, T, sm, alpha, beta, eps2, old_row_p, old_col, old_off, old_data,
                       c_data, kept_out, tid,
                       kLoop ? slot_counter + (blockIdx.x & 7) : nullptr, &next_slot);
    if constexpr (!kLoop) break;
    __syncthreads();  // next_slot is visible; the next set-up rewrites the LDS
    slot = persistent_slot(xrange, n_tiles, __builtin_amdgcn_readfirstlane(next_slot));
    if (slot < 0) break;
  }
}
#undef LAMB_FMAC_BCAST

// Workgroups of a 256-thread kernel resident on the current device at once
// (resident per CU times the CUs): the grid of a persistent launch. Cached
// per kernel and device; 0 when the runtime cannot tell.
int32_t resident_workgroups(const void* kernel) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, int32_t> cache;
  std::lock_guard<std::mutex> lock(mu);
  const auto key = std::make_pair(kernel, dev);
  if (auto it = cache.find(key); it != cache.end()) return it->second;
  int per_cu = 0, n_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess ||
      hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  const int32_t n = (int32_t)std::max(0, per_cu) * std::max(0, n_cu);
  cache[key] = n;
  return n;
}

// Sets a persistent grid's eight slot counters: counter x counts the turns
// already taken by the workgroups of residue x (blockIdx.x mod 8), whose
// first turns are 0, 1, ... (blockIdx.x / 8), one per workgroup. A kernel,
// not a memset, so that graph captures replay it.
__global__ void slot_counter_reset(int32_t* __restrict__ counter, int32_t grid) {
  LAMB_GUARD;
  const int x = threadIdx.x;
  if (x < 8) counter[x] = (grid - x + 7) / 8;
}

void launch_slot_counter_reset(hipStream_t s, int32_t* counter, int32_t grid) {
  hipLaunchKernelGGL(slot_counter_reset, dim3(1), dim3(64), 0, s, counter, grid);
}

// The lanes kernel and the half kernel's run walk address the panels by
// 32-bit byte offsets: both panels, the zero column after the last one
// included, must stay below 4 GB.
bool lanes_panels_fit(uint32_t a_zero, uint32_t b_zero) {
  return ((uint64_t)std::max(a_zero, b_zero) + 1) * 512u <= 0xFFFFFFFFull;
}

hipError_t launch_numeric_f64(hipStream_t s, int32_t n_tiles, const TileDesc* tiles,
                              const int32_t* perm, const Entry* entries, const Band* bands,
                              const double* a_pack, const double* b_pack, uint32_t a_zero,
                              uint32_t b_zero, double alpha, double beta, double eps2,
                              const int64_t* old_row_p, const int32_t* old_col,
                              const int64_t* old_off, const double* old_data, double* c_data,
                              uint16_t* kept, int variant, bool whole, int form,
                              int32_t* slot_counter, int32_t* grid_used) {
  if (grid_used) *grid_used = 0;
  if (n_tiles == 0) return hipSuccess;
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
                       tiles, perm, entries, bands, a_pack, b_pack, a_zero, b_zero, alpha, beta,
                       eps2, old_row_p, old_col, old_off, old_data, c_data, kept);
  };
  // a persistent grid runs the looping instantiation; one workgroup per tile
  // the straight one, whose registers are those of a single tile
  if (variant == kF64Ksplit && whole)
    form != kFormGrid ? go(numeric_f64_ksplit_kernel<true, true>)
               : go(numeric_f64_ksplit_kernel<true, false>);
  else if (variant == kF64Ksplit)
    form != kFormGrid ? go(numeric_f64_ksplit_kernel<false, true>)
               : go(numeric_f64_ksplit_kernel<false, false>);
  else if (variant == kF64Lanes && whole && lanes_panels_fit(a_zero, b_zero))
    form != kFormGrid ? go(numeric_f64_lanes_kernel<true>)
               : go(numeric_f64_lanes_kernel<false>);
  else if (variant == kF64Fine && whole && lanes_panels_fit(a_zero, b_zero))
    form != kFormGrid ? go(numeric_f64_half_kernel<true, true, true, true>)
               : go(numeric_f64_half_kernel<true, false, true, true>);
  else if ((variant == kF64Half || variant == kF64Fine) && whole &&
           lanes_panels_fit(a_zero, b_zero))
    form != kFormGrid ? go(numeric_f64_half_kernel<true, true, true>)
               : go(numeric_f64_half_kernel<true, false, true>);
  else if ((variant == kF64Half || variant == kF64Fine) && whole)
    form != kFormGrid ? go(numeric_f64_half_kernel<true, true>)
               : go(numeric_f64_half_kernel<true, false>);
  else if (variant == kF64Half || variant == kF64Lanes || variant == kF64Fine)
    // the lanes kernel and the fine form: whole products only
    form != kFormGrid ? go(numeric_f64_half_kernel<false, true>)
               : go(numeric_f64_half_kernel<false, false>);
  else
    form != kFormGrid ? go(numeric_f64_kernel<true>) : go(numeric_f64_kernel<false>);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Trace: per-atom partial sums of the diagonal of diagonal blocks.
// ---------------------------------------------------------------------------
__global__ void trace_kernel(int32_t n_atoms, const int64_t* __restrict__ row_p,
                             const int32_t* __restrict__ col, const int64_t* __restrict__ off,
                             const int32_t* __restrict__ blk, const double* __restrict__ data,
                             double* __restrict__ out) {
  LAMB_GUARD;
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
