// Pieces shared by the numeric tile kernels (FP64 and BF16-split): tile
// set-up, the per-lane atom masks of a wave's 32 x 32 quadrant, the live
// block mask of a K step, and the epilogue (alpha / beta, block norms,
// post-filter, C pool writes).
#pragma once

#include "common.h"

namespace lamb {

// LDS of one tile besides the main-loop table / C tile buffer
struct TileSmem {
  int16_t rs_r[kBandAtoms + 1], rs_c[kBandAtoms + 1];  // atom row / col starts
  uint16_t pg_r[kBandAtoms], pg_c[kBandAtoms];          // 4-row / 4-col groups per atom
  int32_t pair_id[kBandAtoms * kBandAtoms];
  int64_t pair_off[kBandAtoms * kBandAtoms];
  int32_t wave_cnt[4];
  int64_t wave_sz[4];
  uint32_t keptw[kBandAtoms];
  int32_t max_rows, max_cols;  // largest row / column atom of the tile
  double zero;                 // read by norm-chain slots outside a block
  int32_t n_exact;             // blocks whose post-filter needs the exact chain
  int16_t exact_id[kBandAtoms * kBandAtoms];
};

// Tile slot (position in the tile order) of the j-th tile of a persistent
// workgroup of residue x = blockIdx.x mod 8, the XCD the hardware deals it
// to. Interleaved (xrange null): slot x + 8 j, below n_tiles, so the XCDs
// share one window of the order. XCD ranges: position xrange[x] + j of the
// range [xrange[x], xrange[x + 1]) that residue x walks alone. -1 past the
// end of the workgroup's sequence.
__device__ __forceinline__ int32_t persistent_slot(const int32_t* __restrict__ xrange,
                                                   int32_t n_tiles, int32_t j) {
  const int32_t x = (int32_t)(blockIdx.x & 7);
  if (xrange != nullptr) {
    const int32_t s = xrange[x] + j;
    return s < xrange[x + 1] ? s : -1;
  }
  const int32_t s = x + 8 * j;
  return s < n_tiles ? s : -1;
}

// v of the lane a DPP control selects (0 where it selects none), for a
// 64-bit value as two 32-bit moves
template <int kCtrl, int kRowMask>
__device__ __forceinline__ double dpp_f64(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, kCtrl, kRowMask,
                                                           0xf, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), kCtrl,
                                                           kRowMask, 0xf, false);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Inclusive prefix sum over the wave's 64 lanes: row prefix sums by
// row_shr 1, 2, 4, 8, then row_bcast 15 and 31 carry the earlier rows in.
__device__ __forceinline__ int wave_scan_dpp(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);
  return v;
}

// Sums over each 16-lane row of the wave, returned to the lanes of that row
// (g = lane >> 4): row prefix sums by row_shr 1, 2, 4, 8 (each value meets
// at most four additions) leave the row totals in lanes 15, 31, 47 and 63.
__device__ __forceinline__ double row_totals_dpp(double v, int g) {
  v += dpp_f64<0x111, 0xf>(v);
  v += dpp_f64<0x112, 0xf>(v);
  v += dpp_f64<0x114, 0xf>(v);
  v += dpp_f64<0x118, 0xf>(v);
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  uint64_t r = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, 16 * q + 15);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), 16 * q + 15);
    if (g == q) r = ((uint64_t)hi << 32) | lo;
  }
  return __builtin_bit_cast(double, r);
}

// threadIdx.x; with kOpaque, opaque to the optimiser. A persistent kernel's
// tile loop takes it opaque once per tile, so that per-lane values computed
// from it stay inside the loop: hoisted out of it, they would be held in
// registers (or spilled) through the walk.
template <bool kOpaque>
__device__ __forceinline__ int tile_tid() {
  int t = threadIdx.x;
  if constexpr (kOpaque) asm volatile("" : "+v"(t));
  return t;
}

struct TileInfo {
  int tile, I, J, e0, ne, ktot, natI, natJ;
  int32_t firstI, firstJ;
  int64_t out_base;
  const TileDesc* tp;
};

// Tile of slot `slot` of the tile order and its band geometry in LDS.
// Workgroup b runs slots b, b + gridDim.x, ...: one slot with a grid of one
// workgroup per tile, many with a persistent grid of the resident
// workgroups. The hardware deals consecutive workgroups round-robin over the
// 8 XCDs, so the tiles running at any moment are a window of the tile order
// spread over all XCDs: the A and B strips of that window are read by all of
// them from the shared last-level cache. (Giving each XCD its own contiguous
// range of the order, so that each L2 sees a compact range, measured 20.1 ms
// on the 6,144-atom box against 18.5 ms.) Ends with the LDS visible to all
// threads.
__device__ __forceinline__ TileInfo tile_setup(const TileDesc* __restrict__ tiles,
                                               const int32_t* __restrict__ perm,
                                               const Band* __restrict__ bands, TileSmem& sm,
                                               int32_t slot, int tid) {
  TileInfo t;
  t.tile = perm[slot];
  t.tp = tiles + t.tile;
  t.I = t.tp->I;
  t.J = t.tp->J;
  t.e0 = t.tp->e0;
  t.ne = t.tp->ne;
  t.ktot = t.tp->ktot;
  t.out_base = t.tp->out_base;
  t.natI = bands[t.I].natoms;
  t.natJ = bands[t.J].natoms;
  t.firstI = bands[t.I].first;
  t.firstJ = bands[t.J].first;
  if (tid <= kBandAtoms) {
    sm.rs_r[tid] = bands[t.I].rs[min(tid, t.natI)];
    sm.rs_c[tid] = bands[t.J].rs[min(tid, t.natJ)];
  }
  if (tid < kBandAtoms) sm.keptw[tid] = 0;
  if (tid == 0) {
    sm.max_rows = 0;
    sm.max_cols = 0;
    sm.zero = 0.0;
    sm.n_exact = 0;
  }
  __syncthreads();
  if (tid < kBandAtoms) {
    uint32_t g = 0;
    if (tid < t.natI && sm.rs_r[tid + 1] > sm.rs_r[tid])
      for (int q = sm.rs_r[tid] / 4; q <= (sm.rs_r[tid + 1] - 1) / 4; ++q) g |= 1u << q;
    sm.pg_r[tid] = (uint16_t)g;
    g = 0;
    if (tid < t.natJ && sm.rs_c[tid + 1] > sm.rs_c[tid])
      for (int q = sm.rs_c[tid] / 4; q <= (sm.rs_c[tid + 1] - 1) / 4; ++q) g |= 1u << q;
    sm.pg_c[tid] = (uint16_t)g;
    if (tid < t.natI) atomicMax(&sm.max_rows, sm.rs_r[tid + 1] - sm.rs_r[tid]);
    if (tid < t.natJ) atomicMax(&sm.max_cols, sm.rs_c[tid + 1] - sm.rs_c[tid]);
  }
  __syncthreads();
  return t;
}

// Mask bit of the band atom holding this lane's operand rows (A, bits 0..15)
// and columns (B, bits 16..31) for the quadrant of wave (wr, wc): rows
// 32wr + 16h + (lane & 15), cols 32wc + 16q + (lane & 15); 0 past the band.
__device__ __forceinline__ void lane_atom_bits(const TileInfo& t, const TileSmem& sm, int wr,
                                               int wc, int lane, uint32_t rbit[2],
                                               uint32_t cbit[2]) {
  rbit[0] = rbit[1] = cbit[0] = cbit[1] = 0u;
  for (int a = 0; a < t.natI; ++a)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = 32 * wr + 16 * h + (lane & 15);
      if (r >= sm.rs_r[a] && r < sm.rs_r[a + 1]) rbit[h] = 1u << a;
    }
  for (int a = 0; a < t.natJ; ++a)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int c = 32 * wc + 16 * q + (lane & 15);
      if (c >= sm.rs_c[a] && c < sm.rs_c[a + 1]) cbit[q] = 1u << (16 + a);
    }
}

// Row / column 4-groups touched by an entry with atom masks rmask, cmask:
// rows in the low 16 bits, columns in the high 16 bits.
__device__ __forceinline__ uint32_t entry_groups(const TileSmem& sm, uint32_t rmask,
                                                 uint32_t cmask) {
  uint32_t gr = 0, gc = 0;
  for (int p = 0; p < kBandAtoms; ++p) {
    if ((rmask >> p) & 1u) gr |= sm.pg_r[p];
    if ((cmask >> p) & 1u) gc |= sm.pg_c[p];
  }
  return gr | (gc << 16);
}

// 4-bit mask of the live 16 x 16 blocks of a wave's 32 x 32 quadrant in a
// step whose row / column 4-groups are g (rows low half, columns high half):
// bit 2h+q = row half h and column half q both hold a live group.
__device__ __forceinline__ uint32_t quadrant_blocks(uint32_t g, uint32_t gsh_r, uint32_t gsh_c) {
  const uint32_t rg = (g & 0xFFFFu) >> gsh_r, cg = (g >> 16) >> gsh_c;
  const uint32_t r0 = (rg & 0xFu) ? 1u : 0u, r1 = (rg & 0xF0u) ? 1u : 0u;
  const uint32_t c0 = (cg & 0xFu) ? 1u : 0u, c1 = (cg & 0xF0u) ? 1u : 0u;
  return (r0 & c0) | ((r0 & c1) << 1) | ((r1 & c0) << 2) | ((r1 & c1) << 3);
}

// Epilogue for a workgroup of NT threads (NT / 64 waves; tid = threadIdx.x).
// Ct holds the tile's 64 x 64 accumulated products (row pitch kPitch) and is
// visible to all threads. Each structural block (i,j) of the tile becomes
// alpha * (A*B)_ij + beta * Cold_ij (Cold found by binary search in its CSR
// row; each operation rounded on its own, no contraction), is written to the
// C pool at out_base + its offset (blocks back to back in row-major order of
// atom pairs, each column-major) and back into Ct. Its sum of squares is then
// taken by one thread as one sequential FMA chain over the block's values in
// column-major order, s = fma(x, x, s) from s = +0, the same order as the
// prefilter's block norms; the block is kept iff that sum is >= eps2.
// kept_out[tile*16 + i] gets the kept columns of band atom i. With a
// slot_counter (a persistent grid), thread 0 takes the next value of that
// counter (the index of the workgroup's next slot in its sequence) and
// stores it to *next_slot (LDS), visible after a barrier.
template <int NT>
__device__ __forceinline__ void tile_epilogue(
    double* Ct, const TileInfo& t, TileSmem& sm, double alpha, double beta, double eps2,
    const int64_t* __restrict__ old_row_p, const int32_t* __restrict__ old_col,
    const int64_t* __restrict__ old_off, const double* __restrict__ old_data,
    double* __restrict__ c_data, uint16_t* __restrict__ kept_out, int tid,
    int32_t* __restrict__ slot_counter = nullptr, int32_t* next_slot = nullptr) {
#pragma clang fp contract(off)
  static_assert(NT % 64 == 0 && NT >= 64 && NT <= 256, "epilogue threads");
  constexpr int kWaves = NT / 64;
  const int lane = tid & 63;
  // a persistent grid's next tile slot, claimed now and published at the
  // end, so that the atomic's round trip overlaps the epilogue
  int32_t claimed = 0;
  if (slot_counter != nullptr && tid == 0) claimed = atomicAdd(slot_counter, 1);
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // structural pairs (row-major over band positions: id = 16 ip + jp) and
  // their offsets, NT ids per pass
  int cbase = 0;
  int64_t sbase = 0;
  for (int id0 = 0; id0 < kBandAtoms * kBandAtoms; id0 += NT) {
    const int id = id0 + tid;
    const int ip = id >> 4, jp = id & 15;
    const bool present = ip < t.natI && jp < t.natJ && ((t.tp->pattern[ip] >> jp) & 1);
    const int64_t sz =
        present ? (int64_t)(sm.rs_r[ip + 1] - sm.rs_r[ip]) * (sm.rs_c[jp + 1] - sm.rs_c[jp]) : 0;
    // inclusive prefix counts by ballot, sizes by a DPP scan (a tile holds
    // at most 64 x 64 values, so 32 bits suffice)
    const uint64_t pres = __ballot(present);
    const int cnt = (int)__builtin_amdgcn_mbcnt_hi(
                        (uint32_t)(pres >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)pres, 0u)) +
                    (present ? 1 : 0);
    const int64_t csz = wave_scan_dpp((int)sz);
    if (lane == 63) {
      sm.wave_cnt[wave] = cnt;
      sm.wave_sz[wave] = csz;
    }
    __syncthreads();
    int cw = cbase;
    int64_t sw = sbase;
    for (int w = 0; w < wave; ++w) {
      cw += sm.wave_cnt[w];
      sw += sm.wave_sz[w];
    }
    if (present) {
      sm.pair_id[cw + cnt - 1] = id;
      sm.pair_off[cw + cnt - 1] = sw + csz - sz;
    }
    for (int w = 0; w < kWaves; ++w) {
      cbase += sm.wave_cnt[w];
      sbase += sm.wave_sz[w];
    }
    __syncthreads();
  }
  const int npairs = cbase;

  // values: one 16-lane row of the wave per block, four blocks per wave at
  // a time, lanes over the block's elements
  const int grp = lane >> 4, gl = lane & 15;
  for (int pi0 = wave; pi0 < npairs; pi0 += 4 * kWaves) {
    const int pi = pi0 + grp * kWaves;
    int pr = 0, pc = 0, m = 1, len = 0;
    int64_t base = 0;
    if (pi < npairs) {
      const int id = sm.pair_id[pi];
      pr = id >> 4;
      pc = id & 15;
      m = sm.rs_r[pr + 1] - sm.rs_r[pr];
      len = m * (sm.rs_c[pc + 1] - sm.rs_c[pc]);
      base = t.out_base + sm.pair_off[pi];
    }
    const double* old = nullptr;
    if (beta != 0.0 && old_row_p != nullptr && pi < npairs) {
      const int32_t gi = t.firstI + pr, gj = t.firstJ + pc;
      int64_t lo = old_row_p[gi], hi = old_row_p[gi + 1] - 1;
      int64_t found = -1;
      while (lo <= hi) {
        const int64_t mid = (lo + hi) >> 1;
        const int32_t cc = old_col[mid];
        if (cc == gj) {
          found = mid;
          break;
        }
        if (cc < gj) lo = mid + 1;
        else hi = mid - 1;
      }
      if (found >= 0) old = old_data + old_off[found];
    }
    // element e = r + m * cc of the column-major block, stepped by 16
    // without a division per element (a zero-row atom has len 0)
    const int mm = m > 0 ? m : 1;
    const int q16 = 16 / mm, r16 = 16 - q16 * mm;
    int cc = gl / mm, r = gl - (gl / mm) * mm;
    double part = 0.0;  // this lane's squares, one FMA each
    for (int e = gl; e < len; e += 16) {
      double* x = &Ct[(sm.rs_r[pr] + r) * kPitch + sm.rs_c[pc] + cc];
      double v = alpha * *x;
      if (old) {
        const double bo = beta * old[e];
        v = v + bo;
      }
      c_data[base + e] = v;
      *x = v;
      part = __builtin_fma(v, v, part);
      r += r16;
      cc += q16;
      if (r >= mm) {
        r -= mm;
        cc += 1;
      }
    }
    // Post-filter decision from a cheap sum when it is certain. Every term
    // is a square, so both this sum (lane partials of at most ceil(len/16)
    // FMAs, then at most four additions) and the contract's sequential FMA
    // chain over the len values are within gamma_k * S of the exact sum S,
    // gamma_k = k u / (1 - k u), u = 2^-53. Their difference is below
    // 2 (len + ceil(len/16) + 8) u * sum; when the sum lies farther than
    // that from eps2, the chain is on the same side of eps2, and the keep
    // bit is exactly the chain's. Other blocks (near eps2, a sum that is
    // tiny, zero, infinite or NaN) get the chain itself below.
    const double sum = row_totals_dpp(part, grp);
    const double margin = sum * ((double)(len + (len + 15) / 16 + 8) * 0x1p-52);
    const bool certain = sum > 0x1p-930 && sum <= 0x1p+1000 && fabs(sum - eps2) > margin;
    if (gl == 0 && pi < npairs) {
      if (!certain) sm.exact_id[atomicAdd(&sm.n_exact, 1)] = (int16_t)pi;
      else if (sum >= eps2) atomicOr(&sm.keptw[pr], 1u << pc);
    }
  }
  __syncthreads();
  // norms of the undecided blocks: one thread per block, one FMA chain over
  // its values in column-major order (the contract's order). Every thread
  // walks the tile's largest block shape, column by column in steps of
  // eight rows; a slot outside its own block reads sm.zero, and
  // fma(0, 0, s) = s leaves the chain unchanged. The eight reads of a step
  // are issued before its FMAs. Waves without a block skip.
  const int n_exact = __builtin_amdgcn_readfirstlane(sm.n_exact);
  if (n_exact > 0) {
    const int mmax = __builtin_amdgcn_readfirstlane(sm.max_rows);
    const int nmax = __builtin_amdgcn_readfirstlane(sm.max_cols);
    for (int pi0 = 0; pi0 < n_exact; pi0 += NT) {
      if (pi0 + 64 * wave >= n_exact) continue;
      const int pi = pi0 + tid;
      int pr = 0, pc = 0, m = 0, n = 0, base = 0;
      if (pi < n_exact) {
        const int id = sm.pair_id[sm.exact_id[pi]];
        pr = id >> 4;
        pc = id & 15;
        m = sm.rs_r[pr + 1] - sm.rs_r[pr];
        n = sm.rs_c[pc + 1] - sm.rs_c[pc];
        base = sm.rs_r[pr] * kPitch + sm.rs_c[pc];
      }
      double s = 0.0;
      for (int cc = 0; cc < nmax; ++cc) {
        const int mc = cc < n ? m : 0;  // rows of this column in this thread's block
        const double* col = Ct + base + cc;
        for (int r0 = 0; r0 < mmax; r0 += 8) {
          double x[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) x[u] = *(r0 + u < mc ? col + (r0 + u) * kPitch : &sm.zero);
#pragma unroll
          for (int u = 0; u < 8; ++u) s = __builtin_fma(x[u], x[u], s);
        }
      }
      if (pi < n_exact && s >= eps2) atomicOr(&sm.keptw[pr], 1u << pc);
    }
  }
  __syncthreads();
  if (tid < kBandAtoms) kept_out[(int64_t)t.tile * kBandAtoms + tid] = (uint16_t)sm.keptw[tid];
  if (slot_counter != nullptr && tid == 0) *next_slot = claimed;
}

}  // namespace lamb
