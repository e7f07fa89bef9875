// Symbolic phase on the device: block norms -> strips -> packed panels ->
// tiles and their (A strip, B strip) entries, and the C index after the
// numeric kernel. Every step is deterministic: entries of a tile are in
// ascending K atom order and C blocks in ascending column order.
#include <cstdlib>
#include <cstring>

#include "common.h"
#include "device_util.h"
#include "symbolic.h"

namespace lamb {

LAMB_GUARD_UNIT(symbolic)


__global__ void fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
  LAMB_GUARD;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

static hipError_t fill32(hipStream_t s, int32_t* p, int64_t n, int32_t v) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// A strips. The blocks of a band's rows, merged by column atom k (ties by
// row position), are the band's strips in order. Each block finds its merged
// position with one binary search per row of its band.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t lower_bound_i32(const int32_t* a, int64_t lo, int64_t hi,
                                                   int32_t key) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void a_merge_rank(int64_t nblk, const int32_t* __restrict__ rows,
                             const int32_t* __restrict__ col, const int64_t* __restrict__ row_p,
                             const int32_t* __restrict__ band_of,
                             const int32_t* __restrict__ pos_in_band,
                             const Band* __restrict__ bands, int64_t* __restrict__ merged) {
  LAMB_GUARD;
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  const int32_t i = rows[b];
  const int32_t k = col[b];
  const int32_t I = band_of[i];
  const int p = pos_in_band[i];
  const int32_t first = bands[I].first;
  const int nat = bands[I].natoms;
  int64_t pos = row_p[first];
  for (int q = 0; q < nat; ++q) {
    const int64_t lo = row_p[first + q], hi = row_p[first + q + 1];
    if (q == p) {
      pos += b - lo;
    } else {
      // rows before p sort ahead on ties
      int64_t c = lower_bound_i32(col, lo, hi, q < p ? k + 1 : k);
      pos += c - lo;
    }
  }
  merged[pos] = b;
}

// strip-start flags over a sequence of blocks: A (merged order) or B (CSR
// order). A strip starts where the band changes or the K atom changes.
__global__ void strip_flags(int64_t nblk, int a_side, const int64_t* __restrict__ merged,
                            const int32_t* __restrict__ rows, const int32_t* __restrict__ col,
                            const int32_t* __restrict__ band_of, int64_t* __restrict__ flag) {
  LAMB_GUARD;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nblk) return;
  int64_t f;
  if (a_side) {
    const int64_t b = merged[t];
    if (t == 0) {
      f = 1;
    } else {
      const int64_t pb = merged[t - 1];
      f = (band_of[rows[b]] != band_of[rows[pb]] || col[b] != col[pb]) ? 1 : 0;
    }
  } else {
    f = (t == 0 || rows[t] != rows[t - 1] || band_of[col[t]] != band_of[col[t - 1]]) ? 1 : 0;
  }
  flag[t] = f;
}

__global__ void strip_init(int64_t n, Strip* __restrict__ s) {
  LAMB_GUARD;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  s[i].mask = 0;
  s[i].pad0 = 0;
  s[i].pad1 = 0;
  s[i].min_nsq = __builtin_inf();
}

// scan[t] = strip index + 1 of the block at sequence slot t (inclusive count)
__global__ void strip_fill(int64_t nblk, int a_side, const int64_t* __restrict__ merged,
                           const int64_t* __restrict__ excl, const int64_t* __restrict__ flag,
                           const int32_t* __restrict__ rows, const int32_t* __restrict__ col,
                           const int64_t* __restrict__ off, const double* __restrict__ nsq,
                           const int32_t* __restrict__ band_of,
                           const int32_t* __restrict__ pos_in_band, Strip* __restrict__ strips) {
  LAMB_GUARD;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nblk) return;
  const int64_t b = a_side ? merged[t] : t;
  const int64_t s = excl[t] + flag[t] - 1;
  const int32_t i = rows[b], j = col[b];
  // A strip (I, k): position of row atom i; B strip (k, J): position of col atom j
  const int p = a_side ? pos_in_band[i] : pos_in_band[j];
  Strip* S = strips + s;
  if (flag[t]) {
    S->k = a_side ? j : i;
    S->band = a_side ? band_of[i] : band_of[j];
  }
  S->off[p] = off[b];
  S->nsq[p] = nsq[b];
  atomicOr(reinterpret_cast<uint32_t*>(&S->mask), 1u << p);
  atomicMin(reinterpret_cast<unsigned long long*>(&S->min_nsq),
            (unsigned long long)__double_as_longlong(nsq[b]));
}

// per-strip panel size: K columns rounded up to a multiple of 8 (one BF16
// slice chunk), times 64; and the strip's head
__global__ void strip_panel_size(int64_t n, const Strip* __restrict__ s,
                                 const int32_t* __restrict__ blk, const Band* __restrict__ bands,
                                 int64_t* __restrict__ sz, StripHead* __restrict__ head) {
  LAMB_GUARD;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  sz[i] = 64 * (int64_t)((blk[s[i].k] + 7) & ~7);
  const int32_t band = s[i].band;
  const uint32_t m = s[i].mask;
  int rows = 0;
  for (int p = 0; p < kBandAtoms; ++p)
    if ((m >> p) & 1u) rows += bands[band].rs[p + 1] - bands[band].rs[p];
  StripHead h;
  h.band = band;
  h.mask = (uint16_t)m;
  h.rows = (uint16_t)rows;
  h.min_nsq = s[i].min_nsq;
  head[i] = h;
}

// ptr[g] = first strip of group g: strip index at the group's first slot
__global__ void group_ptr(int64_t ngroups, const int64_t* __restrict__ slot_of_group,
                          const int64_t* __restrict__ excl, int64_t nslots, int64_t nstrips,
                          int64_t* __restrict__ ptr) {
  LAMB_GUARD;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g > ngroups) return;
  if (g == ngroups) {
    ptr[g] = nstrips;
    return;
  }
  const int64_t t = slot_of_group[g];
  ptr[g] = t < nslots ? excl[t] : nstrips;
}

// slot of band I's first block in the merged A order = row_p[first atom]
__global__ void band_first_slot(int32_t nbands, const Band* __restrict__ bands,
                                const int64_t* __restrict__ row_p, int64_t* __restrict__ out) {
  LAMB_GUARD;
  const int32_t I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I < nbands) out[I] = row_p[bands[I].first];
}

// ---------------------------------------------------------------------------
// Tiles and entries. One wave per (row band I, segment g) walks its share of
// band I's A strips in K order; the lanes take the B strips of that K atom
// (distinct column bands J). Counters per (I, g, J) live in dense scratch.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int rows_of_mask(const int16_t* rs, uint32_t m) {
  int r = 0;
  for (int p = 0; p < kBandAtoms; ++p)
    if ((m >> p) & 1u) r += rs[p + 1] - rs[p];
  return r;
}

// A wave's walk over its A strips [lo, hi), two strips ahead: k2 is the K
// atom of the strip after next; k1 / kd1 / b01 / b11 (its block size and B
// strip range) belong to the next strip, so that the loads for a strip are
// in flight while the one before is processed.
struct StripWalk {
  const Strip* as;
  const int32_t* blk;
  const int64_t* b_ptr;
  int64_t hi;
  int32_t k1 = 0, k2 = 0, kd1 = 0;
  int64_t b01 = 0, b11 = 0;
  __device__ __forceinline__ StripWalk(const Strip* a, const int32_t* bk, const int64_t* bp,
                                       int64_t lo, int64_t h)
      : as(a), blk(bk), b_ptr(bp), hi(h) {
    if (lo < hi) {
      k1 = as[lo].k;
      kd1 = blk[k1];
      b01 = b_ptr[k1];
      b11 = b_ptr[k1 + 1];
    }
    if (lo + 1 < hi) k2 = as[lo + 1].k;
  }
  // called while strip sa is processed: make sa + 1 the next strip
  __device__ __forceinline__ void advance(int64_t sa) {
    if (sa + 1 < hi) {
      k1 = k2;
      kd1 = blk[k1];
      b01 = b_ptr[k1];
      b11 = b_ptr[k1 + 1];
    }
    if (sa + 2 < hi) k2 = as[sa + 2].k;
  }
};

// Calls emit(rm, cm) for every entry of strip pair (SA, SB) under the
// prefilter nsq(A_ik) * nsq(B_kj) >= eps2; returns the dropped pair count.
// The heads decide the common case (no pair cut); the strips' per-atom norms
// are read only when the smallest norms could be cut.
template <typename F>
__device__ __forceinline__ int pair_entries(const StripHead& HA, const StripHead& HB,
                                            const Strip* SA, const Strip* SB, double eps2,
                                            F&& emit) {
  const uint32_t am = HA.mask, bm = HB.mask;
  if (HA.min_nsq * HB.min_nsq >= eps2) {
    emit(am, bm);
    return 0;
  }
  int dropped = 0;
  for (int q = 0; q < kBandAtoms; ++q) {
    if (!((bm >> q) & 1u)) continue;
    uint32_t r = 0;
    for (int p = 0; p < kBandAtoms; ++p)
      if (((am >> p) & 1u) && SA->nsq[p] * SB->nsq[q] >= eps2) r |= 1u << p;
    dropped += __builtin_popcount(am) - __builtin_popcount(r);
  }
  // columns grouped by their kept-row set; each group is one entry
  uint32_t remaining = bm;
  while (remaining) {
    const int q = __builtin_ctz(remaining);
    uint32_t r = 0;
    for (int p = 0; p < kBandAtoms; ++p)
      if (((am >> p) & 1u) && SA->nsq[p] * SB->nsq[q] >= eps2) r |= 1u << p;
    uint32_t cm = 0;
    for (int q2 = q; q2 < kBandAtoms; ++q2) {
      if (!((remaining >> q2) & 1u)) continue;
      uint32_t r2 = 0;
      for (int p = 0; p < kBandAtoms; ++p)
        if (((am >> p) & 1u) && SA->nsq[p] * SB->nsq[q2] >= eps2) r2 |= 1u << p;
      if (r2 == r) cm |= 1u << q2;
    }
    remaining &= ~cm;
    if (r) emit(r, cm);
  }
  return dropped;
}

// Output mask (masked multiplies): mpat[(I * nbands + J) * 8 + p / 2] holds,
// in its 16-bit half p % 2, the column atoms q of band J that row atom p of
// band I may produce. An entry (rm, cm) is cut to the rows with a masked
// column in cm and to those columns; pairs[p] gets row p's masked columns.
// Returns false when nothing is left. Without a mask the entry is unchanged.
__device__ __forceinline__ bool mask_entry(const uint32_t* __restrict__ mpat, int32_t nbands,
                                           int32_t I, int32_t J, uint32_t& rm, uint32_t& cm,
                                           uint32_t pairs[kBandAtoms]) {
  if (!mpat) {
    for (int p = 0; p < kBandAtoms; ++p) pairs[p] = ((rm >> p) & 1u) ? cm : 0u;
    return true;
  }
  const uint32_t* m = mpat + ((int64_t)I * nbands + J) * (kBandAtoms / 2);
  uint32_t r2 = 0, c2 = 0;
  for (int p = 0; p < kBandAtoms; ++p) {
    uint32_t row = 0;
    if ((rm >> p) & 1u) row = (m[p >> 1] >> (16 * (p & 1))) & 0xFFFFu & cm;
    pairs[p] = row;
    if (row) {
      r2 |= 1u << p;
      c2 |= row;
    }
  }
  rm = r2;
  cm = c2;
  return r2 != 0;
}

// mpat bits of every stored block of the mask matrix
__global__ void mask_pattern(int64_t nblk, const int32_t* __restrict__ rows,
                             const int32_t* __restrict__ col, const int32_t* __restrict__ band_of,
                             const int32_t* __restrict__ pos_in_band, int32_t nbands,
                             uint32_t* __restrict__ mpat) {
  LAMB_GUARD;
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  const int32_t i = rows[b], j = col[b];
  const int32_t I = band_of[i], J = band_of[j], p = pos_in_band[i], q = pos_in_band[j];
  atomicOr(&mpat[((int64_t)I * nbands + J) * (kBandAtoms / 2) + (p >> 1)],
           (1u << q) << (16 * (p & 1)));
}

// kLds = false: one wave per (band I, segment g), pattern words and
// counters updated in global memory. kLds = true: one workgroup of nseg waves
// per band I (wave g = segment g); the band's nbands x 8 pattern words and
// nseg x nbands entry / K-column counters live in LDS (dynamic shared memory,
// sym_count_lds(nbands, nseg) bytes) and are stored to global memory at the
// end, in place of global atomics.
template <bool kLds>
__global__ __launch_bounds__(kLds ? 512 : 64) void sym_count(
    int32_t nbands, int nseg, const Band* __restrict__ bands, const int32_t* __restrict__ blk,
    const Strip* __restrict__ as, const StripHead* __restrict__ ah,
    const int64_t* __restrict__ a_ptr, const Strip* __restrict__ bs,
    const StripHead* __restrict__ bh, const int64_t* __restrict__ b_ptr, double eps2,
    const int64_t* __restrict__ old_row_p, const int32_t* __restrict__ old_col,
    const int32_t* __restrict__ band_of, const int32_t* __restrict__ pos_in_band,
    int32_t* __restrict__ cnt, int32_t* __restrict__ ksum, uint32_t* __restrict__ pat,
    unsigned long long* __restrict__ stats, const uint32_t* __restrict__ mpat) {
  LAMB_GUARD;
  extern __shared__ uint32_t spat[];
  const int32_t I = kLds ? (int32_t)blockIdx.x : (int32_t)(blockIdx.x / nseg);
  const int g = kLds ? (int)(threadIdx.x >> 6) : (int)(blockIdx.x % nseg);
  if (I >= nbands) return;
  const int lane = threadIdx.x & 63;
  const int64_t sa0 = a_ptr[I], sa1 = a_ptr[I + 1];
  const int64_t len = sa1 - sa0;
  const int64_t lo = sa0 + len * g / nseg, hi = sa0 + len * (g + 1) / nseg;
  const int64_t nwords = (int64_t)nbands * (kBandAtoms / 2);
  int32_t* scnt = reinterpret_cast<int32_t*>(spat + nwords);  // [nseg][nbands]
  int32_t* sks = scnt + (int64_t)nseg * nbands;                // [nseg][nbands]
  int32_t* c = kLds ? scnt + (int64_t)g * nbands : cnt + ((int64_t)I * nseg + g) * nbands;
  int32_t* ks = kLds ? sks + (int64_t)g * nbands : ksum + ((int64_t)I * nseg + g) * nbands;
  uint32_t* pt = pat + (int64_t)I * nbands * (kBandAtoms / 2);
  __shared__ int16_t rsI[kBandAtoms + 1];
  if (threadIdx.x <= kBandAtoms) rsI[threadIdx.x] = bands[I].rs[min((int)threadIdx.x, bands[I].natoms)];
  if (kLds)
    for (int64_t x = threadIdx.x; x < nwords + 2 * (int64_t)nseg * nbands; x += blockDim.x)
      spat[x] = 0u;
  __syncthreads();
  auto or_word = [&](int64_t x, uint32_t v) {
    if constexpr (kLds)
      atomicOr(&spat[x], v);
    else
      atomicOr(&pt[x], v);
  };
  unsigned long long kept = 0, dropped = 0, partial = 0;
  double flops = 0.0;
  StripWalk walk(as, blk, b_ptr, lo, hi);
  for (int64_t sa = lo; sa < hi; ++sa) {
    const Strip* SA = as + sa;
    const StripHead HA = ah[sa];
    const int32_t k = walk.k1;
    const int kd = walk.kd1;
    const int64_t b0 = walk.b01, b1 = walk.b11;
    walk.advance(sa);
    for (int64_t sb = b0 + lane; sb < b1; sb += 64) {
      const StripHead HB = bh[sb];
      const int32_t J = HB.band;
      int n = 0;
      dropped += pair_entries(HA, HB, SA, bs + sb, eps2, [&](uint32_t rm, uint32_t cm) {
        uint32_t pairs[kBandAtoms];
        if (!mask_entry(mpat, nbands, I, J, rm, cm, pairs)) return;
        // a zero-size K atom adds to the C pattern but holds no entry
        if (kd > 0) ++n;
        // entries whose atoms the prefilter or the mask cut (sym_fill flags
        // them with kEntryPartial)
        if (kd > 0 && (rm != HA.mask || cm != HB.mask)) ++partial;
        kept += (unsigned long long)__builtin_popcount(rm) * __builtin_popcount(cm);
        // rows of rm in band I, columns of cm in band J
        int cr = HB.rows;
        if (cm != HB.mask) {
          cr = 0;
          for (int q = 0; q < kBandAtoms; ++q)
            if ((cm >> q) & 1u) cr += bands[J].rs[q + 1] - bands[J].rs[q];
        }
        const int rr = rm == HA.mask ? (int)HA.rows : rows_of_mask(rsI, rm);
        flops += 2.0 * rr * kd * cr;
        // one atomic per pattern word (two row atoms)
        for (int wd = 0; wd < kBandAtoms / 2; ++wd) {
          const uint32_t v = pairs[2 * wd] | (pairs[2 * wd + 1] << 16);
          if (v) or_word((int64_t)J * (kBandAtoms / 2) + wd, v);
        }
      });
      if (n) {
        atomicAdd(&c[J], n);
        atomicAdd(&ks[J], n * kd);
      }
    }
  }
  if (g == 0 && old_row_p != nullptr) {
    const Band& bI = bands[I];
    for (int p = 0; p < bI.natoms; ++p) {
      const int32_t i = bI.first + p;
      for (int64_t b = old_row_p[i] + lane; b < old_row_p[i + 1]; b += 64) {
        const int32_t j = old_col[b];
        const int32_t J = band_of[j];
        or_word((int64_t)J * (kBandAtoms / 2) + (p >> 1), (1u << pos_in_band[j]) << (16 * (p & 1)));
      }
    }
  }
  // stats: [0] kept triples, [1] dropped, [2] flops (as integer: exact below 2^63)
  for (int o = 32; o > 0; o >>= 1) {
    kept += __shfl_xor(kept, o, 64);
    dropped += __shfl_xor(dropped, o, 64);
    partial += __shfl_xor(partial, o, 64);
    flops += __shfl_xor(flops, o, 64);
  }
  if (lane == 0) {
    atomicAdd(&stats[0], kept);
    atomicAdd(&stats[1], dropped);
    atomicAdd(&stats[2], (unsigned long long)flops);
    if (partial) atomicAdd(&stats[3], partial);
  }
  if (kLds) {
    __syncthreads();
    for (int64_t x = threadIdx.x; x < nwords; x += blockDim.x) pt[x] = spat[x];
    int32_t* gc = cnt + (int64_t)I * nseg * nbands;
    int32_t* gk = ksum + (int64_t)I * nseg * nbands;
    for (int64_t x = threadIdx.x; x < (int64_t)nseg * nbands; x += blockDim.x) {
      gc[x] = scnt[x];
      gk[x] = sks[x];
    }
  }
}

// dynamic LDS of sym_count<true> / sym_fill<true>
static size_t sym_count_lds(int32_t nbands, int nseg) {
  return sizeof(uint32_t) * (size_t)nbands * (kBandAtoms / 2) +
         2 * sizeof(int32_t) * (size_t)nseg * (size_t)nbands;
}
static size_t sym_fill_lds(int32_t nbands, int nseg) {
  return 2 * sizeof(int32_t) * (size_t)nseg * (size_t)nbands;
}
constexpr size_t kSymLdsMax = 96 * 1024;
// LAMB_HIP_SYM_LDS=0 takes the global-memory forms of sym_count / sym_fill
// whatever the band count (they serve more than kSymLdsMax bytes per band)
static bool sym_lds_allowed() {
  const char* v = getenv("LAMB_HIP_SYM_LDS");
  return !(v && strcmp(v, "0") == 0);
}

// Per (row band I, chunk of 64 column bands): tiles present, their entry
// counts, C sizes and C blocks (lanes over J). One wave per (I, chunk); the
// totals of (I, c) go to slot I * nchunk + c of each of the four arrays.
__global__ __launch_bounds__(64) void sym_chunk_totals(
    int32_t nbands, int nseg, int32_t nchunk, const Band* __restrict__ bands,
    const int32_t* __restrict__ cnt, const uint32_t* __restrict__ pat,
    int64_t* __restrict__ ch_tiles, int64_t* __restrict__ ch_entries,
    int64_t* __restrict__ ch_csize, int64_t* __restrict__ ch_cblocks) {
  LAMB_GUARD;
  const int32_t I = (int32_t)(blockIdx.x / nchunk), c = (int32_t)(blockIdx.x % nchunk);
  if (I >= nbands) return;
  const int lane = threadIdx.x;
  const Band& bI = bands[I];
  int64_t nt = 0, ne = 0, cs = 0, cb = 0;
  const int32_t J = c * 64 + lane;
  if (J < nbands) {
    int64_t e = 0;
    for (int g = 0; g < nseg; ++g) e += cnt[((int64_t)I * nseg + g) * nbands + J];
    const uint32_t* pt = pat + ((int64_t)I * nbands + J) * (kBandAtoms / 2);
    bool any = false;
    const Band& bJ = bands[J];
    for (int p = 0; p < bI.natoms; ++p) {
      const uint32_t m = (pt[p >> 1] >> (16 * (p & 1))) & 0xFFFFu;
      if (!m) continue;
      any = true;
      const int mr = bI.rs[p + 1] - bI.rs[p];
      for (int q = 0; q < bJ.natoms; ++q)
        if ((m >> q) & 1u) {
          cs += (int64_t)mr * (bJ.rs[q + 1] - bJ.rs[q]);
          ++cb;
        }
    }
    if (any || e) {
      nt = 1;
      ne = e;
    }
  }
  nt = wave_sum_i64(nt);
  ne = wave_sum_i64(ne);
  cs = wave_sum_i64(cs);
  cb = wave_sum_i64(cb);
  if (lane == 0) {
    const int64_t x = (int64_t)I * nchunk + c;
    ch_tiles[x] = nt;
    ch_entries[x] = ne;
    ch_csize[x] = cs;
    ch_cblocks[x] = cb;
  }
}

// tile_base[I] = first tile of band I (the scanned tile counts of its first
// chunk); tile_base[nbands] = all tiles
__global__ void band_tile_base(int32_t nbands, int32_t nchunk, const int64_t* __restrict__ scanned,
                               int64_t* __restrict__ tile_base) {
  LAMB_GUARD;
  const int32_t I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I <= nbands) tile_base[I] = scanned[(int64_t)I * nchunk];
}

// Tile descriptors (J ascending within a band); turns cnt/ksum into the
// per-(I, g, J) entry cursors and K-column starts used by sym_fill. One wave
// per (I, chunk of 64 column bands), starting at the chunk's scanned tile,
// entry and C-size offsets.
__global__ __launch_bounds__(64) void sym_tiles(
    int32_t nbands, int nseg, int32_t nchunk, const Band* __restrict__ bands,
    int32_t* __restrict__ cnt, int32_t* __restrict__ ksum, const uint32_t* __restrict__ pat,
    const int64_t* __restrict__ tile_base, const int64_t* __restrict__ entry_base,
    const int64_t* __restrict__ csize_base, TileDesc* __restrict__ tiles) {
  LAMB_GUARD;
  const int32_t I = (int32_t)(blockIdx.x / nchunk), c = (int32_t)(blockIdx.x % nchunk);
  if (I >= nbands) return;
  const int lane = threadIdx.x;
  const Band& bI = bands[I];
  const int64_t x0 = (int64_t)I * nchunk + c;
  int64_t t_run = tile_base[x0], e_run = entry_base[x0], c_run = csize_base[x0];
  {
    const int32_t J0 = c * 64;
    const int32_t J = J0 + lane;
    int64_t e = 0, k = 0, cs = 0;
    bool any = false;
    uint32_t pw[kBandAtoms / 2] = {};
    if (J < nbands) {
      for (int g = 0; g < nseg; ++g) {
        e += cnt[((int64_t)I * nseg + g) * nbands + J];
        k += ksum[((int64_t)I * nseg + g) * nbands + J];
      }
      const uint32_t* pt = pat + ((int64_t)I * nbands + J) * (kBandAtoms / 2);
      const Band& bJ = bands[J];
      for (int w = 0; w < kBandAtoms / 2; ++w) pw[w] = pt[w];
      for (int p = 0; p < bI.natoms; ++p) {
        const uint32_t m = (pw[p >> 1] >> (16 * (p & 1))) & 0xFFFFu;
        if (!m) continue;
        any = true;
        const int mr = bI.rs[p + 1] - bI.rs[p];
        for (int q = 0; q < bJ.natoms; ++q)
          if ((m >> q) & 1u) cs += (int64_t)mr * (bJ.rs[q + 1] - bJ.rs[q]);
      }
    }
    const bool present = any || e > 0;
    const int64_t t_off = wave_excl_i64(present ? 1 : 0);
    const int64_t e_off = wave_excl_i64(present ? e : 0);
    const int64_t c_off = wave_excl_i64(present ? cs : 0);
    if (present) {
      TileDesc td;
      td.I = I;
      td.J = J;
      td.e0 = (int32_t)(e_run + e_off);
      td.ne = (int32_t)e;
      td.ktot = (int32_t)k;
      td.pad = 0;
      td.out_base = c_run + c_off;
      for (int p = 0; p < kBandAtoms; ++p)
        td.pattern[p] = (uint16_t)((pw[p >> 1] >> (16 * (p & 1))) & 0xFFFFu);
      tiles[t_run + t_off] = td;
      // segment cursors: entries / K columns of earlier segments come first
      int64_t ecur = e_run + e_off, kcur = 0;
      for (int g = 0; g < nseg; ++g) {
        const int64_t idx = ((int64_t)I * nseg + g) * nbands + J;
        const int32_t ce = cnt[idx], ck = ksum[idx];
        cnt[idx] = (int32_t)ecur;
        ksum[idx] = (int32_t)kcur;
        ecur += ce;
        kcur += ck;
      }
    }
    t_run += wave_sum_i64(present ? 1 : 0);
    e_run += wave_sum_i64(present ? e : 0);
    c_run += wave_sum_i64(present ? cs : 0);
  }
}

// kLds = false: one wave per (band I, segment g) with its cursors in global
// memory. kLds = true: one workgroup of nseg waves per band I, the band's
// cursors copied to LDS (sym_fill_lds(nbands, nseg) bytes of dynamic shared
// memory) so that reserving entry slots takes LDS atomics.
template <bool kLds>
__global__ __launch_bounds__(kLds ? 512 : 64) void sym_fill(
    int32_t nbands, int nseg, const int32_t* __restrict__ blk, const Strip* __restrict__ as,
    const StripHead* __restrict__ ah, const int64_t* __restrict__ a_ptr,
    const Strip* __restrict__ bs, const StripHead* __restrict__ bh,
    const int64_t* __restrict__ b_ptr, const int64_t* __restrict__ a_pk,
    const int64_t* __restrict__ b_pk, double eps2, int32_t* __restrict__ cursor,
    int32_t* __restrict__ kcur, Entry* __restrict__ entries, const uint32_t* __restrict__ mpat) {
  LAMB_GUARD;
  extern __shared__ int32_t scur[];
  const int32_t I = kLds ? (int32_t)blockIdx.x : (int32_t)(blockIdx.x / nseg);
  const int g = kLds ? (int)(threadIdx.x >> 6) : (int)(blockIdx.x % nseg);
  if (I >= nbands) return;
  const int lane = threadIdx.x & 63;
  const int64_t sa0 = a_ptr[I], sa1 = a_ptr[I + 1];
  const int64_t len = sa1 - sa0;
  const int64_t lo = sa0 + len * g / nseg, hi = sa0 + len * (g + 1) / nseg;
  const int64_t nsb = (int64_t)nseg * nbands;
  if (kLds) {
    for (int64_t x = threadIdx.x; x < nsb; x += blockDim.x) {
      scur[x] = cursor[(int64_t)I * nsb + x];
      scur[nsb + x] = kcur[(int64_t)I * nsb + x];
    }
    __syncthreads();
  }
  int32_t* c = kLds ? scur + (int64_t)g * nbands : cursor + ((int64_t)I * nseg + g) * nbands;
  int32_t* kc = kLds ? scur + nsb + (int64_t)g * nbands : kcur + ((int64_t)I * nseg + g) * nbands;
  StripWalk walk(as, blk, b_ptr, lo, hi);
  for (int64_t sa = lo; sa < hi; ++sa) {
    const Strip* SA = as + sa;
    const StripHead HA = ah[sa];
    const int32_t k = walk.k1;
    const int kd = walk.kd1;
    const int64_t b0 = walk.b01, b1 = walk.b11;
    walk.advance(sa);
    (void)k;
    for (int64_t sb = b0 + lane; sb < b1; sb += 64) {
      const Strip* SB = bs + sb;
      const StripHead HB = bh[sb];
      const int32_t J = HB.band;
      if (kd <= 0) continue;
      int n = 0;
      pair_entries(HA, HB, SA, SB, eps2, [&](uint32_t rm, uint32_t cm) {
        uint32_t pairs[kBandAtoms];
        if (mask_entry(mpat, nbands, I, J, rm, cm, pairs)) ++n;
      });
      if (!n) continue;
      // reserve n entry slots and n*kd K columns of tile (I, J); the lanes
      // of one step hold distinct J and steps run in K order
      int32_t e = atomicAdd(&c[J], n);
      int32_t kk = atomicAdd(&kc[J], n * kd);
      pair_entries(HA, HB, SA, SB, eps2, [&](uint32_t rm, uint32_t cm) {
        uint32_t pairs[kBandAtoms];
        if (!mask_entry(mpat, nbands, I, J, rm, cm, pairs)) return;
        Entry en;
        en.a_col = (uint32_t)(a_pk[sa] / 64) |
                   ((rm == HA.mask && cm == HB.mask) ? 0u : kEntryPartial);
        en.b_col = (uint32_t)(b_pk[sb] / 64);
        en.kstart = kk;
        en.rmask = (uint16_t)rm;
        en.cmask = (uint16_t)cm;
        entries[e] = en;
        ++e;
        kk += kd;
      });
    }
  }
}

// ---------------------------------------------------------------------------
// C index from the numeric kernel's kept masks. A band's tiles (J order)
// are cut into kCidxSeg segments; one wave per (row band I, segment),
// lanes 0..15 = row positions.
// ---------------------------------------------------------------------------
constexpr int kCidxSeg = 8;

__device__ __forceinline__ void cidx_segment(const int64_t* __restrict__ tile_base, int32_t I,
                                             int g, int64_t& t0, int64_t& t1) {
  const int64_t a = tile_base[I], n = tile_base[I + 1] - a;
  t0 = a + n * g / kCidxSeg;
  t1 = a + n * (g + 1) / kCidxSeg;
}

// seg_cnt[(I * kCidxSeg + g) * 16 + p] = kept blocks of row p in segment g,
// row_cnt[first + p] = their sum. One wave per band: lane (p, q) counts
// segments q and q + 4 of row p.
__global__ __launch_bounds__(64) void cidx_count(int32_t nbands, const Band* __restrict__ bands,
                                                 const int64_t* __restrict__ tile_base,
                                                 const uint16_t* __restrict__ kept,
                                                 int32_t* __restrict__ seg_cnt,
                                                 int64_t* __restrict__ row_cnt) {
  LAMB_GUARD;
  const int32_t I = blockIdx.x;
  if (I >= nbands) return;
  const int lane = threadIdx.x;
  const int p = lane & 15, q = lane >> 4;  // row p; segments q and q + 4
  const int nat = bands[I].natoms;
  int32_t n[2] = {0, 0};
  if (p < nat)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      int64_t t0, t1;
      cidx_segment(tile_base, I, q + 4 * h, t0, t1);
      for (int64_t t = t0; t < t1; ++t) n[h] += __builtin_popcount(kept[t * kBandAtoms + p]);
      seg_cnt[((int64_t)I * kCidxSeg + q + 4 * h) * 16 + p] = n[h];
    }
  int64_t tot = (int64_t)n[0] + n[1];
  tot += __shfl_xor(tot, 16, 64);
  tot += __shfl_xor(tot, 32, 64);
  if (lane < nat) row_cnt[bands[I].first + lane] = tot;
}

// Writes the C index of (row band I, segment g): for each kept block of the
// segment's tiles its column atom, C pool offset and row atom, at its CSR
// position (a band's tiles are in J order and a tile's columns ascend, so
// each row's blocks come in column order). Passes of 16 tiles. First, lanes
// over the 16 rows of 4 tiles at a time: each (tile, row)'s offset inside
// its tile's C region (laid out row by row), pattern and kept masks, and the
// tile's column band, into LDS. Then lanes over the 16 tiles of 4 rows at a
// time write the index: the lanes of a row write neighbouring CSR positions,
// where lanes over rows would each write a different part of the arrays.
__global__ __launch_bounds__(64) void cidx_fill(int32_t nbands, const Band* __restrict__ bands,
                                                const int64_t* __restrict__ tile_base,
                                                const TileDesc* __restrict__ tiles,
                                                const uint16_t* __restrict__ kept,
                                                const int32_t* __restrict__ seg_cnt,
                                                const int64_t* __restrict__ row_p,
                                                int32_t* __restrict__ col, int64_t* __restrict__ off,
                                                int32_t* __restrict__ rows,
                                                unsigned long long* __restrict__ nelems) {
  LAMB_GUARD;
  const int32_t I = (int32_t)(blockIdx.x / kCidxSeg);
  const int g = (int)(blockIdx.x % kCidxSeg);
  if (I >= nbands) return;
  const int lane = threadIdx.x;
  constexpr int kPass = 16;  // tiles per pass
  __shared__ int64_t s_next[kBandAtoms];               // next CSR position of row p
  __shared__ int32_t s_m[kBandAtoms];                  // rows of band I's atom p
  __shared__ int64_t s_off[kPass][kBandAtoms];         // C pool offset of row p in tile l
  __shared__ uint16_t s_pm[kPass][kBandAtoms], s_km[kPass][kBandAtoms];
  __shared__ int16_t s_rs[kPass][kBandAtoms + 1];      // column starts of tile l's band J
  __shared__ int32_t s_jfirst[kPass], s_jnat[kPass];
  const Band& bI = bands[I];
  const int nat = bI.natoms;
  if (lane < kBandAtoms) {
    int64_t b = 0;
    int m = 0;
    if (lane < nat) {
      b = row_p[bI.first + lane];
      for (int h = 0; h < g; ++h) b += seg_cnt[((int64_t)I * kCidxSeg + h) * 16 + lane];
      m = bI.rs[lane + 1] - bI.rs[lane];
    }
    s_next[lane] = b;
    s_m[lane] = m;
  }
  __syncthreads();
  unsigned long long ne = 0;
  int64_t t0, t1;
  cidx_segment(tile_base, I, g, t0, t1);
  for (int64_t tb = t0; tb < t1; tb += kPass) {
    // ---- row offsets, masks and column bands of the pass's tiles ----
    {
      const int p = lane & 15, slot = lane >> 4;
      for (int sp = 0; sp < kPass / 4; ++sp) {
        const int l = 4 * sp + slot;
        const int64_t t = tb + l;
        const bool tv = t < t1;
        const TileDesc* td = tiles + (tv ? t : tb);
        const Band& bJ = bands[td->J];
        const bool act = tv && p < nat;
        const uint32_t pm = act ? td->pattern[p] : 0u;
        const int m = s_m[p];
        const int jn = bJ.natoms;
        int64_t rowsz = 0;
        for (int q = 0; q < jn; ++q)
          if ((pm >> q) & 1u) rowsz += (int64_t)m * (bJ.rs[q + 1] - bJ.rs[q]);
        // exclusive prefix over the 16 rows of this tile
        int64_t x = rowsz;
        for (int o = 1; o < 16; o <<= 1) {
          const int64_t y = __shfl_up(x, o, 16);
          if (p >= o) x += y;
        }
        s_off[l][p] = td->out_base + x - rowsz;
        s_pm[l][p] = (uint16_t)pm;
        s_km[l][p] = act ? kept[t * kBandAtoms + p] : (uint16_t)0;
        s_rs[l][p] = bJ.rs[min(p, jn)];
        if (p == 0) {
          s_rs[l][kBandAtoms] = bJ.rs[jn];
          s_jfirst[l] = bJ.first;
          s_jnat[l] = tv ? jn : 0;
        }
      }
    }
    __syncthreads();
    // ---- the index: lane (r, l) = row 4 * round + r of tile tb + l ----
    {
      const int l = lane & 15, r = lane >> 4;
      const int jn = s_jnat[l];
      for (int round = 0; round < kBandAtoms / 4; ++round) {
        const int p = 4 * round + r;
        if (4 * round >= nat) break;  // no row of this round in band I
        const uint32_t pm = p < nat ? s_pm[l][p] : 0u;
        const uint32_t km = s_km[l][p];
        const int nk = __builtin_popcount(pm & km);
        int x = nk;  // inclusive prefix over the 16 tiles of this row
        for (int o = 1; o < 16; o <<= 1) {
          const int y = __shfl_up(x, o, 16);
          if (l >= o) x += y;
        }
        const int total = __shfl(x, 15, 16);
        const int64_t first = s_next[p];
        int64_t bb = first + x - nk;
        int64_t o_run = s_off[l][p];
        const int m = s_m[p];
        const int32_t jf = s_jfirst[l];
        for (int q = 0; q < jn; ++q) {
          if (!((pm >> q) & 1u)) continue;
          const int64_t sz = (int64_t)m * (s_rs[l][q + 1] - s_rs[l][q]);
          if ((km >> q) & 1u) {
            col[bb] = jf + q;
            off[bb] = o_run;
            rows[bb] = bI.first + p;
            ++bb;
            ne += (unsigned long long)sz;
          }
          o_run += sz;
        }
        if (l == 15 && p < nat) s_next[p] = first + total;
      }
    }
    __syncthreads();  // the next pass rewrites the LDS
  }
  ne = wave_sum_i64((int64_t)ne);
  if (lane == 0) atomicAdd(nelems, ne);
}

// ---------------------------------------------------------------------------
// Tile schedule. Tiles are stored band-major (I, J). The numeric kernel walks
// them in blocks of rb row bands, column band by column band: order key
// (I / rb, J, I % rb), so the rb tiles that share the B strips of column
// band J run back to back. perm[position] = tile.
// ---------------------------------------------------------------------------
__global__ void tile_cells(int64_t ntiles, int rb, int32_t nbands, const TileDesc* __restrict__ tiles,
                           uint32_t* __restrict__ cell_mask) {
  LAMB_GUARD;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const int32_t I = tiles[t].I, J = tiles[t].J;
  atomicOr(&cell_mask[(int64_t)(I / rb) * nbands + J], 1u << (I % rb));
}

__global__ void cell_counts(int64_t ncells, const uint32_t* __restrict__ cell_mask,
                            int64_t* __restrict__ cnt) {
  LAMB_GUARD;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncells) cnt[c] = __builtin_popcount(cell_mask[c]);
}

__global__ void tile_perm(int64_t ntiles, int rb, int32_t nbands, const TileDesc* __restrict__ tiles,
                          const uint32_t* __restrict__ cell_mask, const int64_t* __restrict__ base,
                          int32_t* __restrict__ perm) {
  LAMB_GUARD;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const int32_t I = tiles[t].I, J = tiles[t].J;
  const int64_t c = (int64_t)(I / rb) * nbands + J;
  const uint32_t below = cell_mask[c] & ((1u << (I % rb)) - 1u);
  perm[base[c] + __builtin_popcount(below)] = (int32_t)t;
}

// XCD ranges: weight of each position of the tile order (its tile's K
// columns, which track the tile's loads and MFMAs to a few per cent), then,
// from the exclusive scan acc of the weights (acc[nt] = total), position p
// goes to XCD floor(8 (acc[p] + w_p / 2) / total). That XCD index never
// decreases along the order, so each XCD gets one contiguous range, and
// xr[x] = its first position: p writes xr[x] for the XCDs x that start at p
// (those after its predecessor's, up to its own). xr holds nt beforehand, so
// XCDs past the last position get the empty range [nt, nt).
__global__ void tile_weights(int64_t nt, const TileDesc* __restrict__ tiles,
                             const int32_t* __restrict__ perm, int64_t* __restrict__ w) {
  LAMB_GUARD;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nt) w[p] = tiles[perm[p]].ktot;
}

__device__ __forceinline__ int xcd_of_position(const int64_t* __restrict__ acc, int64_t nt,
                                               int64_t p) {
  const int64_t tot = acc[nt], a = acc[p], w = acc[p + 1] - a;
  const int64_t x = tot > 0 ? (8 * (2 * a + w)) / (2 * tot) : (8 * p) / nt;
  return (int)(x < 7 ? x : 7);
}

__global__ void xcd_bounds(int64_t nt, const int64_t* __restrict__ acc, int32_t* __restrict__ xr) {
  LAMB_GUARD;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nt) return;
  const int xp = xcd_of_position(acc, nt, p);
  const int xq = p == 0 ? -1 : xcd_of_position(acc, nt, p - 1);
  for (int x = xq + 1; x <= xp; ++x) xr[x] = (int32_t)p;
}

// out[0] = *a, out[1] = *b: two device counts gathered for one read
__global__ void gather_pair(const int64_t* __restrict__ a, const int64_t* __restrict__ b,
                            int64_t* __restrict__ out) {
  LAMB_GUARD;
  if (threadIdx.x == 0) {
    out[0] = *a;
    out[1] = *b;
  }
}

// out[0..3] = the four chunk-total sums (after their scans, at [v*(nb+2)+nb]
// with nb the number of scanned values), out[4..7] = kept, dropped, flops
// and partial-entry statistics
__global__ void gather_totals(const int64_t* __restrict__ bts, int32_t nb,
                              const int64_t* __restrict__ stats, int64_t* __restrict__ out) {
  LAMB_GUARD;
  const int t = threadIdx.x;
  if (t < 4) out[t] = bts[(int64_t)t * (nb + 2) + nb];
  else if (t < 8) out[t] = stats[t - 4];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
#define CK(expr)                             \
  do {                                       \
    hipError_t e_ = (expr);                  \
    if (e_ != hipSuccess) return e_;         \
  } while (0)

static unsigned grid_of(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

namespace {
// scratch of one side of the strip build
struct StripScratch {
  DevBuf* flag;
  DevBuf* excl;
  DevBuf* gslot;
};

// slot flags and their scan: excl[nblk] = the side's strip count
hipError_t strips_count(hipStream_t s, SymScratch& W, const StripScratch& sc, const MatView& M,
                        const DistView& D, bool a_side) {
  const int64_t nblk = M.nblk;
  // the slot sequence: merged (A) or CSR (B) order
  CK(sc.flag->ensure(sizeof(int64_t) * (size_t)(nblk + 1)));
  CK(sc.excl->ensure(sizeof(int64_t) * (size_t)(nblk + 2)));
  CK(W.scan.ensure(sizeof(int64_t) * (size_t)scan_scratch_elems(nblk + 1)));
  if (a_side) {
    CK(W.merged.ensure(sizeof(int64_t) * (size_t)std::max<int64_t>(nblk, 1)));
    if (nblk)
      hipLaunchKernelGGL(a_merge_rank, dim3(grid_of(nblk)), dim3(256), 0, s, nblk, M.rows, M.col,
                         M.row_p, D.band_of, D.pos_in_band, D.bands, (int64_t*)W.merged.p);
  }
  if (nblk)
    hipLaunchKernelGGL(strip_flags, dim3(grid_of(nblk)), dim3(256), 0, s, nblk, a_side ? 1 : 0,
                       (const int64_t*)W.merged.p, M.rows, M.col, D.band_of,
                       (int64_t*)sc.flag->p);
  CK(hipGetLastError());
  return exclusive_scan(s, (const int64_t*)sc.flag->p, (int64_t*)sc.excl->p, nblk,
                        (int64_t*)W.scan.p);
}

// strips, group pointers, heads and packed panel offsets: pk[nstrips] = the
// side's packed panel elements
hipError_t strips_fill_side(hipStream_t s, SymScratch& W, const StripScratch& sc,
                            const MatView& M, const DistView& D, bool a_side, int64_t nstrips,
                            StripSet& out) {
  const int64_t nblk = M.nblk;
  out.n = nstrips;
  CK(out.strips.ensure(sizeof(Strip) * (size_t)std::max<int64_t>(nstrips, 1)));
  if (nstrips)
    hipLaunchKernelGGL(strip_init, dim3(grid_of(nstrips)), dim3(256), 0, s, nstrips,
                       (Strip*)out.strips.p);
  if (nblk)
    hipLaunchKernelGGL(strip_fill, dim3(grid_of(nblk)), dim3(256), 0, s, nblk, a_side ? 1 : 0,
                       (const int64_t*)W.merged.p, (const int64_t*)sc.excl->p,
                       (const int64_t*)sc.flag->p, M.rows, M.col, M.off, M.nsq, D.band_of,
                       D.pos_in_band, (Strip*)out.strips.p);
  CK(hipGetLastError());
  // group pointers: per band (A) or per row atom (B)
  const int64_t ngroups = a_side ? D.nbands : D.n;
  CK(out.ptr.ensure(sizeof(int64_t) * (size_t)(ngroups + 1)));
  CK(sc.gslot->ensure(sizeof(int64_t) * (size_t)std::max<int64_t>(ngroups, 1)));
  if (a_side) {
    if (D.nbands)
      hipLaunchKernelGGL(band_first_slot, dim3(grid_of(D.nbands)), dim3(256), 0, s, D.nbands,
                         D.bands, M.row_p, (int64_t*)sc.gslot->p);
  } else {
    if (D.n)
      CK(hipMemcpyAsync(sc.gslot->p, M.row_p, sizeof(int64_t) * (size_t)D.n,
                        hipMemcpyDeviceToDevice, s));
  }
  hipLaunchKernelGGL(group_ptr, dim3(grid_of(ngroups + 1)), dim3(256), 0, s, ngroups,
                     (const int64_t*)sc.gslot->p, (const int64_t*)sc.excl->p, nblk, nstrips,
                     (int64_t*)out.ptr.p);
  CK(hipGetLastError());
  // packed panel offsets (the flags are consumed: their buffer takes the
  // per-strip panel sizes)
  CK(out.pk.ensure(sizeof(int64_t) * (size_t)(nstrips + 1)));
  CK(sc.flag->ensure(sizeof(int64_t) * (size_t)(std::max(nblk, nstrips) + 1)));
  CK(W.scan.ensure(sizeof(int64_t) * (size_t)scan_scratch_elems(std::max(nblk, nstrips) + 1)));
  CK(out.head.ensure(sizeof(StripHead) * (size_t)std::max<int64_t>(nstrips, 1)));
  if (nstrips)
    hipLaunchKernelGGL(strip_panel_size, dim3(grid_of(nstrips)), dim3(256), 0, s, nstrips,
                       (const Strip*)out.strips.p, D.blk, D.bands, (int64_t*)sc.flag->p,
                       (StripHead*)out.head.p);
  return exclusive_scan(s, (const int64_t*)sc.flag->p, (int64_t*)out.pk.p, nstrips,
                        (int64_t*)W.scan.p);
}
}  // namespace

hipError_t build_strips(lamb_hip_ctx* ctx, SymScratch& W, const MatView& MA, const MatView& MB,
                        const DistView& D, StripSet& A, StripSet& B) {
  hipStream_t s = ctx->stream;
  const StripScratch sa{&W.flag, &W.excl, &W.gslot}, sb{&W.flag_b, &W.excl_b, &W.gslot_b};
  CK(W.sizes.ensure(sizeof(int64_t) * 2));
  int64_t* sizes = (int64_t*)W.sizes.p;
  CK(strips_count(s, W, sa, MA, D, true));
  CK(strips_count(s, W, sb, MB, D, false));
  int64_t n[2] = {0, 0};  // strips of A, of B
  hipLaunchKernelGGL(gather_pair, dim3(1), dim3(64), 0, s, (const int64_t*)W.excl.p + MA.nblk,
                     (const int64_t*)W.excl_b.p + MB.nblk, sizes);
  CK(hipGetLastError());
  CK(read_sizes(ctx, sizes, n, 2));
  CK(strips_fill_side(s, W, sa, MA, D, true, n[0], A));
  CK(strips_fill_side(s, W, sb, MB, D, false, n[1], B));
  int64_t pk[2] = {0, 0};  // packed panel elements of A, of B
  hipLaunchKernelGGL(gather_pair, dim3(1), dim3(64), 0, s, (const int64_t*)A.pk.p + n[0],
                     (const int64_t*)B.pk.p + n[1], sizes);
  CK(hipGetLastError());
  CK(read_sizes(ctx, sizes, pk, 2));
  A.pk_total = pk[0];
  B.pk_total = pk[1];
  return hipSuccess;
}

hipError_t build_tiles(lamb_hip_ctx* ctx, SymScratch& W, const DistView& D, const StripSet& A,
                       const StripSet& B, const MatView* Cold, double eps2, SymResult& r,
                       const MatView* mask) {
  hipStream_t s = ctx->stream;
  const int32_t nb = D.nbands;
  const int nseg = 8;
  const int64_t dense = (int64_t)nb * nseg * nb;
  const uint32_t* mp = nullptr;
  if (mask) {
    const int64_t words = std::max<int64_t>((int64_t)nb * nb * (kBandAtoms / 2), 1);
    CK(W.mpat.ensure(sizeof(uint32_t) * (size_t)words));
    CK(fill32(s, (int32_t*)W.mpat.p, words, 0));
    if (mask->nblk)
      hipLaunchKernelGGL(mask_pattern, dim3(grid_of(mask->nblk)), dim3(256), 0, s, mask->nblk,
                         mask->rows, mask->col, D.band_of, D.pos_in_band, nb,
                         (uint32_t*)W.mpat.p);
    CK(hipGetLastError());
    mp = (const uint32_t*)W.mpat.p;
  }
  CK(W.cnt.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(dense, 1)));
  CK(W.ksum.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(dense, 1)));
  CK(W.pat.ensure(sizeof(uint32_t) * (size_t)std::max<int64_t>((int64_t)nb * nb * (kBandAtoms / 2), 1)));
  CK(W.stats.ensure(sizeof(unsigned long long) * 16));
  // a band's pattern words and counters in LDS when they fit
  const size_t lds_cnt = sym_count_lds(nb, nseg);
  const bool cnt_lds = lds_cnt <= kSymLdsMax && sym_lds_allowed();
  if (!cnt_lds) {
    CK(fill32(s, (int32_t*)W.cnt.p, dense, 0));
    CK(fill32(s, (int32_t*)W.ksum.p, dense, 0));
    CK(fill32(s, (int32_t*)W.pat.p, (int64_t)nb * nb * (kBandAtoms / 2), 0));
  }
  CK(fill32(s, (int32_t*)W.stats.p, 8, 0));
  if (nb) {
    if (cnt_lds)
      hipLaunchKernelGGL(sym_count<true>, dim3((unsigned)nb), dim3(64 * nseg), lds_cnt, s, nb, nseg, D.bands,
                       D.blk, (const Strip*)A.strips.p, (const StripHead*)A.head.p, (const int64_t*)A.ptr.p,
                       (const Strip*)B.strips.p, (const StripHead*)B.head.p, (const int64_t*)B.ptr.p, eps2,
                       Cold ? Cold->row_p : nullptr, Cold ? Cold->col : nullptr, D.band_of,
                       D.pos_in_band, (int32_t*)W.cnt.p, (int32_t*)W.ksum.p,
                       (uint32_t*)W.pat.p, (unsigned long long*)W.stats.p, mp);
    else
      hipLaunchKernelGGL(sym_count<false>, dim3((unsigned)(nb * nseg)), dim3(64), 0, s, nb, nseg, D.bands,
                       D.blk, (const Strip*)A.strips.p, (const StripHead*)A.head.p, (const int64_t*)A.ptr.p,
                       (const Strip*)B.strips.p, (const StripHead*)B.head.p, (const int64_t*)B.ptr.p, eps2,
                       Cold ? Cold->row_p : nullptr, Cold ? Cold->col : nullptr, D.band_of,
                       D.pos_in_band, (int32_t*)W.cnt.p, (int32_t*)W.ksum.p,
                       (uint32_t*)W.pat.p, (unsigned long long*)W.stats.p, mp);
  }
  CK(hipGetLastError());
  // per-(band, 64-band chunk) totals, scanned in (band, chunk) order: each
  // chunk's scanned value is where its tiles, entries and C blocks start
  const int32_t nch = std::max<int32_t>((nb + 63) / 64, 1);
  const int64_t nx = (int64_t)nb * nch;
  CK(W.band4.ensure(sizeof(int64_t) * 4 * (size_t)(nx + 1)));
  CK(W.band4s.ensure(sizeof(int64_t) * 4 * (size_t)(nx + 2)));
  int64_t* bt = (int64_t*)W.band4.p;
  int64_t* bts = (int64_t*)W.band4s.p;
  if (nb)
    hipLaunchKernelGGL(sym_chunk_totals, dim3((unsigned)nx), dim3(64), 0, s, nb, nseg, nch,
                       D.bands, (const int32_t*)W.cnt.p, (const uint32_t*)W.pat.p, bt,
                       bt + (nx + 1), bt + 2 * (nx + 1), bt + 3 * (nx + 1));
  CK(hipGetLastError());
  CK(W.scan.ensure(sizeof(int64_t) * (size_t)scan_scratch_elems(nx + 1)));
  for (int v = 0; v < 4; ++v)
    CK(exclusive_scan(s, bt + v * (nx + 1), bts + v * (nx + 2), nx, (int64_t*)W.scan.p));
  // the four totals and four statistics, gathered for one read
  int64_t* gathered = (int64_t*)W.stats.p + 8;
  hipLaunchKernelGGL(gather_totals, dim3(1), dim3(64), 0, s, bts, (int32_t)nx,
                     (const int64_t*)W.stats.p, gathered);
  CK(hipGetLastError());
  int64_t tot[8];
  CK(read_sizes(ctx, gathered, tot, 8));
  r.n_tiles = tot[0];
  r.n_entries = tot[1];
  r.c_size = tot[2];
  r.n_struct = tot[3];
  r.kept = tot[4];
  r.dropped = tot[5];
  r.flops = (double)(unsigned long long)tot[6];
  r.n_partial = tot[7];
  CK(r.tiles.ensure(sizeof(TileDesc) * (size_t)std::max<int64_t>(r.n_tiles, 1)));
  CK(r.entries.ensure(sizeof(Entry) * (size_t)std::max<int64_t>(r.n_entries, 1)));
  CK(r.tile_base.ensure(sizeof(int64_t) * (size_t)(nb + 1)));
  hipLaunchKernelGGL(band_tile_base, dim3(grid_of(nb + 1)), dim3(256), 0, s, nb, nch,
                     (const int64_t*)bts, (int64_t*)r.tile_base.p);
  if (nb)
    hipLaunchKernelGGL(sym_tiles, dim3((unsigned)nx), dim3(64), 0, s, nb, nseg, nch, D.bands,
                       (int32_t*)W.cnt.p, (int32_t*)W.ksum.p, (const uint32_t*)W.pat.p, bts,
                       bts + (nx + 2), bts + 2 * (nx + 2), (TileDesc*)r.tiles.p);
  CK(hipGetLastError());
  if (nb) {
    const size_t lds_fill = sym_fill_lds(nb, nseg);
    if (lds_fill <= kSymLdsMax && sym_lds_allowed())
      hipLaunchKernelGGL(sym_fill<true>, dim3((unsigned)nb), dim3(64 * nseg), lds_fill, s, nb,
                         nseg, D.blk, (const Strip*)A.strips.p, (const StripHead*)A.head.p, (const int64_t*)A.ptr.p,
                         (const Strip*)B.strips.p, (const StripHead*)B.head.p, (const int64_t*)B.ptr.p,
                         (const int64_t*)A.pk.p, (const int64_t*)B.pk.p, eps2,
                         (int32_t*)W.cnt.p, (int32_t*)W.ksum.p, (Entry*)r.entries.p, mp);
    else
      hipLaunchKernelGGL(sym_fill<false>, dim3((unsigned)(nb * nseg)), dim3(64), 0, s, nb, nseg,
                         D.blk, (const Strip*)A.strips.p, (const StripHead*)A.head.p, (const int64_t*)A.ptr.p,
                         (const Strip*)B.strips.p, (const StripHead*)B.head.p, (const int64_t*)B.ptr.p,
                         (const int64_t*)A.pk.p, (const int64_t*)B.pk.p, eps2,
                         (int32_t*)W.cnt.p, (int32_t*)W.ksum.p, (Entry*)r.entries.p, mp);
  }
  return hipGetLastError();
}

hipError_t build_c_index(lamb_hip_ctx* ctx, SymScratch& W, const DistView& D,
                         const SymResult& r, const uint16_t* kept, CIndex& c) {
  hipStream_t s = ctx->stream;
  const int32_t nb = D.nbands;
  CK(c.row_p.ensure(sizeof(int64_t) * (size_t)(D.n + 1)));
  CK(W.flag.ensure(sizeof(int64_t) * (size_t)(D.n + 1)));
  CK(W.scan.ensure(sizeof(int64_t) * (size_t)scan_scratch_elems(D.n + 1)));
  CK(W.stats.ensure(sizeof(unsigned long long) * 4));
  CK(W.cidx.ensure(sizeof(int32_t) * (size_t)kCidxSeg * 16 * (size_t)std::max(nb, 1)));
  if (nb)
    hipLaunchKernelGGL(cidx_count, dim3(nb), dim3(64), 0, s, nb, D.bands,
                       (const int64_t*)r.tile_base.p, kept, (int32_t*)W.cidx.p,
                       (int64_t*)W.flag.p);
  CK(hipGetLastError());
  CK(exclusive_scan(s, (const int64_t*)W.flag.p, (int64_t*)c.row_p.p, D.n, (int64_t*)W.scan.p));
  int64_t nblk = 0;
  CK(read_sizes(ctx, (int64_t*)c.row_p.p + D.n, &nblk, 1));
  c.nblk = nblk;
  CK(c.col.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(nblk, 1)));
  CK(c.off.ensure(sizeof(int64_t) * (size_t)std::max<int64_t>(nblk, 1)));
  CK(c.rows.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(nblk, 1)));
  CK(fill32(s, (int32_t*)W.stats.p, 2, 0));
  if (nb)
    hipLaunchKernelGGL(cidx_fill, dim3((unsigned)(nb * kCidxSeg)), dim3(64), 0, s, nb, D.bands,
                       (const int64_t*)r.tile_base.p, (const TileDesc*)r.tiles.p, kept,
                       (const int32_t*)W.cidx.p, (const int64_t*)c.row_p.p, (int32_t*)c.col.p,
                       (int64_t*)c.off.p, (int32_t*)c.rows.p, (unsigned long long*)W.stats.p);
  CK(hipGetLastError());
  int64_t ne = 0;
  CK(read_sizes(ctx, W.stats.p, &ne, 1));
  c.nelems = ne;
  return hipSuccess;
}


hipError_t build_tile_order(lamb_hip_ctx* ctx, SymScratch& W, const DistView& D, SymResult& r,
                            int rb) {
  hipStream_t s = ctx->stream;
  const int32_t nb = D.nbands;
  const int64_t nt = r.n_tiles;
  CK(r.perm.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(nt, 1)));
  if (nt == 0) return hipSuccess;
  const int64_t ncells = (int64_t)((nb + rb - 1) / rb) * nb;
  CK(W.cells.ensure(sizeof(uint32_t) * (size_t)ncells));
  CK(W.cellcnt.ensure(sizeof(int64_t) * (size_t)(ncells + 1)));
  CK(W.cellbase.ensure(sizeof(int64_t) * (size_t)(ncells + 1)));
  CK(W.scan.ensure(sizeof(int64_t) * (size_t)scan_scratch_elems(ncells + 1)));
  CK(fill32(s, (int32_t*)W.cells.p, ncells, 0));
  hipLaunchKernelGGL(tile_cells, dim3(grid_of(nt)), dim3(256), 0, s, nt, rb, nb,
                     (const TileDesc*)r.tiles.p, (uint32_t*)W.cells.p);
  hipLaunchKernelGGL(cell_counts, dim3(grid_of(ncells)), dim3(256), 0, s, ncells,
                     (const uint32_t*)W.cells.p, (int64_t*)W.cellcnt.p);
  CK(hipGetLastError());
  CK(exclusive_scan(s, (const int64_t*)W.cellcnt.p, (int64_t*)W.cellbase.p, ncells,
                    (int64_t*)W.scan.p));
  hipLaunchKernelGGL(tile_perm, dim3(grid_of(nt)), dim3(256), 0, s, nt, rb, nb,
                     (const TileDesc*)r.tiles.p, (const uint32_t*)W.cells.p,
                     (const int64_t*)W.cellbase.p, (int32_t*)r.perm.p);
  return hipGetLastError();
}

hipError_t build_xcd_ranges(lamb_hip_ctx* ctx, SymScratch& W, const SymResult& r, int32_t* xr) {
  hipStream_t s = ctx->stream;
  const int64_t nt = r.n_tiles;
  CK(W.tw.ensure(sizeof(int64_t) * (size_t)(nt + 1)));
  CK(W.scan.ensure(sizeof(int64_t) * (size_t)scan_scratch_elems(nt + 1)));
  CK(fill32(s, xr, 9, (int32_t)nt));
  if (nt == 0) return hipSuccess;
  hipLaunchKernelGGL(tile_weights, dim3(grid_of(nt)), dim3(256), 0, s, nt,
                     (const TileDesc*)r.tiles.p, (const int32_t*)r.perm.p, (int64_t*)W.tw.p);
  CK(hipGetLastError());
  CK(exclusive_scan(s, (const int64_t*)W.tw.p, (int64_t*)W.tw.p, nt, (int64_t*)W.scan.p));
  hipLaunchKernelGGL(xcd_bounds, dim3(grid_of(nt)), dim3(256), 0, s, nt, (const int64_t*)W.tw.p,
                     xr);
  return hipGetLastError();
}

}  // namespace lamb
