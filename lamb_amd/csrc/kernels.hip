// HIP kernels for gfx950: block norms, FP64 numeric phase, trace.
#include "common.h"

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
// 64 x 64 values. The tile's K dimension is the concatenation of the K atoms
// of its entries, streamed through LDS kChunkK columns at a time:
//   As[q][r] = packed A strip value (tile row r, K column q),
//   Bs[q][c] = packed B strip value (K column q, tile col c).
// Pipeline per chunk c (one barrier): issue the global loads of chunk c+1
// into registers, wave 0 builds the column table of chunk c+2, all waves
// run the MFMAs of chunk c, then the registers go to the other LDS buffer.
//
// Each wave owns a 32 x 32 quadrant as 64 4x4 sub-tiles, held as 16
// accumulators of v_mfma_f64_4x4x4_4b (4 independent 4x4x4 blocks per
// instruction): accumulator h*8+g holds the 4 row groups of half h (blocks
// b = 0..3) for column group g. Per 4-column K step a row-group mask and a
// column-group mask tell which 4x4x4 products are non-zero; an instruction
// is skipped when its column group or all four of its row groups are zero.
//
// v_mfma_f64_4x4x4_4b lane layout (lane l): A[a][c] of block b at
// l = a + 4b + 16c; B[c][t] of block b at l = t + 4b + 16c; D[a][t] of
// block b at l = 16a + 4b + t.
// ---------------------------------------------------------------------------
struct ColSrc {
  int64_t a, b;  // packed source offsets of this K column (-1: zero column)
  uint32_t rmask, cmask;
  uint32_t groups;  // row-group mask | column-group mask << 16
  uint32_t pad;
};

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(256, 4) void numeric_f64_kernel(
    const TileDesc* __restrict__ tiles, const Entry* __restrict__ entries,
    const Band* __restrict__ bands, const double* __restrict__ a_pack,
    const double* __restrict__ b_pack, double alpha, double beta, double eps2,
    const int64_t* __restrict__ old_row_p, const int32_t* __restrict__ old_col,
    const int64_t* __restrict__ old_off, const double* __restrict__ old_data,
    double* __restrict__ c_data, uint16_t* __restrict__ kept_out) {
  constexpr int KC = kChunkK;
  constexpr int PANEL = KC * kPitch;
  constexpr int NT = 4;  // column tables in flight (chunks c .. c+3)
  constexpr int CT = kBandRows * kPitch;
  __shared__ double lds[(4 * PANEL > CT) ? 4 * PANEL : CT];  // [A0 B0 A1 B1]; epilogue C tile
  __shared__ ColSrc tab[NT][KC];
  __shared__ uint32_t smr[NT][KC / 4], smc[NT][KC / 4];
  __shared__ int16_t rs_r[kBandAtoms + 1], rs_c[kBandAtoms + 1];
  __shared__ uint16_t pg_r[kBandAtoms], pg_c[kBandAtoms];
  __shared__ int32_t pair_id[kBandAtoms * kBandAtoms];
  __shared__ int64_t pair_off[kBandAtoms * kBandAtoms];
  __shared__ int32_t wave_cnt[4];
  __shared__ int64_t wave_sz[4];
  __shared__ uint32_t keptw[kBandAtoms];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wr = wave >> 1, wc = wave & 1;

  // Workgroups are dealt round-robin over the 8 XCDs; remap so that each XCD
  // walks one contiguous range of tiles (shared A strips stay in its L2).
  const int nb = gridDim.x;
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int per = nb >> 3, rem = nb & 7;
  const int tile = (xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per) + slot;
  const TileDesc* tp = tiles + tile;
  const int I = tp->I, J = tp->J, e0 = tp->e0, ne = tp->ne, ktot = tp->ktot;
  const int64_t out_base = tp->out_base;
  const int natI = bands[I].natoms, natJ = bands[J].natoms;
  const int32_t firstI = bands[I].first, firstJ = bands[J].first;

  if (tid <= kBandAtoms) {
    rs_r[tid] = bands[I].rs[min(tid, natI)];
    rs_c[tid] = bands[J].rs[min(tid, natJ)];
  }
  if (tid < kBandAtoms) keptw[tid] = 0;
  __syncthreads();
  if (tid < kBandAtoms) {
    uint32_t g = 0;
    if (tid < natI && rs_r[tid + 1] > rs_r[tid])
      for (int q = rs_r[tid] / 4; q <= (rs_r[tid + 1] - 1) / 4; ++q) g |= 1u << q;
    pg_r[tid] = (uint16_t)g;
    g = 0;
    if (tid < natJ && rs_c[tid + 1] > rs_c[tid])
      for (int q = rs_c[tid] / 4; q <= (rs_c[tid + 1] - 1) / 4; ++q) g |= 1u << q;
    pg_c[tid] = (uint16_t)g;
  }
  // band position of this thread's staging row (A) and column (B)
  int rpos = -1, cpos = -1;
  for (int a = 0; a < natI; ++a)
    if (lane >= rs_r[a] && lane < rs_r[a + 1]) rpos = a;
  for (int a = 0; a < natJ; ++a)
    if (lane >= rs_c[a] && lane < rs_c[a + 1]) cpos = a;
  __syncthreads();

  const int nchunks = (ktot + KC - 1) / KC;
  const int e_end = e0 + ne;

  // Column table of chunk ch into tab[buf] / smr[buf] / smc[buf], built by
  // wave 0. The entry cursor is a wave-uniform register of wave 0; lanes
  // exchange data through LDS only across an explicit wavefront fence and
  // wave barrier.
  int cursor = e0;
  auto build_table = [&](int ch, int buf) {
    const int Q0 = ch * KC;
    if (lane < KC && Q0 + lane >= ktot) {
      ColSrc z;
      z.a = -1;
      z.b = -1;
      z.rmask = 0;
      z.cmask = 0;
      z.groups = 0;
      z.pad = 0;
      tab[buf][lane] = z;
    }
    const int e = cursor + lane;
    const bool valid = lane <= KC && e < e_end;
    int ks = 0, kd = 0;
    int64_t asrc = 0, bsrc = 0;
    uint32_t rm = 0, cm = 0;
    if (valid) {
      const Entry* en = entries + e;
      ks = en->kstart;
      kd = en->kdim;
      asrc = en->a_src;
      bsrc = en->b_src;
      rm = en->rmask;
      cm = en->cmask;
    }
    const int lo = max(ks, Q0), hi = min(ks + kd, Q0 + KC);
    if (valid && lo < hi) {
      uint32_t gr = 0, gc = 0;
      for (int p = 0; p < kBandAtoms; ++p) {
        if ((rm >> p) & 1u) gr |= pg_r[p];
        if ((cm >> p) & 1u) gc |= pg_c[p];
      }
      for (int Q = lo; Q < hi; ++Q) {
        ColSrc cs;
        cs.a = asrc + (int64_t)(Q - ks) * 64;
        cs.b = bsrc + (int64_t)(Q - ks) * 64;
        cs.rmask = rm;
        cs.cmask = cm;
        cs.groups = gr | (gc << 16);
        cs.pad = 0;
        tab[buf][Q - Q0] = cs;
      }
    }
    const uint64_t done = __ballot(valid && ks + kd <= Q0 + KC);
    cursor += (int)__popcll(done);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < KC / 4) {
      uint32_t g = 0;
#pragma unroll
      for (int t = 0; t < 4; ++t) g |= tab[buf][4 * lane + t].groups;
      smr[buf][lane] = g & 0xFFFFu;
      smc[buf][lane] = g >> 16;
    }
  };

  if (wave == 0)
    for (int ch = 0; ch < min(nchunks, NT - 1); ++ch) build_table(ch, ch);
  __syncthreads();

  // Two register sets of staged loads: chunk c+1 and c+2 are in flight while
  // chunk c runs on the MFMA pipes. Loads are predicated on the entry's atom
  // masks, so rows / columns of absent atoms cost no memory traffic.
  constexpr int PER = KC / 4;  // K columns per wave per chunk
  double ra0[PER], rb0[PER], ra1[PER], rb1[PER];
#define LAMB_ISSUE(RA, RB, ch)                                             \
  {                                                                        \
    const int t_ = (ch) % NT;                                              \
    _Pragma("unroll") for (int p = 0; p < PER; ++p) {                      \
      const int q = wave + 4 * p;                                          \
      const int64_t sa = tab[t_][q].a, sb = tab[t_][q].b;                  \
      const uint32_t rmk = tab[t_][q].rmask, cmk = tab[t_][q].cmask;       \
      RA[p] = (sa >= 0 && rpos >= 0 && ((rmk >> rpos) & 1u)) ? a_pack[sa + lane] : 0.0; \
      RB[p] = (sb >= 0 && cpos >= 0 && ((cmk >> cpos) & 1u)) ? b_pack[sb + lane] : 0.0; \
    }                                                                      \
  }
#define LAMB_COMMIT(RA, RB, ch)                                            \
  {                                                                        \
    double* As_ = lds + ((ch) & 1) * 2 * PANEL;                            \
    double* Bs_ = As_ + PANEL;                                             \
    _Pragma("unroll") for (int p = 0; p < PER; ++p) {                      \
      const int q = wave + 4 * p;                                          \
      As_[q * kPitch + lane] = RA[p];                                      \
      Bs_[q * kPitch + lane] = RB[p];                                      \
    }                                                                      \
  }
#define LAMB_STEP(RA_NEXT2, RB_NEXT2, RA_NEXT1, RB_NEXT1, c)               \
  {                                                                        \
    if ((c) + 2 < nchunks) LAMB_ISSUE(RA_NEXT2, RB_NEXT2, (c) + 2);        \
    if (wave == 0 && (c) + NT - 1 < nchunks) build_table((c) + NT - 1, ((c) + NT - 1) % NT); \
    {                                                                      \
      const double* As = lds + ((c) & 1) * 2 * PANEL;                      \
      const double* Bs = As + PANEL;                                       \
      const int t3 = (c) % NT;                                             \
      const int nsteps = (min(KC, ktot - (c) * KC) + 3) >> 2;              \
      for (int s = 0; s < nsteps; ++s) {                                   \
        const uint32_t mr = __builtin_amdgcn_readfirstlane(smr[t3][s]) >> (8 * wr); \
        const uint32_t mc = __builtin_amdgcn_readfirstlane(smc[t3][s]) >> (8 * wc); \
        if ((mr & 0xFFu) == 0 || (mc & 0xFFu) == 0) continue;             \
        const int krow = (4 * s + (lane >> 4)) * kPitch;                   \
        double bop[8];                                                     \
        _Pragma("unroll") for (int g = 0; g < 8; ++g)                      \
          bop[g] = ((mc >> g) & 1u) ? Bs[krow + 32 * wc + 4 * g + (lane & 3)] : 0.0; \
        _Pragma("unroll") for (int h = 0; h < 2; ++h) {                    \
          if ((mr >> (4 * h)) & 0xFu) {                                    \
            const double aop = As[krow + 32 * wr + 16 * h + (lane & 15)];  \
            _Pragma("unroll") for (int g = 0; g < 8; ++g)                  \
              if ((mc >> g) & 1u)                                          \
                acc[h * 8 + g] = __builtin_amdgcn_mfma_f64_4x4x4f64(aop, bop[g], acc[h * 8 + g], 0, 0, 0); \
          }                                                                \
        }                                                                  \
      }                                                                    \
    }                                                                      \
    if ((c) + 1 < nchunks) LAMB_COMMIT(RA_NEXT1, RB_NEXT1, (c) + 1);       \
    __syncthreads();                                                       \
  }

  double acc[16];
#pragma unroll
  for (int x = 0; x < 16; ++x) acc[x] = 0.0;

  if (nchunks > 0) LAMB_ISSUE(ra0, rb0, 0);
  if (nchunks > 1) LAMB_ISSUE(ra1, rb1, 1);
  if (nchunks > 0) LAMB_COMMIT(ra0, rb0, 0);
  __syncthreads();
  // chunk c (even): next+2 goes to set 0 (free), next+1 sits in set 1.
  for (int c = 0; c < nchunks; c += 2) {
    LAMB_STEP(ra0, rb0, ra1, rb1, c);
    if (c + 1 < nchunks) LAMB_STEP(ra1, rb1, ra0, rb0, c + 1);
  }
#undef LAMB_ISSUE
#undef LAMB_COMMIT
#undef LAMB_STEP

  // ---- epilogue: C tile to LDS ----
  double* Ct = lds;
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int g = 0; g < 8; ++g) {
      const int row = 32 * wr + 16 * h + 4 * ((lane >> 2) & 3) + (lane >> 4);
      const int col = 32 * wc + 4 * g + (lane & 3);
      Ct[row * kPitch + col] = acc[h * 8 + g];
    }
  // structural pairs (row-major over band positions) and their offsets
  const int ip = tid >> 4, jp = tid & 15;
  const bool present = ip < natI && jp < natJ && ((tp->pattern[ip] >> jp) & 1);
  const int64_t sz =
      present ? (int64_t)(rs_r[ip + 1] - rs_r[ip]) * (rs_c[jp + 1] - rs_c[jp]) : 0;
  int cnt = present ? 1 : 0;
  int64_t csz = sz;
  for (int o = 1; o < 64; o <<= 1) {
    const int c2 = __shfl_up(cnt, o, 64);
    const int64_t s2 = __shfl_up(csz, o, 64);
    if (lane >= o) {
      cnt += c2;
      csz += s2;
    }
  }
  if (lane == 63) {
    wave_cnt[wave] = cnt;
    wave_sz[wave] = csz;
  }
  __syncthreads();
  int cbase = 0;
  int64_t sbase = 0;
  for (int w = 0; w < wave; ++w) {
    cbase += wave_cnt[w];
    sbase += wave_sz[w];
  }
  const int npairs = wave_cnt[0] + wave_cnt[1] + wave_cnt[2] + wave_cnt[3];
  if (present) {
    pair_id[cbase + cnt - 1] = tid;
    pair_off[cbase + cnt - 1] = sbase + csz - sz;
  }
  __syncthreads();

  for (int pi = wave; pi < npairs; pi += 4) {
    const int id = pair_id[pi];
    const int pr = id >> 4, pc = id & 15;
    const int m = rs_r[pr + 1] - rs_r[pr];
    const int n = rs_c[pc + 1] - rs_c[pc];
    const int len = m * n;
    const int64_t base = out_base + pair_off[pi];
    const double* old = nullptr;
    if (beta != 0.0 && old_row_p != nullptr) {
      const int32_t gi = firstI + pr, gj = firstJ + pc;
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
    double part = 0.0;
    for (int e = lane; e < len; e += 64) {
      const int cc = e / m, r = e - cc * m;
      double v = alpha * Ct[(rs_r[pr] + r) * kPitch + rs_c[pc] + cc];
      if (old) v = v + beta * old[e];
      c_data[base + e] = v;
      part = __builtin_fma(v, v, part);
    }
    const double nsq = wave_sum(part);
    if (lane == 0 && nsq >= eps2) atomicOr(&keptw[pr], 1u << pc);
  }
  __syncthreads();
  if (tid < kBandAtoms) kept_out[(int64_t)tile * kBandAtoms + tid] = (uint16_t)keptw[tid];
}

hipError_t launch_numeric_f64(hipStream_t s, int32_t n_tiles, const TileDesc* tiles,
                              const Entry* entries, const Band* bands,
                              const double* a_pack, const double* b_pack, double alpha,
                              double beta, double eps2, const int64_t* old_row_p,
                              const int32_t* old_col, const int64_t* old_off,
                              const double* old_data, double* c_data, uint16_t* kept) {
  if (n_tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(numeric_f64_kernel, dim3((unsigned)n_tiles), dim3(256), 0, s, tiles,
                     entries, bands, a_pack, b_pack, alpha, beta, eps2, old_row_p, old_col,
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
