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
// FP64 numeric phase.
//
// One workgroup (4 waves) owns one C tile = row band I x col band J, at most
// 64 x 64 values. The tile's K dimension is the concatenation of the K atoms
// of its entries; it is staged through LDS kChunkK columns at a time:
//   As[q][r] = A strip value (tile row r, K column q), zero where absent,
//   Bs[q][c] = B strip value (K column q, tile col c).
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
struct ColInfo {
  int32_t a_strip, b_strip;
  int16_t kk, kdim;
  uint16_t rmask, cmask;
};

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(256) void numeric_f64_kernel(
    const TileDesc* __restrict__ tiles, const Entry* __restrict__ entries,
    const Strip* __restrict__ a_strips, const Strip* __restrict__ b_strips,
    const Band* __restrict__ bands, const int32_t* __restrict__ blk,
    const double* __restrict__ a_data, const double* __restrict__ b_data, double alpha,
    double beta, double eps2, const int64_t* __restrict__ old_row_p,
    const int32_t* __restrict__ old_col, const int64_t* __restrict__ old_off,
    const double* __restrict__ old_data, double* __restrict__ c_data,
    uint16_t* __restrict__ kept_out) {
  constexpr int KC = kChunkK;
  __shared__ double lds[2 * KC * kPitch];  // As | Bs, later the C tile [64][kPitch]
  __shared__ ColInfo colinfo[KC];
  __shared__ uint32_t smask_r[KC / 4], smask_c[KC / 4];
  __shared__ int8_t rowpos[kBandRows], rowin[kBandRows], colpos[kBandRows], colin[kBandRows];
  __shared__ int16_t rs_r[kBandAtoms + 1], rs_c[kBandAtoms + 1];
  __shared__ uint16_t pg_r[kBandAtoms], pg_c[kBandAtoms];
  __shared__ int32_t e_cursor;
  __shared__ int32_t pair_id[kBandAtoms * kBandAtoms];
  __shared__ int64_t pair_off[kBandAtoms * kBandAtoms];
  __shared__ int32_t wave_tot[4];
  __shared__ uint32_t keptw[kBandAtoms];

  double* As = lds;
  double* Bs = lds + KC * kPitch;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;

  const TileDesc T = tiles[blockIdx.x];
  const Band bI = bands[T.I];
  const Band bJ = bands[T.J];

  if (tid <= kBandAtoms) {
    rs_r[tid] = bI.rs[tid < bI.natoms ? tid : bI.natoms];
    rs_c[tid] = bJ.rs[tid < bJ.natoms ? tid : bJ.natoms];
  }
  if (tid < kBandRows) {
    int p = -1;
    for (int a = 0; a < bI.natoms; ++a)
      if (tid >= bI.rs[a] && tid < bI.rs[a + 1]) p = a;
    rowpos[tid] = (int8_t)p;
    rowin[tid] = (int8_t)(p >= 0 ? tid - bI.rs[p] : 0);
  } else if (tid < 2 * kBandRows) {
    const int c = tid - kBandRows;
    int p = -1;
    for (int a = 0; a < bJ.natoms; ++a)
      if (c >= bJ.rs[a] && c < bJ.rs[a + 1]) p = a;
    colpos[c] = (int8_t)p;
    colin[c] = (int8_t)(p >= 0 ? c - bJ.rs[p] : 0);
  } else if (tid < 2 * kBandRows + kBandAtoms) {
    const int p = tid - 2 * kBandRows;
    uint32_t g = 0;
    if (p < bI.natoms && bI.rs[p + 1] > bI.rs[p])
      for (int q = bI.rs[p] / 4; q <= (bI.rs[p + 1] - 1) / 4; ++q) g |= 1u << q;
    pg_r[p] = (uint16_t)g;
  } else if (tid < 2 * kBandRows + 2 * kBandAtoms) {
    const int p = tid - 2 * kBandRows - kBandAtoms;
    uint32_t g = 0;
    if (p < bJ.natoms && bJ.rs[p + 1] > bJ.rs[p])
      for (int q = bJ.rs[p] / 4; q <= (bJ.rs[p + 1] - 1) / 4; ++q) g |= 1u << q;
    pg_c[p] = (uint16_t)g;
  }
  if (tid == 0) e_cursor = T.e0;
  if (tid < kBandAtoms) keptw[tid] = 0;
  __syncthreads();

  double acc[16];
#pragma unroll
  for (int x = 0; x < 16; ++x) acc[x] = 0.0;

  const int nchunks = (T.ktot + KC - 1) / KC;
  const int e_end = T.e0 + T.ne;
  for (int ch = 0; ch < nchunks; ++ch) {
    const int Q0 = ch * KC;
    if (wave == 0) {
      const int q = lane;
      const int Q = Q0 + q;
      const bool valid = q < KC && Q < T.ktot;
      uint32_t gr = 0, gc = 0;
      int e = e_cursor;
      if (valid) {
        while (e + 1 < e_end && entries[e + 1].kstart <= Q) ++e;
        const Entry en = entries[e];
        const int k = a_strips[en.a_strip].k;
        ColInfo ci;
        ci.a_strip = en.a_strip;
        ci.b_strip = en.b_strip;
        ci.kk = (int16_t)(Q - en.kstart);
        ci.kdim = (int16_t)blk[k];
        ci.rmask = en.rmask;
        ci.cmask = en.cmask;
        colinfo[q] = ci;
        for (int p = 0; p < kBandAtoms; ++p) {
          if ((en.rmask >> p) & 1) gr |= pg_r[p];
          if ((en.cmask >> p) & 1) gc |= pg_c[p];
        }
      } else if (q < KC) {
        ColInfo ci;
        ci.a_strip = 0;
        ci.b_strip = 0;
        ci.kk = 0;
        ci.kdim = 0;
        ci.rmask = 0;
        ci.cmask = 0;
        colinfo[q] = ci;
      }
      // OR the group masks over each group of 4 K columns
      gr |= __shfl_xor(gr, 1, 64);
      gr |= __shfl_xor(gr, 2, 64);
      gc |= __shfl_xor(gc, 1, 64);
      gc |= __shfl_xor(gc, 2, 64);
      if (q < KC && (q & 3) == 0) {
        smask_r[q >> 2] = gr;
        smask_c[q >> 2] = gc;
      }
      const int last = min(KC, T.ktot - Q0) - 1;
      if (q == last) e_cursor = e;
    }
    __syncthreads();
    // stage A: As[q][r]
    {
      const int r = tid & 63;
      const int pos = rowpos[r];
#pragma unroll
      for (int p = 0; p < KC / 4; ++p) {
        const int q = (tid >> 6) + 4 * p;
        const ColInfo ci = colinfo[q];
        double v = 0.0;
        if (pos >= 0 && ((ci.rmask >> pos) & 1)) {
          const int m = rs_r[pos + 1] - rs_r[pos];
          const int64_t o = a_strips[ci.a_strip].off[pos];
          v = a_data[o + (int64_t)ci.kk * m + rowin[r]];
        }
        As[q * kPitch + r] = v;
      }
    }
    // stage B: Bs[q][c]
    {
      const int q = tid & 31;
      const ColInfo ci = colinfo[q];
#pragma unroll
      for (int p = 0; p < kBandRows / 8; ++p) {
        const int c = (tid >> 5) + 8 * p;
        const int pos = colpos[c];
        double v = 0.0;
        if (pos >= 0 && ((ci.cmask >> pos) & 1)) {
          const int64_t o = b_strips[ci.b_strip].off[pos];
          v = b_data[o + (int64_t)colin[c] * ci.kdim + ci.kk];
        }
        Bs[q * kPitch + c] = v;
      }
    }
    __syncthreads();
    const int nsteps = (min(KC, T.ktot - Q0) + 3) / 4;
    for (int s = 0; s < nsteps; ++s) {
      const uint32_t mr = __builtin_amdgcn_readfirstlane(smask_r[s]);
      const uint32_t mc = __builtin_amdgcn_readfirstlane(smask_c[s]);
      const int krow = (4 * s + (lane >> 4)) * kPitch;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (((mr >> (8 * wr + 4 * h)) & 0xFu) == 0) continue;
        const double a_op = As[krow + 32 * wr + 16 * h + (lane & 15)];
#pragma unroll
        for (int g = 0; g < 8; ++g) {
          if (((mc >> (8 * wc + g)) & 1u) == 0) continue;
          const double b_op = Bs[krow + 32 * wc + 4 * g + (lane & 3)];
          acc[h * 8 + g] = __builtin_amdgcn_mfma_f64_4x4x4f64(a_op, b_op, acc[h * 8 + g], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }

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
  const bool present = ip < bI.natoms && jp < bJ.natoms && ((T.pattern[ip] >> jp) & 1);
  const int64_t sz = present ? (int64_t)(bI.rs[ip + 1] - bI.rs[ip]) * (bJ.rs[jp + 1] - bJ.rs[jp]) : 0;
  // wave-inclusive scans of (present, size)
  int cnt = present ? 1 : 0;
  int64_t acc_sz = sz;
  for (int o = 1; o < 64; o <<= 1) {
    const int c2 = __shfl_up(cnt, o, 64);
    const int64_t s2 = __shfl_up(acc_sz, o, 64);
    if (lane >= o) {
      cnt += c2;
      acc_sz += s2;
    }
  }
  __shared__ int64_t wave_sz[4];
  if (lane == 63) {
    wave_tot[wave] = cnt;
    wave_sz[wave] = acc_sz;
  }
  __syncthreads();
  int cbase = 0;
  int64_t sbase = 0;
  for (int w = 0; w < wave; ++w) {
    cbase += wave_tot[w];
    sbase += wave_sz[w];
  }
  const int npairs = wave_tot[0] + wave_tot[1] + wave_tot[2] + wave_tot[3];
  if (present) {
    pair_id[cbase + cnt - 1] = tid;
    pair_off[cbase + cnt - 1] = sbase + acc_sz - sz;
  }
  __syncthreads();

  const int32_t atom_i0 = bI.first, atom_j0 = bJ.first;
  for (int pi = wave; pi < npairs; pi += 4) {
    const int id = pair_id[pi];
    const int pi_ = id >> 4, pj = id & 15;
    const int m = rs_r[pi_ + 1] - rs_r[pi_];
    const int n = rs_c[pj + 1] - rs_c[pj];
    const int len = m * n;
    const int64_t base = T.out_base + pair_off[pi];
    const double* old = nullptr;
    if (beta != 0.0 && old_row_p != nullptr) {
      const int32_t gi = atom_i0 + pi_, gj = atom_j0 + pj;
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
      const int c = e / m, r = e - c * m;
      double v = alpha * Ct[(rs_r[pi_] + r) * kPitch + rs_c[pj] + c];
      if (old) v = v + beta * old[e];
      c_data[base + e] = v;
      part = __builtin_fma(v, v, part);
    }
    const double nsq = wave_sum(part);
    if (lane == 0 && nsq >= eps2) atomicOr(&keptw[pi_], 1u << pj);
  }
  __syncthreads();
  if (tid < kBandAtoms) kept_out[(int64_t)blockIdx.x * kBandAtoms + tid] = (uint16_t)keptw[tid];
}

hipError_t launch_numeric_f64(hipStream_t s, int32_t n_tiles, const TileDesc* tiles,
                              const Entry* entries, const Strip* a_strips,
                              const Strip* b_strips, const Band* bands, const int32_t* d_blk,
                              const double* a_data, const double* b_data, double alpha,
                              double beta, double eps2, const int64_t* old_row_p,
                              const int32_t* old_col, const int64_t* old_off,
                              const double* old_data, double* c_data, uint16_t* kept) {
  if (n_tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(numeric_f64_kernel, dim3((unsigned)n_tiles), dim3(256), 0, s, tiles,
                     entries, a_strips, b_strips, bands, d_blk, a_data, b_data, alpha, beta,
                     eps2, old_row_p, old_col, old_off, old_data, c_data, kept);
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
