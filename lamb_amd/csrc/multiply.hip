// Multiply driver: block norms, strips, tile lists (symbolic phase), the
// numeric kernel and the C index.
#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <numeric>
#include <vector>

#include "common.h"
#include "internal.h"

using namespace lamb;

#define HIPCHK(ctx, expr, what)                            \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return hip_error(ctx, e_, what); \
  } while (0)

// ---------------------------------------------------------------------------
// spatial order: recursive coordinate bisection along the longest extent
// ---------------------------------------------------------------------------
namespace {
void rcb(const float* xyz, int32_t* idx, int32_t n) {
  if (n <= 1) return;
  float lo[3] = {xyz[3 * idx[0]], xyz[3 * idx[0] + 1], xyz[3 * idx[0] + 2]};
  float hi[3] = {lo[0], lo[1], lo[2]};
  for (int32_t t = 1; t < n; ++t)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], xyz[3 * idx[t] + d]);
      hi[d] = std::max(hi[d], xyz[3 * idx[t] + d]);
    }
  int dim = 0;
  for (int d = 1; d < 3; ++d)
    if (hi[d] - lo[d] > hi[dim] - lo[dim]) dim = d;
  const int32_t mid = n / 2;
  std::nth_element(idx, idx + mid, idx + n, [&](int32_t a, int32_t b) {
    const float xa = xyz[3 * a + dim], xb = xyz[3 * b + dim];
    return xa != xb ? xa < xb : a < b;
  });
  rcb(xyz, idx, mid);
  rcb(xyz, idx + mid, n - mid);
}
}  // namespace

void lamb::spatial_order(int32_t n, const float* coords, std::vector<int32_t>& order) {
  order.resize(n);
  std::iota(order.begin(), order.end(), 0);
  rcb(coords, order.data(), n);
}

// ---------------------------------------------------------------------------
// strips
// ---------------------------------------------------------------------------
static int32_t ensure_nsq(lamb_hip_mat* m) {
  lamb_hip_ctx* ctx = m->dist->ctx;
  if (m->nsq_valid) return LAMB_HIP_OK;
  HIPCHK(ctx, m->d_nsq.ensure(sizeof(double) * std::max<int64_t>(m->nblk, 1)), "nsq alloc");
  HIPCHK(ctx, launch_block_nsq(ctx->stream, m->nblk, (const int32_t*)m->d_rows.p,
                               (const int32_t*)m->d_col.p, (const int32_t*)m->dist->d_blk.p,
                               (const int64_t*)m->d_off.p, (const double*)m->d_data.p,
                               (double*)m->d_nsq.p), "nsq kernel");
  m->nsq.resize((size_t)m->nblk);
  if (m->nblk)
    HIPCHK(ctx, hipMemcpyAsync(m->nsq.data(), m->d_nsq.p, sizeof(double) * m->nblk,
                               hipMemcpyDeviceToHost, ctx->stream), "nsq download");
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "nsq sync");
  m->nsq_valid = true;
  return LAMB_HIP_OK;
}

// A strips: per row band I, the blocks of its rows merged by column atom k.
static void build_a_strips(const lamb_hip_mat* A, std::vector<Strip>& strips,
                           std::vector<int64_t>& band_ptr) {
  const lamb_hip_dist* d = A->dist;
  const int32_t nb = (int32_t)d->bands.size();
  strips.clear();
  band_ptr.assign(nb + 1, 0);
  std::vector<int64_t> cur(kBandAtoms), end(kBandAtoms);
  for (int32_t I = 0; I < nb; ++I) {
    const Band& b = d->bands[I];
    for (int p = 0; p < b.natoms; ++p) {
      cur[p] = A->row_p[b.first + p];
      end[p] = A->row_p[b.first + p + 1];
    }
    while (true) {
      int32_t kmin = INT32_MAX;
      for (int p = 0; p < b.natoms; ++p)
        if (cur[p] < end[p]) kmin = std::min(kmin, A->col[cur[p]]);
      if (kmin == INT32_MAX) break;
      Strip s{};
      s.k = kmin;
      s.band = I;
      s.min_nsq = INFINITY;
      for (int p = 0; p < b.natoms; ++p) {
        s.off[p] = -1;
        if (cur[p] < end[p] && A->col[cur[p]] == kmin) {
          s.mask |= (uint16_t)(1u << p);
          s.off[p] = A->off[cur[p]];
          s.nsq[p] = A->nsq[cur[p]];
          s.min_nsq = std::min(s.min_nsq, s.nsq[p]);
          ++cur[p];
        }
      }
      strips.push_back(s);
    }
    band_ptr[I + 1] = (int64_t)strips.size();
  }
}

// B strips: per row atom k, runs of its blocks that fall in one col band J.
static void build_b_strips(const lamb_hip_mat* B, std::vector<Strip>& strips,
                           std::vector<int64_t>& row_ptr) {
  const lamb_hip_dist* d = B->dist;
  strips.clear();
  row_ptr.assign(d->n + 1, 0);
  for (int32_t k = 0; k < d->n; ++k) {
    int64_t b = B->row_p[k];
    const int64_t e = B->row_p[k + 1];
    while (b < e) {
      const int32_t J = d->band_of[B->col[b]];
      Strip s{};
      s.k = k;
      s.band = J;
      s.min_nsq = INFINITY;
      for (int p = 0; p < kBandAtoms; ++p) s.off[p] = -1;
      while (b < e && d->band_of[B->col[b]] == J) {
        const int p = d->pos_in_band[B->col[b]];
        s.mask |= (uint16_t)(1u << p);
        s.off[p] = B->off[b];
        s.nsq[p] = B->nsq[b];
        s.min_nsq = std::min(s.min_nsq, s.nsq[p]);
        ++b;
      }
      strips.push_back(s);
    }
    row_ptr[k + 1] = (int64_t)strips.size();
  }
}

static inline int popc(uint32_t x) { return __builtin_popcount(x); }

extern "C" int32_t lamb_hip_multiply(lamb_hip_ctx* ctx, const lamb_hip_mat* A_,
                                     const lamb_hip_mat* B_, lamb_hip_mat* C, double alpha,
                                     double beta, double eps, int32_t flags,
                                     lamb_hip_stats* stats) {
  auto t_start = std::chrono::steady_clock::now();
  if (!ctx || !A_ || !B_ || !C) return LAMB_HIP_FAILURE;
  lamb_hip_mat* A = const_cast<lamb_hip_mat*>(A_);
  lamb_hip_mat* B = const_cast<lamb_hip_mat*>(B_);
  lamb_hip_dist* d = A->dist;
  if (B->dist != d || C->dist != d)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "multiply: A, B, C must share one distribution");
  if (C == A || C == B)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS, "multiply: C aliases A or B");
  if (flags != LAMB_HIP_FP64)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_UNIMPLEMENTED,
                     "multiply: flags=%d not implemented", flags);
  if (d->max_blk > kBandRows)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_UNIMPLEMENTED,
                     "multiply: block dimension %d above %d", d->max_blk, kBandRows);
  if (!(eps >= 0.0))
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS, "multiply: eps_filter < 0");
  (void)hipSetDevice(ctx->device);
  int32_t rc;
  if ((rc = ensure_nsq(A))) return rc;
  if ((rc = ensure_nsq(B))) return rc;
  const double eps2 = eps * eps;
  const bool use_old = beta != 0.0 && C->nblk > 0;

  // ---- strips ----
  std::vector<Strip> as, bs;
  std::vector<int64_t> a_ptr, b_ptr;
  build_a_strips(A, as, a_ptr);
  build_b_strips(B, bs, b_ptr);
  // packed panel offsets: [kdim][64] doubles per strip
  std::vector<int64_t> a_pk(as.size() + 1, 0), b_pk(bs.size() + 1, 0);
  for (size_t t = 0; t < as.size(); ++t) a_pk[t + 1] = a_pk[t] + 64 * (int64_t)d->blk_int[as[t].k];
  for (size_t t = 0; t < bs.size(); ++t) b_pk[t + 1] = b_pk[t] + 64 * (int64_t)d->blk_int[bs[t].k];

  // ---- symbolic: tiles and their (A strip, B strip) entries ----
  const int32_t nbands = (int32_t)d->bands.size();
  std::vector<TileDesc> tiles;
  std::vector<Entry> entries;
  std::vector<std::vector<Entry>> bucket(nbands);
  std::vector<int32_t> ktot(nbands, 0);
  std::vector<std::array<uint16_t, kBandAtoms>> pat(nbands);
  std::vector<char> touched_flag(nbands, 0);
  std::vector<int32_t> touched;
  int64_t kept = 0, filtered = 0, nstruct = 0;
  double flops = 0.0;
  int64_t c_size = 0;
  auto rows_of = [&](const Band& b, uint32_t mask) {
    int r = 0;
    for (int p = 0; p < b.natoms; ++p)
      if ((mask >> p) & 1) r += b.rs[p + 1] - b.rs[p];
    return r;
  };
  for (int32_t I = 0; I < nbands; ++I) {
    const Band& bI = d->bands[I];
    touched.clear();
    auto touch = [&](int32_t J) {
      if (!touched_flag[J]) {
        touched_flag[J] = 1;
        touched.push_back(J);
        bucket[J].clear();
        ktot[J] = 0;
        pat[J].fill(0);
      }
    };
    auto emit = [&](int32_t sa, int32_t sb, int32_t J, int kdim, uint16_t rm, uint16_t cm) {
      if (kdim > 0) {
        Entry e{};
        e.a_col = (uint32_t)(a_pk[sa] / 64);
        e.b_col = (uint32_t)(b_pk[sb] / 64);
        e.kstart = ktot[J];
        e.rmask = rm;
        e.cmask = cm;
        bucket[J].push_back(e);
        ktot[J] += kdim;
      }
      for (int p = 0; p < kBandAtoms; ++p)
        if ((rm >> p) & 1) pat[J][p] |= cm;
      kept += (int64_t)popc(rm) * popc(cm);
      flops += 2.0 * rows_of(bI, rm) * kdim * rows_of(d->bands[J], cm);
    };
    for (int64_t sa = a_ptr[I]; sa < a_ptr[I + 1]; ++sa) {
      const Strip& SA = as[sa];
      const int kdim = d->blk_int[SA.k];
      for (int64_t sb = b_ptr[SA.k]; sb < b_ptr[SA.k + 1]; ++sb) {
        const Strip& SB = bs[sb];
        const int32_t J = SB.band;
        if (SA.min_nsq * SB.min_nsq >= eps2) {
          touch(J);
          emit((int32_t)sa, (int32_t)sb, J, kdim, SA.mask, SB.mask);
          continue;
        }
        // exact per-(i,j) prefilter: group columns by their kept-row set
        uint16_t rowset[kBandAtoms];
        int ncols = 0;
        uint16_t colsets[kBandAtoms];
        uint16_t keys[kBandAtoms];
        for (int q = 0; q < kBandAtoms; ++q) {
          if (!((SB.mask >> q) & 1)) continue;
          uint16_t r = 0;
          for (int p = 0; p < kBandAtoms; ++p)
            if (((SA.mask >> p) & 1) && SA.nsq[p] * SB.nsq[q] >= eps2) r |= (uint16_t)(1u << p);
          filtered += popc(SA.mask) - popc(r);
          rowset[q] = r;
          if (!r) continue;
          int g = 0;
          while (g < ncols && keys[g] != r) ++g;
          if (g == ncols) {
            keys[ncols] = r;
            colsets[ncols] = 0;
            ++ncols;
          }
          colsets[g] |= (uint16_t)(1u << q);
        }
        (void)rowset;
        if (ncols) touch(J);
        for (int g = 0; g < ncols; ++g) emit((int32_t)sa, (int32_t)sb, J, kdim, keys[g], colsets[g]);
      }
    }
    if (use_old) {
      for (int p = 0; p < bI.natoms; ++p) {
        const int32_t i = bI.first + p;
        for (int64_t b = C->row_p[i]; b < C->row_p[i + 1]; ++b) {
          const int32_t j = C->col[b];
          const int32_t J = d->band_of[j];
          touch(J);
          pat[J][p] |= (uint16_t)(1u << d->pos_in_band[j]);
        }
      }
    }
    std::sort(touched.begin(), touched.end());
    for (int32_t J : touched) {
      touched_flag[J] = 0;
      const Band& bJ = d->bands[J];
      TileDesc t{};
      t.I = I;
      t.J = J;
      t.e0 = (int32_t)entries.size();
      t.ne = (int32_t)bucket[J].size();
      t.ktot = ktot[J];
      t.out_base = c_size;
      for (int p = 0; p < kBandAtoms; ++p) {
        t.pattern[p] = pat[J][p];
        if (p < bI.natoms)
          for (int q = 0; q < bJ.natoms; ++q)
            if ((pat[J][p] >> q) & 1) {
              c_size += (int64_t)(bI.rs[p + 1] - bI.rs[p]) * (bJ.rs[q + 1] - bJ.rs[q]);
              ++nstruct;
            }
      }
      entries.insert(entries.end(), bucket[J].begin(), bucket[J].end());
      tiles.push_back(t);
    }
  }
  if (entries.size() > (size_t)INT32_MAX || as.size() > (size_t)INT32_MAX ||
      a_pk.back() / 64 > (int64_t)UINT32_MAX || b_pk.back() / 64 > (int64_t)UINT32_MAX ||
      bs.size() > (size_t)INT32_MAX || tiles.size() > (size_t)INT32_MAX)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_OVERFLOW_ERROR,
                     "multiply: symbolic sizes exceed 32-bit indices");

  // ---- device buffers ----
  hipStream_t s = ctx->stream;
  const int32_t n_tiles = (int32_t)tiles.size();
  HIPCHK(ctx, ctx->d_tiles.upload(s, tiles.data(), sizeof(TileDesc) * tiles.size()), "tiles");
  HIPCHK(ctx, ctx->d_entries.upload(s, entries.data(), sizeof(Entry) * entries.size()), "entries");
  HIPCHK(ctx, ctx->d_astrips.upload(s, as.data(), sizeof(Strip) * as.size()), "a strips");
  HIPCHK(ctx, ctx->d_bstrips.upload(s, bs.data(), sizeof(Strip) * bs.size()), "b strips");
  HIPCHK(ctx, ctx->d_apk_off.upload(s, a_pk.data(), sizeof(int64_t) * a_pk.size()), "a pk off");
  HIPCHK(ctx, ctx->d_bpk_off.upload(s, b_pk.data(), sizeof(int64_t) * b_pk.size()), "b pk off");
  HIPCHK(ctx, ctx->d_apack.ensure(sizeof(double) * (size_t)std::max<int64_t>(a_pk.back(), 1)),
         "a pack alloc");
  HIPCHK(ctx, ctx->d_bpack.ensure(sizeof(double) * (size_t)std::max<int64_t>(b_pk.back(), 1)),
         "b pack alloc");
  HIPCHK(ctx, hipEventRecord(ctx->ev[2], s), "event");
  HIPCHK(ctx, launch_pack_strips(s, (int64_t)as.size(), (const Strip*)ctx->d_astrips.p,
                                 (const int64_t*)ctx->d_apk_off.p, (const Band*)d->d_bands.p,
                                 (const int32_t*)d->d_blk.p, (const double*)A->d_data.p, true,
                                 (double*)ctx->d_apack.p), "pack A");
  HIPCHK(ctx, launch_pack_strips(s, (int64_t)bs.size(), (const Strip*)ctx->d_bstrips.p,
                                 (const int64_t*)ctx->d_bpk_off.p, (const Band*)d->d_bands.p,
                                 (const int32_t*)d->d_blk.p, (const double*)B->d_data.p, false,
                                 (double*)ctx->d_bpack.p), "pack B");
  HIPCHK(ctx, hipEventRecord(ctx->ev[3], s), "event");
  HIPCHK(ctx, ctx->d_kept.ensure(sizeof(uint16_t) * kBandAtoms * std::max(n_tiles, 1)), "kept");
  DevBuf newdata;
  HIPCHK(ctx, newdata.ensure(sizeof(double) * (size_t)std::max<int64_t>(c_size, 1)), "C pool");

  HIPCHK(ctx, hipEventRecord(ctx->ev[0], s), "event");
  HIPCHK(ctx, launch_numeric_f64(s, n_tiles, (const TileDesc*)ctx->d_tiles.p,
                                 (const Entry*)ctx->d_entries.p, (const Band*)d->d_bands.p,
                                 (const double*)ctx->d_apack.p, (const double*)ctx->d_bpack.p,
                                 alpha, beta, eps2,
                                 use_old ? (const int64_t*)C->d_row_p.p : nullptr,
                                 use_old ? (const int32_t*)C->d_col.p : nullptr,
                                 use_old ? (const int64_t*)C->d_off.p : nullptr,
                                 use_old ? (const double*)C->d_data.p : nullptr,
                                 (double*)newdata.p, (uint16_t*)ctx->d_kept.p),
         "numeric kernel");
  HIPCHK(ctx, hipEventRecord(ctx->ev[1], s), "event");
  std::vector<uint16_t> keptmask((size_t)n_tiles * kBandAtoms);
  if (n_tiles)
    HIPCHK(ctx, hipMemcpyAsync(keptmask.data(), ctx->d_kept.p, sizeof(uint16_t) * keptmask.size(),
                               hipMemcpyDeviceToHost, s), "kept download");
  HIPCHK(ctx, hipStreamSynchronize(s), "numeric sync");
  float ms_num = 0.f, ms_pack = 0.f;
  HIPCHK(ctx, hipEventElapsedTime(&ms_num, ctx->ev[0], ctx->ev[1]), "event time");
  HIPCHK(ctx, hipEventElapsedTime(&ms_pack, ctx->ev[2], ctx->ev[3]), "event time");

  // ---- C index (internal CSR, rows ascending, cols ascending) ----
  std::vector<int64_t> row_p(d->n + 1, 0);
  std::vector<int32_t> col;
  std::vector<int64_t> off;
  std::vector<int32_t> rows;
  int64_t nelems = 0;
  {
    size_t t0 = 0;
    std::vector<int64_t> pairoff((size_t)kBandAtoms * kBandAtoms);
    std::vector<std::vector<int64_t>> tile_pairoff;
    while (t0 < tiles.size()) {
      const int32_t I = tiles[t0].I;
      size_t t1 = t0;
      while (t1 < tiles.size() && tiles[t1].I == I) ++t1;
      const Band& bI = d->bands[I];
      // pair offsets per tile of this band
      tile_pairoff.assign(t1 - t0, std::vector<int64_t>((size_t)kBandAtoms * kBandAtoms, -1));
      for (size_t t = t0; t < t1; ++t) {
        const Band& bJ = d->bands[tiles[t].J];
        int64_t o = 0;
        for (int p = 0; p < bI.natoms; ++p)
          for (int q = 0; q < bJ.natoms; ++q)
            if ((tiles[t].pattern[p] >> q) & 1) {
              tile_pairoff[t - t0][p * kBandAtoms + q] = o;
              o += (int64_t)(bI.rs[p + 1] - bI.rs[p]) * (bJ.rs[q + 1] - bJ.rs[q]);
            }
      }
      for (int p = 0; p < bI.natoms; ++p) {
        const int32_t i = bI.first + p;
        const int m = bI.rs[p + 1] - bI.rs[p];
        for (size_t t = t0; t < t1; ++t) {
          const Band& bJ = d->bands[tiles[t].J];
          const uint16_t km = keptmask[t * kBandAtoms + p];
          for (int q = 0; q < bJ.natoms; ++q)
            if ((km >> q) & 1) {
              col.push_back(bJ.first + q);
              off.push_back(tiles[t].out_base + tile_pairoff[t - t0][p * kBandAtoms + q]);
              rows.push_back(i);
              nelems += (int64_t)m * (bJ.rs[q + 1] - bJ.rs[q]);
            }
        }
        row_p[i + 1] = (int64_t)col.size();
      }
      t0 = t1;
    }
    for (int32_t i = 0; i < d->n; ++i) row_p[i + 1] = std::max(row_p[i + 1], row_p[i]);
  }
  // swap in the new C
  C->d_data.release();
  C->d_data = newdata;
  newdata.p = nullptr;
  newdata.bytes = 0;
  C->row_p.swap(row_p);
  C->col.swap(col);
  C->off.swap(off);
  C->rows.swap(rows);
  C->nblk = (int64_t)C->col.size();
  C->ndata = c_size;
  C->nelems = nelems;
  C->invalidate_derived();
  HIPCHK(ctx, C->upload_index(s), "C index upload");
  HIPCHK(ctx, hipStreamSynchronize(s), "C index sync");

  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->n_triples_kept = kept;
    stats->n_triples_filtered = filtered;
    stats->n_c_blocks_struct = nstruct;
    stats->n_c_blocks_kept = C->nblk;
    stats->n_c_elems_kept = nelems;
    stats->n_tiles = n_tiles;
    stats->n_tile_entries = (int64_t)entries.size();
    stats->flops = flops;
    stats->ms_numeric = ms_num;
    stats->ms_symbolic = ms_pack;
    stats->ms_total = std::chrono::duration<double, std::milli>(
                          std::chrono::steady_clock::now() - t_start).count();
  }
  return LAMB_HIP_OK;
}
