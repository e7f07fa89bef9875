// Multiply driver: block norms, strips, tile lists (symbolic phase), the
// numeric kernel and the C index.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <vector>

#include "common.h"
#include "internal.h"
#include "symbolic.h"

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
// multiply driver (everything on the device; the host only reads the few
// totals it needs to size the next allocation)
// ---------------------------------------------------------------------------
namespace lamb {
struct SymState {
  SymScratch W;
  StripSet sa, sb;
  SymResult r;
  void release() {
    W.release();
    sa.release();
    sb.release();
    r.release();
  }
};
}  // namespace lamb

static SymState& sym_state(lamb_hip_ctx* ctx) {
  if (!ctx->sym) ctx->sym = new SymState();
  return *static_cast<SymState*>(ctx->sym);
}

void lamb::release_sym_state(lamb_hip_ctx* ctx) {
  if (ctx->sym) {
    static_cast<SymState*>(ctx->sym)->release();
    delete static_cast<SymState*>(ctx->sym);
    ctx->sym = nullptr;
  }
}

static int32_t ensure_nsq(lamb_hip_mat* m) {
  lamb_hip_ctx* ctx = m->dist->ctx;
  if (m->nsq_valid) return LAMB_HIP_OK;
  HIPCHK(ctx, m->d_nsq.ensure(sizeof(double) * std::max<int64_t>(m->nblk, 1)), "nsq alloc");
  HIPCHK(ctx, launch_block_nsq(ctx->stream, m->nblk, (const int32_t*)m->d_rows.p,
                               (const int32_t*)m->d_col.p, (const int32_t*)m->dist->d_blk.p,
                               (const int64_t*)m->d_off.p, (const double*)m->d_data.p,
                               (double*)m->d_nsq.p), "nsq kernel");
  m->nsq_valid = true;
  return LAMB_HIP_OK;
}

bool lamb::values_missing(const lamb_hip_mat* m) {
  return m->nblk > 0 && m->ndata > 0 && m->d_data.p == nullptr;
}

int32_t lamb::refuse_missing_values(lamb_hip_ctx* ctx, const char* what) {
  return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_PRECONDITION_FAILED,
                   "%s: the matrix's values were released", what);
}

// phase timing events (not recorded while a graph is being captured)
static hipError_t record_event(lamb_hip_ctx* ctx, int k) {
  if (ctx->size_mode == kSizesReplay) return hipSuccess;
  return hipEventRecord(ctx->ev[k], ctx->stream);
}

// FP64 kernel variant forced by LAMB_HIP_F64_KERNEL (quadrant | half | ksplit |
// lanes | fine), or -1
static int f64_forced() {
  const char* v = getenv("LAMB_HIP_F64_KERNEL");
  if (v && strcmp(v, "quadrant") == 0) return kF64Quadrant;
  if (v && strcmp(v, "ksplit") == 0) return kF64Ksplit;
  if (v && strcmp(v, "half") == 0) return kF64Half;
  if (v && strcmp(v, "lanes") == 0) return kF64Lanes;
  if (v && strcmp(v, "fine") == 0) return kF64Fine;
  return -1;
}

// The FP64 variant of a product: a function of the workload only, so that a
// product's bits do not depend on timings (the variants add K products in
// different orders). Tiles of a dense product (flops per tile at or above
// 1e7, e.g. the iterates of a purification chain, about 1.8e8) take the
// K-split kernel when no entry needs per-lane masks; every other product the
// half kernel. Measured: 6,144-atom water box (9.0e5 flop per tile) half
// 16.0 ms, quadrant 17.8, K-split 20.3; a dense 16 A product of 3,072 atoms
// K-split 251 ms, half 279, quadrant 344.
constexpr double kDenseTileFlops = 1.0e7;
static int f64_rule(double flops_per_tile, bool whole) {
  return (whole && flops_per_tile >= kDenseTileFlops) ? kF64Ksplit : kF64Half;
}

// Tile-order block sizes (row bands per block) the autotuner times. The order
// only maps tiles to workgroups, so every candidate writes the same bits.
static const int kRowBlocks[] = {4, 8, 16};

// half-octave bucket of a positive size
static int64_t bucket(double x) { return (int64_t)std::floor(2.0 * std::log2(std::max(x, 1.0))); }

static DistView dist_view(const lamb_hip_dist* d) {
  DistView v;
  v.n = d->n;
  v.nbands = (int32_t)d->bands.size();
  v.blk = (const int32_t*)d->d_blk.p;
  v.bands = (const Band*)d->d_bands.p;
  v.band_of = (const int32_t*)d->d_band_of.p;
  v.pos_in_band = (const int32_t*)d->d_pos_in_band.p;
  return v;
}

static MatView mat_view(const lamb_hip_mat* m) {
  MatView v;
  v.nblk = m->nblk;
  v.row_p = (const int64_t*)m->d_row_p.p;
  v.col = (const int32_t*)m->d_col.p;
  v.off = (const int64_t*)m->d_off.p;
  v.rows = (const int32_t*)m->d_rows.p;
  v.nsq = (const double*)m->d_nsq.p;
  return v;
}

extern "C" int32_t lamb_hip_multiply(lamb_hip_ctx* ctx, const lamb_hip_mat* A,
                                     const lamb_hip_mat* B, lamb_hip_mat* C, double alpha,
                                     double beta, double eps, int32_t flags,
                                     lamb_hip_stats* stats) {
  return lamb::multiply_masked(ctx, A, B, C, alpha, beta, eps, flags, stats, nullptr);
}

int32_t lamb::multiply_masked(lamb_hip_ctx* ctx, const lamb_hip_mat* A_, const lamb_hip_mat* B_,
                              lamb_hip_mat* C, double alpha, double beta, double eps,
                              int32_t flags, lamb_hip_stats* stats, const lamb_hip_mat* mask,
                              double post_eps) {
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
  if (values_missing(A) || values_missing(B) || (beta != 0.0 && values_missing(C)))
    return refuse_missing_values(ctx, "multiply");
  if (mask && (mask->dist != d || mask == C))
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "multiply: the mask must share the distribution and not be C");
  const int mode = flags & 0xFF;
  int slices = (flags >> 8) & 0xF;
  if ((flags & ~0xFFF) != 0 || (mode != LAMB_HIP_FP64 && mode != LAMB_HIP_BF16_SPLIT) ||
      (mode == LAMB_HIP_FP64 && slices != 0))
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_UNIMPLEMENTED,
                     "multiply: flags=0x%x not implemented", flags);
  const bool split = mode == LAMB_HIP_BF16_SPLIT;
  if (split && slices == 0) slices = 5;
  if (split && (slices < 3 || slices > 5))
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "multiply: BF16 split needs 3..5 slices, got %d", slices);
  if (d->max_blk > kBandRows)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_UNIMPLEMENTED,
                     "multiply: block dimension %d above %d", d->max_blk, kBandRows);
  if (!(eps >= 0.0))
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS, "multiply: eps_filter < 0");
  (void)hipSetDevice(ctx->device);
  hipStream_t s = ctx->stream;
  int32_t rc;
  if ((rc = ensure_nsq(A))) return rc;
  if ((rc = ensure_nsq(B))) return rc;
  const double eps2 = eps * eps;
  const double post2 = post_eps < 0.0 ? eps2 : post_eps * post_eps;
  const bool use_old = beta != 0.0 && C->nblk > 0;
  SymState& S = sym_state(ctx);
  const DistView D = dist_view(d);
  const MatView MA = mat_view(A), MB = mat_view(B), MC = mat_view(C);

  HIPCHK(ctx, record_event(ctx, 2), "event");
  // ---- strips and packed panels ----
  HIPCHK(ctx, build_strips(ctx, S.W, MA, MB, D, S.sa, S.sb), "strips");
  // panel columns are addressed by 32-bit indices (31 bits on the A side,
  // whose top bit flags partial entries); one zero column follows the last
  // panel of each side
  if (S.sa.pk_total / 64 >= (int64_t)INT32_MAX || S.sb.pk_total / 64 >= (int64_t)UINT32_MAX ||
      (int64_t)slices * (S.sa.pk_total / 512 + 1) * 64 >= (int64_t)INT32_MAX * 64)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_OVERFLOW_ERROR,
                     "multiply: packed panels exceed 32-bit column indices");
  uint32_t a_zero, b_zero;
  if (!split) {
    a_zero = (uint32_t)(S.sa.pk_total / 64);
    b_zero = (uint32_t)(S.sb.pk_total / 64);
    HIPCHK(ctx, ctx->d_apack.ensure(sizeof(double) * (size_t)(S.sa.pk_total + 64)),
           "a pack alloc");
    HIPCHK(ctx, ctx->d_bpack.ensure(sizeof(double) * (size_t)(S.sb.pk_total + 64)),
           "b pack alloc");
    HIPCHK(ctx, launch_pack_strips(s, S.sa.n, (const Strip*)S.sa.strips.p,
                                   (const int64_t*)S.sa.pk.p, D.bands, D.blk,
                                   (const double*)A->d_data.p, true, (double*)ctx->d_apack.p),
           "pack A");
    HIPCHK(ctx, launch_pack_strips(s, S.sb.n, (const Strip*)S.sb.strips.p,
                                   (const int64_t*)S.sb.pk.p, D.bands, D.blk,
                                   (const double*)B->d_data.p, false, (double*)ctx->d_bpack.p),
           "pack B");
  } else {
    // exponent per A row and B column, then S BF16 slice chunks per strip
    // (chunks of 8 K columns; the chunk after the last one is zero)
    const int64_t nca = S.sa.pk_total / 512, ncb = S.sb.pk_total / 512;
    a_zero = (uint32_t)nca;
    b_zero = (uint32_t)ncb;
    const int64_t* elem0 = (const int64_t*)d->d_elem0.p;
    HIPCHK(ctx, ctx->d_aexp.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(d->n_elem, 1)),
           "a exponents alloc");
    HIPCHK(ctx, ctx->d_bexp.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(d->n_elem, 1)),
           "b exponents alloc");
    HIPCHK(ctx, launch_slice_exponents(s, MA.nblk, MA.rows, MA.col, D.blk, elem0, MA.off,
                                       (const double*)A->d_data.p, true, d->n_elem,
                                       (int32_t*)ctx->d_aexp.p), "A exponents");
    HIPCHK(ctx, launch_slice_exponents(s, MB.nblk, MB.rows, MB.col, D.blk, elem0, MB.off,
                                       (const double*)B->d_data.p, false, d->n_elem,
                                       (int32_t*)ctx->d_bexp.p), "B exponents");
    HIPCHK(ctx, ctx->d_aslice.ensure((size_t)slices * (size_t)(nca + 1) * 64 * 16), "a slices");
    HIPCHK(ctx, ctx->d_bslice.ensure((size_t)slices * (size_t)(ncb + 1) * 64 * 16), "b slices");
    HIPCHK(ctx, launch_pack_slices(s, slices, S.sa.n, (const Strip*)S.sa.strips.p,
                                   (const int64_t*)S.sa.pk.p, D.bands, D.blk, elem0,
                                   (const int32_t*)ctx->d_aexp.p, (const double*)A->d_data.p,
                                   true, nca, ctx->d_aslice.p), "slice A");
    HIPCHK(ctx, launch_pack_slices(s, slices, S.sb.n, (const Strip*)S.sb.strips.p,
                                   (const int64_t*)S.sb.pk.p, D.bands, D.blk, elem0,
                                   (const int32_t*)ctx->d_bexp.p, (const double*)B->d_data.p,
                                   false, ncb, ctx->d_bslice.p), "slice B");
  }
  // ---- tiles and entries ----
  const MatView MM = mask ? mat_view(mask) : MatView();
  HIPCHK(ctx, build_tiles(ctx, S.W, D, S.sa, S.sb, use_old ? &MC : nullptr, eps2, S.r,
                          mask ? &MM : nullptr),
         "tiles");
  if (S.r.n_entries > INT32_MAX || S.r.n_tiles > INT32_MAX)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_OVERFLOW_ERROR,
                     "multiply: symbolic sizes exceed 32-bit indices");
  // FP64 variant, tile order and launch form. The variant (the per-wave tile
  // shape) comes from the rule unless shape tuning is on or a loaded table
  // names it: variants add K products in different orders, so only an
  // explicit request lets timings choose it. The autotuner then times the
  // candidate tile orders and the three launch forms (one workgroup per
  // tile, a persistent grid over the interleaved order, a persistent grid
  // with one range of the order per XCD), on the first product of a workload
  // class (form,
  // variant, largest block, half-octave buckets of flops per tile and of C
  // blocks per tile) and keeps the fastest for the context's lifetime; a
  // graph capture with no tuned class uses 8 row bands per block and one
  // workgroup per tile. Order and launch form never change the bits.
  const int32_t n_tiles = (int32_t)S.r.n_tiles;
  const double nt = std::max<double>(1.0, (double)n_tiles);
  const bool whole = S.r.n_partial == 0;
  int variant = 0;
  // the per-wave tile shape (FP64 variant): forced, from the context's shape
  // table (shape tuning on, or a loaded table), or by the rule. A class is
  // (form, whole, the distribution's block sizes: the (m, n, k) shapes of
  // its blocks, half-octave buckets of flops and of C blocks per tile).
  const std::array<int64_t, 5> shape_key = {
      (int64_t)flags | ((int64_t)whole << 32), (int64_t)d->size_mask[0],
      (int64_t)d->size_mask[1], bucket(S.r.flops / nt), bucket((double)S.r.n_struct / nt)};
  bool tune_shape = false;
  if (!split) {
    const int forced = f64_forced();
    if (forced >= 0) {
      variant = forced;
    } else if (auto it = ctx->shape_tune.find(shape_key); it != ctx->shape_tune.end()) {
      variant = it->second;
    } else {
      variant = f64_rule(S.r.flops / nt, whole);
      tune_shape = ctx->tune_shapes != 0 && ctx->size_mode != kSizesReplay && n_tiles > 0;
    }
  }
  // the lanes kernel serves whole products with panels below 4 GB; any other
  // product forced or tuned to it runs the half kernel, and reports it
  const bool lanes_ok = whole && lanes_panels_fit(a_zero, b_zero);
  if ((variant == kF64Lanes || variant == kF64Fine) && !lanes_ok) variant = kF64Half;
  lamb_hip_ctx::TuneChoice choice;
  bool tune = false;
  const char* env_rb = getenv("LAMB_HIP_TILE_ROWBLOCK");
  const char* env_pe = getenv("LAMB_HIP_PERSISTENT");
  if (env_rb || env_pe) {
    if (env_rb) choice.rowblock = std::max(1, std::min(atoi(env_rb), 32));
    if (env_pe) choice.form = std::max(kFormGrid, std::min(atoi(env_pe), kFormXcdRanges));
  }
  HIPCHK(ctx, build_tile_order(ctx, S.W, D, S.r, choice.rowblock), "tile order");
  HIPCHK(ctx, record_event(ctx, 3), "event");

  // ---- numeric ----
  HIPCHK(ctx, ctx->d_kept.ensure(sizeof(uint16_t) * kBandAtoms * (size_t)std::max(n_tiles, 1)),
         "kept");
  // the new C goes into C's spare buffers, swapped in at the end (a graph
  // capture allocates and frees nothing: the spares are grown beforehand)
  DevBuf newdata = C->s_data;
  C->s_data = DevBuf();
  HIPCHK(ctx, newdata.ensure(sizeof(double) * (size_t)std::max<int64_t>(S.r.c_size, 1)),
         "C pool");
  HIPCHK(ctx, ctx->d_slot_counter.ensure(kSlotControlWords * sizeof(int32_t)), "slot counters");
  int32_t grid = 0;
  int32_t* slot_counter = (int32_t*)ctx->d_slot_counter.p;
  // every launch in the XCD ranges form first splits the current tile order
  // (three small kernels), so the bounds always match the order it walks
  auto launch_numeric = [&](int form) -> hipError_t {
    if (form == kFormXcdRanges) {
      const hipError_t e = build_xcd_ranges(ctx, S.W, S.r, slot_counter + kXcdRangeWords);
      if (e != hipSuccess) return e;
    }
    if (!split)
      return launch_numeric_f64(s, n_tiles, (const TileDesc*)S.r.tiles.p,
                                (const int32_t*)S.r.perm.p, (const Entry*)S.r.entries.p, D.bands,
                                (const double*)ctx->d_apack.p, (const double*)ctx->d_bpack.p,
                                a_zero, b_zero, alpha, beta, post2,
                                use_old ? MC.row_p : nullptr, use_old ? MC.col : nullptr,
                                use_old ? MC.off : nullptr,
                                use_old ? (const double*)C->d_data.p : nullptr,
                                (double*)newdata.p, (uint16_t*)ctx->d_kept.p, variant, whole,
                                form, slot_counter, &grid);
    return launch_numeric_bf16(s, slices, n_tiles, (const TileDesc*)S.r.tiles.p,
                               (const int32_t*)S.r.perm.p, (const Entry*)S.r.entries.p, D.bands,
                               ctx->d_aslice.p, ctx->d_bslice.p, a_zero, b_zero, alpha, beta,
                               post2, use_old ? MC.row_p : nullptr, use_old ? MC.col : nullptr,
                               use_old ? MC.off : nullptr,
                               use_old ? (const double*)C->d_data.p : nullptr,
                               (double*)newdata.p, (uint16_t*)ctx->d_kept.p, form,
                               slot_counter, &grid);
  };
  // time one launch of the numeric kernel (every candidate writes the whole
  // C pool and kept masks; the last launch's result is the call's)
  auto timed = [&](int form, float* ms) -> hipError_t {
    hipError_t e;
    if ((e = record_event(ctx, 0)) != hipSuccess) return e;
    if ((e = launch_numeric(form)) != hipSuccess) return e;
    if ((e = record_event(ctx, 1)) != hipSuccess) return e;
    if ((e = hipEventSynchronize(ctx->ev[1])) != hipSuccess) return e;
    return hipEventElapsedTime(ms, ctx->ev[0], ctx->ev[1]);
  };
  if (tune_shape) {
    // shape tuning: each variant once with the default tile order and one
    // workgroup per tile; the fastest becomes the class's shape
    float best = 0.f;
    int best_v = variant;
    for (int v : {kF64Half, kF64Quadrant, kF64Ksplit, kF64Lanes, kF64Fine}) {
      if ((v == kF64Lanes || v == kF64Fine) && !lanes_ok) continue;
      variant = v;
      float ms = 0.f;
      HIPCHK(ctx, timed(choice.form, &ms), "numeric kernel (shape tuning)");
      if (v == kF64Half || ms < best) {
        best = ms;
        best_v = v;
      }
    }
    variant = best_v;
    ctx->shape_tune[shape_key] = best_v;
  }
  if (!split) ctx->last_variant = variant;
  const std::array<int64_t, 4> key = {
      (int64_t)flags | ((int64_t)whole << 32) | ((int64_t)variant << 40), d->max_blk,
      bucket(S.r.flops / nt), bucket((double)S.r.n_struct / nt)};
  if (!env_rb && !env_pe) {
    if (auto it = ctx->tune.find(key); it != ctx->tune.end()) {
      choice = it->second;
      if (choice.rowblock != 8)
        HIPCHK(ctx, build_tile_order(ctx, S.W, D, S.r, choice.rowblock), "tile order");
    } else if (ctx->size_mode != kSizesReplay && n_tiles > 0) {
      tune = true;
    }
  }
  if (tune) {
    // every candidate writes the same C pool and kept masks; the last
    // launch's result is the call's, its time the call's numeric time
    // Two stages, five launches: the tile orders in the interleaved
    // persistent form, then the fastest order in the other two forms (the
    // orders rank alike in every form). Within timing noise the XCD ranges
    // form wins: it fetches a fifth to a quarter fewer bytes from beyond the
    // L2 than the interleaved forms, which leaves the fabric to concurrent
    // work, so the interleaved forms' times are compared 0.5 % higher.
    float best = 0.f;
    lamb_hip_ctx::TuneChoice best_choice = choice;
    int32_t best_grid = 0;
    bool first = true;
    auto consider = [&](int32_t rowblock, int form) -> hipError_t {
      float ms = 0.f;
      const hipError_t e = timed(form, &ms);
      if (e != hipSuccess) return e;
      if (form != kFormXcdRanges) ms *= 1.005f;
      if (first || ms < best) {
        best = ms;
        best_choice.rowblock = rowblock;
        best_choice.form = form;
        best_grid = grid;
      }
      first = false;
      return hipSuccess;
    };
    constexpr int kOrders = (int)(sizeof(kRowBlocks) / sizeof(kRowBlocks[0]));
    for (int c = 0; c < kOrders; ++c) {
      HIPCHK(ctx, build_tile_order(ctx, S.W, D, S.r, kRowBlocks[c]), "tile order");
      HIPCHK(ctx, consider(kRowBlocks[c], kFormPersistent), "numeric kernel (tuning)");
    }
    const int32_t order = best_choice.rowblock;
    if (order != kRowBlocks[kOrders - 1])
      HIPCHK(ctx, build_tile_order(ctx, S.W, D, S.r, order), "tile order");
    HIPCHK(ctx, consider(order, kFormGrid), "numeric kernel (tuning)");
    HIPCHK(ctx, consider(order, kFormXcdRanges), "numeric kernel (tuning)");
    ctx->tune[key] = best_choice;
    choice = best_choice;
    grid = best_grid;
  } else {
    HIPCHK(ctx, record_event(ctx, 0), "event");
    HIPCHK(ctx, launch_numeric(choice.form),
           split ? "numeric kernel (BF16 split)" : "numeric kernel");
    HIPCHK(ctx, record_event(ctx, 1), "event");
  }
  ctx->last_rowblock = choice.rowblock;
  ctx->last_grid = grid;
  ctx->last_form = choice.form;
  ctx->last_tiles = n_tiles;

  // ---- C index ----
  CIndex ci;
  ci.row_p = C->s_row_p;
  ci.col = C->s_col;
  ci.off = C->s_off;
  ci.rows = C->s_rows;
  C->s_row_p = C->s_col = C->s_off = C->s_rows = DevBuf();
  HIPCHK(ctx, build_c_index(ctx, S.W, D, S.r, (const uint16_t*)ctx->d_kept.p, ci), "C index");
  HIPCHK(ctx, record_event(ctx, 4), "event");
  HIPCHK(ctx, stream_sync(ctx), "multiply sync");
  float ms_num = 0.f, ms_sym = 0.f, ms_idx = 0.f;
  if (ctx->size_mode != kSizesReplay) {
    HIPCHK(ctx, hipEventElapsedTime(&ms_num, ctx->ev[0], ctx->ev[1]), "event time");
    HIPCHK(ctx, hipEventElapsedTime(&ms_sym, ctx->ev[2], ctx->ev[3]), "event time");
    HIPCHK(ctx, hipEventElapsedTime(&ms_idx, ctx->ev[1], ctx->ev[4]), "event time");
  }

  // swap in the new C; the old buffers become the spares (freed unless a
  // graph is being recorded or captured)
  std::swap(C->d_data, newdata);
  C->s_data = newdata;
  std::swap(C->d_row_p, ci.row_p);
  std::swap(C->d_col, ci.col);
  std::swap(C->d_off, ci.off);
  std::swap(C->d_rows, ci.rows);
  C->s_row_p = ci.row_p;
  C->s_col = ci.col;
  C->s_off = ci.off;
  C->s_rows = ci.rows;
  // The spares become the next multiply's C buffers, so that a repeated
  // product allocates nothing (freeing and allocating a 5 GB C pool cost
  // 1.4 ms per call on the 6,144-atom box). They are released when device
  // memory runs low (less than a quarter of it free).
  if (ctx->size_mode == kSizesLive) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || free_b < total_b / 4)
      C->release_spares();
    (void)hipGetLastError();
  }
  C->nblk = ci.nblk;
  C->ndata = S.r.c_size;
  C->nelems = ci.nelems;
  C->host_index_valid = false;
  C->row_p.clear();
  C->col.clear();
  C->off.clear();
  C->rows.clear();
  C->invalidate_derived();

  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->n_triples_kept = S.r.kept;
    stats->n_triples_filtered = S.r.dropped;
    stats->n_c_blocks_struct = S.r.n_struct;
    stats->n_c_blocks_kept = C->nblk;
    stats->n_c_elems_kept = C->nelems;
    stats->n_tiles = n_tiles;
    stats->n_tile_entries = S.r.n_entries;
    stats->flops = S.r.flops;
    stats->ms_numeric = ms_num;
    stats->ms_symbolic = ms_sym;
    stats->ms_index = ms_idx;
    stats->ms_total = std::chrono::duration<double, std::milli>(
                          std::chrono::steady_clock::now() - t_start).count();
  }
  return LAMB_HIP_OK;
}

extern "C" int32_t lamb_hip_tuning_info(const lamb_hip_ctx* ctx, int32_t* last_variant,
                                        int32_t* n_classes) {
  if (!ctx) return LAMB_HIP_FAILURE;
  if (last_variant) *last_variant = ctx->last_variant;
  if (n_classes) *n_classes = (int32_t)ctx->tune.size();
  return LAMB_HIP_OK;
}

// ---------------------------------------------------------------------------
// sector products and the post-filter on its own
// ---------------------------------------------------------------------------
namespace {
// The blocks of X whose row atom lies in subset sr and column atom in subset
// sc, as a temporary matrix with its own index that borrows X's value pool
// (drop_view gives the pool back without freeing it).
int32_t sector_view(lamb_hip_mat* X, int32_t sr, int32_t sc, lamb_hip_mat& V) {
  lamb_hip_dist* d = X->dist;
  lamb_hip_ctx* ctx = d->ctx;
  HIPCHK(ctx, X->download_index(ctx->stream), "sector view: index download");
  V.dist = d;
  V.row_p.assign((size_t)d->n + 1, 0);
  V.col.clear();
  V.off.clear();
  V.rows.clear();
  int64_t ne = 0;
  for (int32_t p = 0; p < d->n; ++p) {
    if (d->subset_of[d->iperm[p]] == sr)
      for (int64_t b = X->row_p[p]; b < X->row_p[p + 1]; ++b) {
        const int32_t q = X->col[b];
        if (d->subset_of[d->iperm[q]] != sc) continue;
        V.col.push_back(q);
        V.off.push_back(X->off[b]);
        V.rows.push_back(p);
        ne += (int64_t)d->blk_int[p] * d->blk_int[q];
      }
    V.row_p[p + 1] = (int64_t)V.col.size();
  }
  V.nblk = (int64_t)V.col.size();
  V.ndata = X->ndata;
  V.nelems = ne;
  V.host_index_valid = true;
  V.nsq_valid = false;
  HIPCHK(ctx, V.upload_index(ctx->stream), "sector view: index upload");
  V.d_data.p = X->d_data.p;
  V.d_data.bytes = X->d_data.bytes;
  return LAMB_HIP_OK;
}

void drop_view(lamb_hip_mat& V) {
  V.d_data.p = nullptr;
  V.d_data.bytes = 0;
  V.release_device();
}
}  // namespace

extern "C" int32_t lamb_hip_sector_multiply(lamb_hip_ctx* ctx, const lamb_hip_mat* A,
                                            const lamb_hip_mat* B, lamb_hip_mat* C,
                                            int32_t sm, int32_t sk, int32_t sn, double eps,
                                            int32_t flags, lamb_hip_stats* stats) {
  if (!ctx || !A || !B || !C) return LAMB_HIP_FAILURE;
  lamb_hip_dist* d = A->dist;
  if (B->dist != d || C->dist != d)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "sector_multiply: A, B, C must share one distribution");
  if (C == A || C == B)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "sector_multiply: C aliases A or B");
  for (int32_t s : {sm, sk, sn})
    if (s < 1 || s > d->n_subsets)
      return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                       "sector_multiply: subset %d outside 1..%d", s, d->n_subsets);
  if (!(eps >= 0.0))
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "sector_multiply: eps_filter < 0");
  if (values_missing(A) || values_missing(B) || values_missing(C))
    return refuse_missing_values(ctx, "sector_multiply");
  (void)hipSetDevice(ctx->device);
  lamb_hip_mat VA, VB;
  int32_t rc = sector_view(const_cast<lamb_hip_mat*>(A), sm, sk, VA);
  if (!rc) rc = sector_view(const_cast<lamb_hip_mat*>(B), sk, sn, VB);
  // C += A(m,k) B(k,n): beta = 1 keeps C's other blocks exactly (alpha * 0 +
  // 1 * old), the post-filter threshold 0 keeps every block
  if (!rc)
    rc = multiply_masked(ctx, &VA, &VB, C, 1.0, C->nblk > 0 ? 1.0 : 0.0, eps, flags, stats,
                         nullptr, 0.0);
  drop_view(VA);
  drop_view(VB);
  return rc;
}

extern "C" int32_t lamb_hip_matrix_filter(lamb_hip_mat* C, double eps, int64_t* n_dropped) {
  if (!C) return LAMB_HIP_FAILURE;
  lamb_hip_ctx* ctx = C->dist->ctx;
  if (!(eps >= 0.0))
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS, "matrix_filter: eps_filter < 0");
  if (values_missing(C)) return refuse_missing_values(ctx, "matrix_filter");
  (void)hipSetDevice(ctx->device);
  int32_t rc;
  if ((rc = ensure_nsq(C))) return rc;
  HIPCHK(ctx, C->download_index(ctx->stream), "matrix_filter: index download");
  std::vector<double> nsq((size_t)C->nblk);
  if (C->nblk)
    HIPCHK(ctx, hipMemcpyAsync(nsq.data(), C->d_nsq.p, sizeof(double) * nsq.size(),
                               hipMemcpyDeviceToHost, ctx->stream), "matrix_filter: nsq");
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "matrix_filter sync");
  const lamb_hip_dist* d = C->dist;
  const double eps2 = eps * eps;
  std::vector<int64_t> row_p((size_t)d->n + 1, 0);
  std::vector<int32_t> col, rows;
  std::vector<int64_t> off;
  int64_t ne = 0, dropped = 0;
  for (int32_t p = 0; p < d->n; ++p) {
    for (int64_t b = C->row_p[p]; b < C->row_p[p + 1]; ++b) {
      if (nsq[b] >= eps2) {
        col.push_back(C->col[b]);
        off.push_back(C->off[b]);
        rows.push_back(p);
        ne += (int64_t)d->blk_int[p] * d->blk_int[C->col[b]];
      } else {
        ++dropped;
      }
    }
    row_p[p + 1] = (int64_t)col.size();
  }
  if (dropped) {
    C->row_p = std::move(row_p);
    C->col = std::move(col);
    C->off = std::move(off);
    C->rows = std::move(rows);
    C->nblk = (int64_t)C->col.size();
    C->nelems = ne;
    C->host_index_valid = true;
    C->invalidate_derived();
    HIPCHK(ctx, C->upload_index(ctx->stream), "matrix_filter: index upload");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "matrix_filter sync");
  }
  if (n_dropped) *n_dropped = dropped;
  return LAMB_HIP_OK;
}

extern "C" int32_t lamb_hip_tuning_order(const lamb_hip_ctx* ctx, int32_t* last_rowblock) {
  if (!ctx) return LAMB_HIP_FAILURE;
  if (last_rowblock) *last_rowblock = ctx->last_rowblock;
  return LAMB_HIP_OK;
}

extern "C" int32_t lamb_hip_tuning_mode(lamb_hip_ctx* ctx, int32_t tune_shapes) {
  if (!ctx) return LAMB_HIP_FAILURE;
  if (tune_shapes != 0 && tune_shapes != 1)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "tuning_mode: tune_shapes must be 0 or 1, got %d", tune_shapes);
  ctx->tune_shapes = tune_shapes;
  return LAMB_HIP_OK;
}

// Tuning table file: a header line, then one line per tuned class.
//   shape k0 k1 k2 k3 k4 variant
//   order k0 k1 k2 k3 rowblock form
static const char kTuneHeader[] = "lamb_hip tuning table 1";

extern "C" int32_t lamb_hip_tuning_save(const lamb_hip_ctx* ctx, const char* path) {
  if (!ctx) return LAMB_HIP_FAILURE;
  lamb_hip_ctx* c = const_cast<lamb_hip_ctx*>(ctx);
  if (!path) return set_error(c, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS, "tuning_save: NULL path");
  FILE* f = fopen(path, "w");
  if (!f)
    return set_error(c, LAMB_HIP_FAILURE, LAMB_HIP_EXTERNAL_ERROR,
                     "tuning_save: cannot open %s for writing", path);
  fprintf(f, "%s\n", kTuneHeader);
  for (const auto& [k, v] : ctx->shape_tune)
    fprintf(f, "shape %lld %lld %lld %lld %lld %d\n", (long long)k[0], (long long)k[1],
            (long long)k[2], (long long)k[3], (long long)k[4], v);
  for (const auto& [k, v] : ctx->tune)
    fprintf(f, "order %lld %lld %lld %lld %d %d\n", (long long)k[0], (long long)k[1],
            (long long)k[2], (long long)k[3], v.rowblock, v.form);
  const bool ok = fclose(f) == 0;
  return ok ? LAMB_HIP_OK
            : set_error(c, LAMB_HIP_FAILURE, LAMB_HIP_EXTERNAL_ERROR,
                        "tuning_save: write to %s failed", path);
}

extern "C" int32_t lamb_hip_tuning_load(lamb_hip_ctx* ctx, const char* path) {
  if (!ctx) return LAMB_HIP_FAILURE;
  if (!path) return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS, "tuning_load: NULL path");
  FILE* f = fopen(path, "r");
  if (!f)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_EXTERNAL_ERROR,
                     "tuning_load: cannot open %s", path);
  std::map<std::array<int64_t, 5>, int32_t> shapes;
  std::map<std::array<int64_t, 4>, lamb_hip_ctx::TuneChoice> orders;
  char line[512];
  int lineno = 0;
  int32_t rc = LAMB_HIP_OK;
  while (fgets(line, sizeof line, f)) {
    ++lineno;
    if (lineno == 1) {
      if (strncmp(line, kTuneHeader, sizeof kTuneHeader - 1) != 0) {
        rc = set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                       "tuning_load: %s is not a tuning table", path);
        break;
      }
      continue;
    }
    long long k[5];
    int a = 0, b = 0;
    char tail = 0;
    if (sscanf(line, "shape %lld %lld %lld %lld %lld %d %c", &k[0], &k[1], &k[2], &k[3], &k[4],
               &a, &tail) == 6 &&
        (a == kF64Quadrant || a == kF64Ksplit || a == kF64Half || a == kF64Lanes ||
         a == kF64Fine)) {
      shapes[{k[0], k[1], k[2], k[3], k[4]}] = a;
    } else if (sscanf(line, "order %lld %lld %lld %lld %d %d %c", &k[0], &k[1], &k[2], &k[3], &a,
                      &b, &tail) == 6 &&
               a >= 1 && a <= 32 && b >= kFormGrid && b <= kFormXcdRanges) {
      lamb_hip_ctx::TuneChoice t;
      t.rowblock = a;
      t.form = b;
      orders[{k[0], k[1], k[2], k[3]}] = t;
    } else {
      rc = set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "tuning_load: %s line %d is not a tuning entry", path, lineno);
      break;
    }
  }
  fclose(f);
  if (rc == LAMB_HIP_OK && lineno == 0)
    rc = set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS, "tuning_load: %s is empty", path);
  if (rc != LAMB_HIP_OK) return rc;
  for (const auto& [k, v] : shapes) ctx->shape_tune[k] = v;
  for (const auto& [k, v] : orders) ctx->tune[k] = v;
  return LAMB_HIP_OK;
}

extern "C" int32_t lamb_hip_tuning_launch(const lamb_hip_ctx* ctx, int32_t* form,
                                          int32_t* workgroups, int32_t* tiles) {
  if (!ctx) return LAMB_HIP_FAILURE;
  if (form) *form = ctx->last_form;
  if (workgroups) *workgroups = ctx->last_grid;
  if (tiles) *tiles = ctx->last_tiles;
  return LAMB_HIP_OK;
}
