// Size reads of the host from the device, and McWeeny chains captured as
// HIP graphs.
//
// The multiply reads a few sizes back to the host (strip and panel counts,
// tile / entry / C totals, C block counts) to size its buffers and grids.
// A graph cannot contain such reads, so a chain is captured in two passes:
// the chain runs once eagerly while every size read is logged (record), then
// it runs again under stream capture where every size read returns the
// logged value and enqueues a check kernel that raises a guard flag when the
// device computes a different value (replay). Buffers whose role rotates
// through the chain are grown to the chain's maximum before the capture, so
// the captured work allocates and frees nothing. A replay whose guard flag
// is raised is reported as an error (its results are not those of the
// chain); otherwise the graph computed exactly what the eager chain does.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "device_util.h"
#include "internal.h"

using namespace lamb;

namespace lamb {

// the flag itself is never guarded: it must be clearable while raised
__global__ void guard_reset_kernel(unsigned long long* flag) {
  if (threadIdx.x == 0) *flag = 0ull;
}

hipError_t install_guard(int device, unsigned long long** flag) {
  static unsigned long long* flags[256] = {};
  if (device < 0 || device >= 256) return hipErrorInvalidDevice;
  if (!flags[device]) {
    unsigned long long* f = nullptr;
    hipError_t e = hipMalloc(&f, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(f, 0, sizeof(unsigned long long));
    for (auto set : {lamb_guard_set_symbolic, lamb_guard_set_scan, lamb_guard_set_kernels,
                     lamb_guard_set_bf16split, lamb_guard_set_chain, lamb_guard_set_adjoint})
      if (e == hipSuccess) e = set(f);
    if (e != hipSuccess) return e;
    flags[device] = f;
  }
  *flag = flags[device];
  return hipSuccess;
}

__global__ void size_guard_kernel(const int64_t* __restrict__ dev, int n, int64_t e0, int64_t e1,
                                  int64_t e2, int64_t e3, int64_t e4, int64_t e5, int64_t e6,
                                  int64_t e7, unsigned long long* __restrict__ flag) {
  const int t = threadIdx.x;
  if (t >= n) return;
  const int64_t e[8] = {e0, e1, e2, e3, e4, e5, e6, e7};
  if (dev[t] != e[t]) atomicOr(flag, 1ull);
}

hipError_t read_sizes(lamb_hip_ctx* ctx, const void* src, int64_t* dst, int n) {
  if (ctx->size_mode == kSizesReplay) {
    if (n > 8 || ctx->size_cursor + (size_t)n > ctx->size_log.size()) return hipErrorInvalidValue;
    int64_t e[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int t = 0; t < n; ++t) e[t] = dst[t] = ctx->size_log[ctx->size_cursor++];
    hipLaunchKernelGGL(size_guard_kernel, dim3(1), dim3(64), 0, ctx->stream,
                       (const int64_t*)src, n, e[0], e[1], e[2], e[3], e[4], e[5], e[6], e[7],
                       ctx->guard);
    return hipGetLastError();
  }
  hipError_t err = hipMemcpyAsync(dst, src, sizeof(int64_t) * (size_t)n, hipMemcpyDeviceToHost,
                                  ctx->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(ctx->stream);
  if (err == hipSuccess && ctx->size_mode == kSizesRecord)
    ctx->size_log.insert(ctx->size_log.end(), dst, dst + n);
  return err;
}

hipError_t stream_sync(lamb_hip_ctx* ctx) {
  if (ctx->size_mode == kSizesReplay) return hipSuccess;
  return hipStreamSynchronize(ctx->stream);
}

}  // namespace lamb

// ---------------------------------------------------------------------------
// McWeeny chains as graphs
// ---------------------------------------------------------------------------
namespace lamb {

// raise flag where two patterns (row_p[n+1], col[nblk], off[nblk]) differ
__global__ void pattern_diff_kernel(int64_t n1, const int64_t* __restrict__ ra,
                                    const int64_t* __restrict__ rb, int64_t nb,
                                    const int32_t* __restrict__ ca, const int32_t* __restrict__ cb,
                                    const int64_t* __restrict__ oa, const int64_t* __restrict__ ob,
                                    unsigned long long* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool diff = false;
  if (i < n1) diff |= ra[i] != rb[i];
  if (i < nb) diff |= ca[i] != cb[i] || oa[i] != ob[i];
  if (diff) atomicOr(flag, 1ull);
}

}  // namespace lamb

struct lamb_hip_graph {
  lamb_hip_ctx* ctx = nullptr;
  lamb_hip_dist* dist = nullptr;
  int32_t steps = 0, flags = 0;
  double eps = 0.0;
  lamb_hip_mat in, W, P2, T;  // captured input, iterate, square, update
  std::vector<int64_t> log;
  DevBuf tr_part, idem_part;  // [steps][n_atoms] per-atom partials
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  lamb_hip_stats rec;  // statistics of the recorded chain
  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    for (lamb_hip_mat* m : {&in, &W, &P2, &T}) m->release_device();
    tr_part.release();
    idem_part.release();
  }
};

namespace {

#define GCHK(ctx, expr, what)                              \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return hip_error(ctx, e_, what); \
  } while (0)

struct ChainMax {
  int64_t nblk = 0, ndata = 0;
  void see(const lamb_hip_mat& m) {
    nblk = std::max(nblk, m.nblk);
    ndata = std::max(ndata, m.ndata);
  }
};

// The chain on the graph's matrices, in the context's current size mode:
// W <- in, then per step P2 = W W, idem partials of (P2, W),
// T = copy(P2) = -2 P2 W + 3 T, swap(W, T), trace partials of W.
int32_t run_chain(lamb_hip_graph* g, ChainMax* mx, lamb_hip_stats* acc) {
  lamb_hip_ctx* ctx = g->ctx;
  const int32_t n = g->dist->n;
  int32_t rc = lamb_hip_matrix_copy(&g->in, &g->W);
  if (rc) return rc;
  if (mx) mx->see(g->W);
  for (int32_t k = 0; k < g->steps; ++k) {
    lamb_hip_stats s;
    rc = lamb_hip_multiply(ctx, &g->W, &g->W, &g->P2, 1.0, 0.0, g->eps, g->flags, &s);
    if (rc) return rc;
    if (acc) acc->flops += s.flops;
    GCHK(ctx, launch_diff_partials(ctx->stream, &g->P2, &g->W,
                                   (double*)g->idem_part.p + (int64_t)k * n), "idem partials");
    rc = lamb_hip_matrix_copy(&g->P2, &g->T);
    if (rc) return rc;
    rc = lamb_hip_multiply(ctx, &g->P2, &g->W, &g->T, -2.0, 3.0, g->eps, g->flags, &s);
    if (rc) return rc;
    if (acc) acc->flops += s.flops;
    std::swap(g->W, g->T);
    g->W.invalidate_derived();
    GCHK(ctx, launch_trace_partials(ctx->stream, &g->W, (double*)g->tr_part.p + (int64_t)k * n),
         "trace partials");
    if (mx) {
      mx->see(g->W);
      mx->see(g->P2);
      mx->see(g->T);
    }
  }
  return LAMB_HIP_OK;
}

// grow every rotating buffer of a matrix to the chain's maximum
hipError_t grow(lamb_hip_mat& m, const ChainMax& mx, int32_t n) {
  const size_t nb = (size_t)std::max<int64_t>(mx.nblk, 1), nd = (size_t)std::max<int64_t>(mx.ndata, 1);
  hipError_t e;
  for (DevBuf* b : {&m.d_data, &m.s_data})
    if ((e = b->ensure(sizeof(double) * nd)) != hipSuccess) return e;
  for (DevBuf* b : {&m.d_row_p, &m.s_row_p})
    if ((e = b->ensure(sizeof(int64_t) * (size_t)(n + 1))) != hipSuccess) return e;
  for (DevBuf* b : {&m.d_col, &m.s_col, &m.d_rows, &m.s_rows})
    if ((e = b->ensure(sizeof(int32_t) * nb)) != hipSuccess) return e;
  if ((e = m.d_off.ensure(sizeof(int64_t) * nb)) != hipSuccess) return e;
  if ((e = m.s_off.ensure(sizeof(int64_t) * nb)) != hipSuccess) return e;
  return m.d_nsq.ensure(sizeof(double) * nb);
}

}  // namespace

extern "C" int32_t lamb_hip_mcweeny_capture(lamb_hip_ctx* ctx, const lamb_hip_mat* P,
                                            int32_t steps, double eps_filter, int32_t flags,
                                            lamb_hip_graph** out) {
  if (!ctx || !P || !out || steps < 1) return LAMB_HIP_FAILURE;
  if (values_missing(P)) return refuse_missing_values(ctx, "mcweeny_capture");
  if (ctx->size_mode != kSizesLive)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_PRECONDITION_FAILED,
                     "mcweeny_capture: another capture is in progress");
  (void)hipSetDevice(ctx->device);
  auto g = new lamb_hip_graph();
  g->ctx = ctx;
  g->dist = P->dist;
  g->steps = steps;
  g->flags = flags;
  g->eps = eps_filter;
  for (lamb_hip_mat* m : {&g->in, &g->W, &g->P2, &g->T}) m->dist = P->dist;
  const int32_t n = P->dist->n;
  auto fail = [&](int32_t rc) {
    ctx->size_mode = kSizesLive;
    g->release();
    delete g;
    return rc;
  };
  int32_t rc = lamb_hip_matrix_copy(P, &g->in);
  if (rc) return fail(rc);
  if (g->tr_part.ensure(sizeof(double) * (size_t)std::max<int64_t>((int64_t)steps * n, 1)) ||
      g->idem_part.ensure(sizeof(double) * (size_t)std::max<int64_t>((int64_t)steps * n, 1)))
    return fail(set_error(ctx, LAMB_HIP_MEMORY, LAMB_HIP_EXTERNAL_ERROR, "capture: buffers"));
  // 1. record the chain's sizes
  ctx->size_mode = kSizesRecord;
  ctx->size_log.clear();
  ChainMax mx;
  memset(&g->rec, 0, sizeof(g->rec));
  rc = run_chain(g, &mx, &g->rec);
  ctx->size_mode = kSizesLive;
  if (rc) return fail(rc);
  g->log = ctx->size_log;
  // 2. every buffer whose role rotates gets the chain's maximum
  for (lamb_hip_mat* m : {&g->W, &g->P2, &g->T})
    if (grow(*m, mx, n) != hipSuccess)
      return fail(set_error(ctx, LAMB_HIP_MEMORY, LAMB_HIP_EXTERNAL_ERROR, "capture: grow"));
  if (hipStreamSynchronize(ctx->stream) != hipSuccess)
    return fail(set_error(ctx, LAMB_HIP_FATAL, LAMB_HIP_EXTERNAL_ERROR, "capture: sync"));
  // 3. capture the same chain with the recorded sizes
  ctx->size_mode = kSizesReplay;
  ctx->size_cursor = 0;
  if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return fail(set_error(ctx, LAMB_HIP_FATAL, LAMB_HIP_EXTERNAL_ERROR, "capture: begin"));
  hipLaunchKernelGGL(guard_reset_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->guard);
  rc = run_chain(g, nullptr, nullptr);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
  ctx->size_mode = kSizesLive;
  g->graph = graph;
  if (rc) return fail(rc);
  if (ec != hipSuccess || !graph)
    return fail(set_error(ctx, LAMB_HIP_FATAL, LAMB_HIP_EXTERNAL_ERROR, "capture: %s",
                          hipGetErrorString(ec)));
  if (ctx->size_cursor != g->log.size())
    return fail(set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_INTERNAL_ERROR,
                          "capture: %zu of %zu recorded sizes used", ctx->size_cursor,
                          g->log.size()));
  if (hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0) != hipSuccess)
    return fail(set_error(ctx, LAMB_HIP_FATAL, LAMB_HIP_EXTERNAL_ERROR, "capture: instantiate"));
  *out = g;
  return LAMB_HIP_OK;
}

extern "C" int32_t lamb_hip_graph_replay(lamb_hip_graph* g, const lamb_hip_mat* P_in,
                                         lamb_hip_mat* P_out, double* traces, double* idem,
                                         double* ms) {
  if (!g || !P_in || !P_out) return LAMB_HIP_FAILURE;
  lamb_hip_ctx* ctx = g->ctx;
  if (P_in->dist != g->dist || P_out->dist != g->dist)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "graph_replay: matrices of another distribution");
  if (values_missing(P_in)) return refuse_missing_values(ctx, "graph_replay");
  if (P_in->nblk != g->in.nblk || P_in->ndata != g->in.ndata)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_PRECONDITION_FAILED,
                     "graph_replay: input pattern differs from the captured one");
  (void)hipSetDevice(ctx->device);
  hipStream_t s = ctx->stream;
  const int32_t n = g->dist->n;
  // the input must have the captured pattern: compare, then copy its values
  hipLaunchKernelGGL(guard_reset_kernel, dim3(1), dim3(64), 0, s, ctx->guard);
  const int64_t cover = std::max<int64_t>((int64_t)n + 1, P_in->nblk);
  hipLaunchKernelGGL(pattern_diff_kernel, dim3((unsigned)((cover + 255) / 256)), dim3(256), 0, s,
                     (int64_t)n + 1, (const int64_t*)P_in->d_row_p.p,
                     (const int64_t*)g->in.d_row_p.p, P_in->nblk, (const int32_t*)P_in->d_col.p,
                     (const int32_t*)g->in.d_col.p, (const int64_t*)P_in->d_off.p,
                     (const int64_t*)g->in.d_off.p, ctx->guard);
  int64_t differs = 0;
  GCHK(ctx, hipMemcpyAsync(&differs, ctx->guard, sizeof(int64_t), hipMemcpyDeviceToHost, s),
       "replay: pattern check");
  GCHK(ctx, hipStreamSynchronize(s), "replay: sync");
  if (differs) {
    hipLaunchKernelGGL(guard_reset_kernel, dim3(1), dim3(64), 0, s, ctx->guard);
    GCHK(ctx, hipStreamSynchronize(s), "replay: guard reset");
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_PRECONDITION_FAILED,
                     "graph_replay: input pattern differs from the captured one");
  }
  if (P_in != &g->in && P_in->ndata)
    GCHK(ctx, hipMemcpyAsync(g->in.d_data.p, P_in->d_data.p, sizeof(double) * P_in->ndata,
                             hipMemcpyDeviceToDevice, s), "replay: input values");
  GCHK(ctx, hipEventRecord(ctx->ev[5], s), "event");
  GCHK(ctx, hipGraphLaunch(g->exec, s), "replay: launch");
  GCHK(ctx, hipEventRecord(ctx->ev[6], s), "event");
  GCHK(ctx, hipStreamSynchronize(s), "replay: sync");
  float t = 0.f;
  GCHK(ctx, hipEventElapsedTime(&t, ctx->ev[5], ctx->ev[6]), "event time");
  if (ms) *ms = t;
  int64_t guard = 0;
  GCHK(ctx, hipMemcpy(&guard, ctx->guard, sizeof(int64_t), hipMemcpyDeviceToHost),
       "replay: guard");
  if (guard) {
    // every kernel after the first differing size returned at entry; clear
    // the flag so that later work runs
    hipLaunchKernelGGL(guard_reset_kernel, dim3(1), dim3(64), 0, s, ctx->guard);
    GCHK(ctx, hipStreamSynchronize(s), "replay: guard reset");
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_PRECONDITION_FAILED,
                     "graph_replay: the chain's sizes differ from the recorded chain "
                     "(different filter decisions); run it eagerly");
  }
  const int32_t G = g->dist->n_groups;
  std::vector<double> part((size_t)g->steps * n);
  for (int pass = 0; pass < 2; ++pass) {
    double* dst = pass == 0 ? traces : idem;
    if (!dst || part.empty()) continue;
    GCHK(ctx, hipMemcpy(part.data(), pass == 0 ? g->tr_part.p : g->idem_part.p,
                        sizeof(double) * part.size(), hipMemcpyDeviceToHost), "partials");
    for (int32_t k = 0; k < g->steps; ++k) {
      double* row = dst + (int64_t)k * G;
      std::fill(row, row + G, 0.0);
      for (int32_t p = 0; p < n; ++p) row[g->dist->group_int[p]] += part[(size_t)k * n + p];
      if (pass == 1)
        for (int32_t q = 0; q < G; ++q) row[q] = std::sqrt(row[q]);
    }
  }
  return lamb_hip_matrix_copy(&g->W, P_out);
}

extern "C" int32_t lamb_hip_graph_info(const lamb_hip_graph* g, int32_t* steps, int64_t* sizes,
                                       double* flops) {
  if (!g) return LAMB_HIP_FAILURE;
  if (steps) *steps = g->steps;
  if (sizes) *sizes = (int64_t)g->log.size();
  if (flops) *flops = g->rec.flops;
  return LAMB_HIP_OK;
}

extern "C" int32_t lamb_hip_graph_destroy(lamb_hip_graph* g) {
  if (!g) return LAMB_HIP_OK;
  (void)hipSetDevice(g->ctx->device);
  g->release();
  delete g;
  return LAMB_HIP_OK;
}
