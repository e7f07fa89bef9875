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
                       (unsigned long long*)ctx->d_guard.p);
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
