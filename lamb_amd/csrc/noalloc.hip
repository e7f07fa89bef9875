// Entry points that allocate nothing, for callers that cannot allocate
// (code under HIP graph capture, serving loops, harnesses that forbid it).
//
// Their scratch memory is one file-scope device array per GPU, `scratch`,
// which each GPU the library runs on holds until the process ends. Host code
// never names the array itself (in host code its name is only a stand-in for
// the device symbol): lamb_hip_init calls noalloc_setup once per device with
// that device current, which takes the array's device address with
// hipGetSymbolAddress and keeps it per device; kernels get that address as an
// argument. Every call on a GPU shares that GPU's array and leaves its last
// contents there, so calls that overlap on one GPU (several contexts or host
// threads at once) race.
#include <mutex>

#include "common.h"
#include "internal.h"

namespace lamb {

// partial sums, one double per atom: 8 MiB per GPU
constexpr int64_t kNoAllocMaxAtoms = LAMB_HIP_NOALLOC_MAX_ATOMS;
__device__ __align__(256) double scratch[kNoAllocMaxAtoms];

namespace {
constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

// tr = sum over global atoms g in ascending order of part[perm[g]] (perm:
// global -> internal atom), from +0: the order lamb_hip_trace sums in on the
// host, so both entry points give the same bits
__global__ void ordered_sum_kernel(int32_t n, const int32_t* __restrict__ perm,
                                   const double* __restrict__ part, double* __restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double t = 0.0;
  for (int32_t g = 0; g < n; ++g) t += part[perm[g]];
  *out = t;
}
}  // namespace

hipError_t noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch[device] = static_cast<double*>(p);
  return hipSuccess;
}

}  // namespace lamb

using namespace lamb;

extern "C" int32_t lamb_hip_trace_noalloc(const lamb_hip_mat* m, double* dev_trace) {
  if (!m) return LAMB_HIP_FAILURE;
  lamb_hip_dist* d = m->dist;
  lamb_hip_ctx* ctx = d->ctx;
  if (!dev_trace)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS, "trace_noalloc: NULL output");
  if (values_missing(m)) return refuse_missing_values(ctx, "trace_noalloc");
  if ((int64_t)d->n > kNoAllocMaxAtoms)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_WRONG_ARGS,
                     "trace_noalloc: %d atoms exceed the scratch array of %lld atoms", d->n,
                     (long long)kNoAllocMaxAtoms);
  double* part = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch[ctx->device]
                                                                  : nullptr;
  if (!part)
    return set_error(ctx, LAMB_HIP_FAILURE, LAMB_HIP_INTERNAL_ERROR,
                     "trace_noalloc: no scratch address for device %d", ctx->device);
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_error(ctx, e, "trace_noalloc device");
  if (d->n > 0) {
    // every partial part[0 .. n) is written before it is read
    e = launch_trace(ctx->stream, d->n, (const int64_t*)m->d_row_p.p, (const int32_t*)m->d_col.p,
                     (const int64_t*)m->d_off.p, (const int32_t*)d->d_blk.p,
                     (const double*)m->d_data.p, part);
    if (e != hipSuccess) return hip_error(ctx, e, "trace_noalloc partials");
  }
  hipLaunchKernelGGL(ordered_sum_kernel, dim3(1), dim3(64), 0, ctx->stream, d->n,
                     (const int32_t*)d->d_perm.p, (const double*)part, dev_trace);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_error(ctx, e, "trace_noalloc sum");
  return LAMB_HIP_OK;
}
