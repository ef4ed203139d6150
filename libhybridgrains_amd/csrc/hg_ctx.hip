// hg_ctx.hip -- context, errors and kernel timing for libhg.
#include "hg_internal.h"

#include <cstring>

hg_status hg_fail(hg_ctx *ctx, hg_status code, const std::string &msg) {
  if (ctx) ctx->last_error = msg;
  return code;
}

extern "C" hg_status hg_ctx_create(int32_t device, hg_ctx **out) {
  if (!out) return HG_ERR_INVALID;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return HG_ERR_HIP;
  if (device < 0 || device >= n) return HG_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return HG_ERR_HIP;
  hg_ctx *c = new hg_ctx();
  c->device = device;
  const hg_status st = hg_observe_init(c);
  if (st != HG_OK) {
    delete c;
    return st;
  }
  *out = c;
  return HG_OK;
}

extern "C" hg_status hg_ctx_destroy(hg_ctx *ctx) {
  delete ctx;
  return HG_OK;
}

extern "C" const char *hg_last_error(const hg_ctx *ctx) {
  return ctx ? ctx->last_error.c_str() : "null context";
}

extern "C" const char *hg_version(void) { return "hg 0.1 gfx950"; }

// ---- KernelProfiler -----------------------------------------------------------

int KernelProfiler::kernel_id(const char *name) {
  for (size_t i = 0; i < names.size(); ++i)
    if (names[i] == name) return int(i);
  names.emplace_back(name);
  total_ms.push_back(0.0);
  launches.push_back(0);
  return int(names.size() - 1);
}

hipEvent_t KernelProfiler::take() {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int KernelProfiler::begin(const char *name, hipStream_t s) {
  if (!on) return -1;
  Rec r;
  r.kernel = kernel_id(name);
  r.a = take();
  r.b = take();
  if (!r.a || !r.b) return -1;
  hipEventRecord(r.a, s);
  pending.push_back(r);
  return int(pending.size() - 1);
}

void KernelProfiler::end(int rec, hipStream_t s) {
  if (rec < 0) return;
  hipEventRecord(pending[size_t(rec)].b, s);
}

hipError_t KernelProfiler::collect() {
  for (Rec &r : pending) {
    hipError_t e = hipEventSynchronize(r.b);
    if (e != hipSuccess) return e;
    float ms = 0.0f;
    e = hipEventElapsedTime(&ms, r.a, r.b);
    if (e != hipSuccess) return e;
    total_ms[size_t(r.kernel)] += double(ms);
    launches[size_t(r.kernel)] += 1;
    pool.push_back(r.a);
    pool.push_back(r.b);
  }
  pending.clear();
  return hipSuccess;
}

void KernelProfiler::reset() {
  for (Rec &r : pending) {
    pool.push_back(r.a);
    pool.push_back(r.b);
  }
  pending.clear();
  for (double &t : total_ms) t = 0.0;
  for (int64_t &l : launches) l = 0;
}

KernelProfiler::~KernelProfiler() {
  reset();
  for (hipEvent_t e : pool) hipEventDestroy(e);
}

hg_status hg_profiler_report(KernelProfiler &prof, hg_ctx *ctx, int32_t max_kernels,
                             char *names, double *ms, int64_t *launches,
                             int32_t *nkernels) {
  HG_HIP(ctx, prof.collect());
  const int32_t n = int32_t(prof.names.size());
  if (nkernels) *nkernels = n;
  for (int32_t i = 0; i < n && i < max_kernels; ++i) {
    if (names) {
      std::memset(names + 32 * i, 0, 32);
      std::strncpy(names + 32 * i, prof.names[size_t(i)].c_str(), 31);
    }
    if (ms) ms[i] = prof.total_ms[size_t(i)];
    if (launches) launches[i] = prof.launches[size_t(i)];
  }
  return HG_OK;
}

// ---- stream hold (for profiling) ---------------------------------------------

// One wavefront waits until the constant-rate clock has advanced by ticks;
// the wait ends whatever the clock's start value (unsigned differences).
__global__ __launch_bounds__(64) void hg_hold_kernel(uint64_t ticks) {
  const uint64_t t0 = wall_clock64();
  while (wall_clock64() - t0 < ticks) __builtin_amdgcn_s_sleep(100);
}

hg_status hg_stream_hold(hg_ctx *ctx, hipStream_t s, double us) {
  if (!(us >= 0.0)) return hg_fail(ctx, HG_ERR_INVALID, "hold time must be >= 0");
  us = us > 1.0e6 ? 1.0e6 : us;
  int khz = 0;
  HG_HIP(ctx, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
  if (khz <= 0) return hg_fail(ctx, HG_ERR_HIP, "device clock rate unknown");
  const uint64_t ticks = uint64_t(us * 1e-3 * double(khz));
  hipLaunchKernelGGL(hg_hold_kernel, dim3(1), dim3(64), 0, s, ticks);
  HG_HIP(ctx, hipGetLastError());
  return HG_OK;
}

// ---- streaming copy (bandwidth reference) -------------------------------------

__global__ __launch_bounds__(256) void hg_copy_kernel(const float4 *__restrict__ src,
                                                      float4 *__restrict__ dst, uint64_t n) {
  const uint64_t stride = uint64_t(gridDim.x) * 256u;
  for (uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

extern "C" hg_status hg_bench_copy(hg_ctx *ctx, uint64_t nbytes, int32_t iters, double *gbps) {
  if (!ctx || nbytes < 16 || iters <= 0 || !gbps) return hg_fail(ctx, HG_ERR_INVALID, "bad copy arguments");
  HG_HIP(ctx, hipSetDevice(ctx->device));
  const uint64_t n = nbytes / 16;
  void *a = nullptr, *b = nullptr;
  HG_HIP(ctx, hipMalloc(&a, n * 16));
  if (hipMalloc(&b, n * 16) != hipSuccess) {
    hipFree(a);
    return hg_fail(ctx, HG_ERR_NOMEM, "copy buffer");
  }
  hipMemset(a, 0, n * 16);
  hipStream_t s = nullptr;
  hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  hipDeviceSynchronize();
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  const uint32_t blocks = 256u * 8u;
  hipLaunchKernelGGL(hg_copy_kernel, dim3(blocks), dim3(256), 0, s, static_cast<const float4 *>(a),
                     static_cast<float4 *>(b), n);
  hipEventRecord(e0, s);
  for (int i = 0; i < iters; ++i)
    hipLaunchKernelGGL(hg_copy_kernel, dim3(blocks), dim3(256), 0, s,
                       static_cast<const float4 *>(a), static_cast<float4 *>(b), n);
  hipEventRecord(e1, s);
  hipError_t e = hipEventSynchronize(e1);
  float ms = 0.0f;
  hipEventElapsedTime(&ms, e0, e1);
  *gbps = double(2 * n * 16) * iters / (double(ms) * 1e-3) / 1e9;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipStreamDestroy(s);
  hipFree(a);
  hipFree(b);
  if (e != hipSuccess) return hg_fail(ctx, HG_ERR_HIP, hipGetErrorString(e));
  return HG_OK;
}
