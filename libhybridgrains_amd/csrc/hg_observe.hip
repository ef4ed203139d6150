// hg_observe.hip -- scene observables computed on the device: the kinetic
// energy, the front (largest x) and top (largest y) of the material, and its
// height profile over equal x-bins.  These are what a long granular collapse
// is judged by (runout, final pile profile, kinetic-energy curve), read from
// the resident state without downloading it.
//
// One pass over the scene's points or bodies: each workgroup reduces its
// share of the kinetic energy in a fixed tree order and writes it, with its
// largest x and y, to a partials array; each point raises its bin's height
// with an integer atomic maximum on an order-preserving encoding of the
// double (order-independent); one workgroup then sums the partials in index
// order.  The launch shape is fixed, so results are bitwise reproducible.
#include "hg_dem_handle.h"
#include "hg_mpm_handle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#define HG_TRY_OBS(expr)          \
  do {                            \
    hg_status s_ = (expr);        \
    if (s_ != HG_OK) return s_;   \
  } while (0)

namespace {

constexpr int kObsThreads = 256;
constexpr uint32_t kObsBlocks = 1024;

// scratch words: [kObsBlocks] kinetic-energy partials, [kObsBlocks] largest
// x, [kObsBlocks] largest y, [4] results, [nbins] bin heights
size_t obs_words(uint32_t nbins) { return 3 * size_t(kObsBlocks) + 4 + size_t(nbins); }

// Monotonic map of doubles onto unsigned integers (larger double, larger
// code); 0 is below every finite or infinite value and marks an empty bin.
__host__ __device__ inline uint64_t obs_code(double v) {
  uint64_t u;
  __builtin_memcpy(&u, &v, 8);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__host__ __device__ inline double obs_value(uint64_t c) {
  const uint64_t u = (c >> 63) ? (c & 0x7fffffffffffffffull) : ~c;
  double v;
  __builtin_memcpy(&v, &u, 8);
  return v;
}

struct ObsArgs {
  const double *x, *y;
  const void *vx, *vy, *m;
  const uint32_t *id;  // where not null, m is indexed by id[i]
  double m_uniform;    // every item's mass where m is null
  uint32_t n;
  double x0, inv_w;
  uint32_t nbins;
  uint64_t *s;  // scratch
};

__global__ __launch_bounds__(kObsThreads) void obs_clear_kernel(uint64_t *s, uint32_t nbins) {
  const uint32_t i = blockIdx.x * kObsThreads + threadIdx.x;
  if (i < nbins) s[3 * kObsBlocks + 4 + i] = 0ull;
}

template <typename T>
__global__ __launch_bounds__(kObsThreads) void obs_points_kernel(ObsArgs a) {
  __shared__ double s_ke[kObsThreads], s_mx[kObsThreads], s_my[kObsThreads];
  double ke = 0.0, mx = -INFINITY, my = -INFINITY;
  uint64_t *bins = a.s + 3 * kObsBlocks + 4;
  for (uint32_t i = blockIdx.x * kObsThreads + threadIdx.x; i < a.n;
       i += gridDim.x * kObsThreads) {
    const double x = a.x[i], y = a.y[i];
    const double vx = double(static_cast<const T *>(a.vx)[i]);
    const double vy = double(static_cast<const T *>(a.vy)[i]);
    const double m =
        a.m ? double(static_cast<const T *>(a.m)[a.id ? a.id[i] : i]) : a.m_uniform;
    ke += 0.5 * m * (vx * vx + vy * vy);
    mx = fmax(mx, x);
    my = fmax(my, y);
    const double fb = floor((x - a.x0) * a.inv_w);
    if (fb >= 0.0 && fb < double(a.nbins)) atomicMax(&bins[uint32_t(fb)], obs_code(y));
  }
  const int t = threadIdx.x;
  s_ke[t] = ke;
  s_mx[t] = mx;
  s_my[t] = my;
  __syncthreads();
  for (int h = kObsThreads / 2; h > 0; h >>= 1) {
    if (t < h) {
      s_ke[t] += s_ke[t + h];
      s_mx[t] = fmax(s_mx[t], s_mx[t + h]);
      s_my[t] = fmax(s_my[t], s_my[t + h]);
    }
    __syncthreads();
  }
  if (t == 0) {
    double *p = reinterpret_cast<double *>(a.s);
    p[blockIdx.x] = s_ke[0];
    p[kObsBlocks + blockIdx.x] = s_mx[0];
    p[2 * kObsBlocks + blockIdx.x] = s_my[0];
  }
}

// Sums the partials of the nb workgroups in index order (each thread a
// strided share, then a fixed tree).
__global__ __launch_bounds__(kObsThreads) void obs_final_kernel(uint64_t *s, uint32_t nb) {
  __shared__ double s_ke[kObsThreads], s_mx[kObsThreads], s_my[kObsThreads];
  const double *p = reinterpret_cast<const double *>(s);
  const int t = threadIdx.x;
  double ke = 0.0, mx = -INFINITY, my = -INFINITY;
  for (uint32_t b = uint32_t(t); b < nb; b += kObsThreads) {
    ke += p[b];
    mx = fmax(mx, p[kObsBlocks + b]);
    my = fmax(my, p[2 * kObsBlocks + b]);
  }
  s_ke[t] = ke;
  s_mx[t] = mx;
  s_my[t] = my;
  __syncthreads();
  for (int h = kObsThreads / 2; h > 0; h >>= 1) {
    if (t < h) {
      s_ke[t] += s_ke[t + h];
      s_mx[t] = fmax(s_mx[t], s_mx[t + h]);
      s_my[t] = fmax(s_my[t], s_my[t + h]);
    }
    __syncthreads();
  }
  if (t == 0) {
    double *r = reinterpret_cast<double *>(s + 3 * kObsBlocks);
    r[0] = s_ke[0];
    r[1] = s_mx[0];
    r[2] = s_my[0];
    r[3] = 0.0;
  }
}

// Queues the three kernels on stream st over the arrays in a (a.s is the
// scratch), copies the results into out (3 + nbins doubles) and waits.
template <typename T>
hg_status obs_run(hg_ctx *ctx, hipStream_t st, ObsArgs a, double *out) {
  const uint32_t nb = std::max(1u, std::min(kObsBlocks, (a.n + kObsThreads - 1) / kObsThreads));
  if (a.nbins)
    hipLaunchKernelGGL(obs_clear_kernel, dim3((a.nbins + kObsThreads - 1) / kObsThreads),
                       dim3(kObsThreads), 0, st, a.s, a.nbins);
  hipLaunchKernelGGL(obs_points_kernel<T>, dim3(nb), dim3(kObsThreads), 0, st, a);
  hipLaunchKernelGGL(obs_final_kernel, dim3(1), dim3(kObsThreads), 0, st, a.s, nb);
  HG_HIP(ctx, hipGetLastError());
  // the results and the bin codes are 8 bytes each: copied straight into out
  HG_HIP(ctx, hipMemcpyAsync(out, a.s + 3 * kObsBlocks, 3 * 8, hipMemcpyDeviceToHost, st));
  if (a.nbins)
    HG_HIP(ctx, hipMemcpyAsync(out + 3, a.s + 3 * kObsBlocks + 4, size_t(a.nbins) * 8,
                               hipMemcpyDeviceToHost, st));
  HG_HIP(ctx, hipStreamSynchronize(st));
  for (uint32_t k = 0; k < a.nbins; ++k) {
    uint64_t c;
    std::memcpy(&c, &out[3 + k], 8);
    out[3 + k] = c ? obs_value(c) : NAN;
  }
  return HG_OK;
}

hg_status check_args(hg_ctx *ctx, double bin_width, uint32_t nbins, const double *out) {
  if (!out) return hg_fail(ctx, HG_ERR_INVALID, "null output");
  if (nbins && !(bin_width > 0.0)) return hg_fail(ctx, HG_ERR_INVALID, "bin width must be > 0");
  return HG_OK;
}

// The handle's scratch, grown to hold nbins bins (device memory the handle
// owns until it is destroyed).
hg_status grow(hg_ctx *ctx, std::vector<void *> &allocs, uint64_t &bytes, uint64_t *&scr,
               size_t &have, uint32_t nbins) {
  const size_t need = obs_words(nbins);
  if (scr && have >= need) return HG_OK;
  void *p = nullptr;
  const hipError_t e = hipMalloc(&p, need * 8);
  if (e != hipSuccess) return hg_fail(ctx, HG_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  allocs.push_back(p);
  bytes += need * 8;
  scr = static_cast<uint64_t *>(p);
  have = need;
  return HG_OK;
}

ObsArgs mpm_args(hg_mpm *h, uint32_t scene, double x0, double bin_width, uint32_t nbins) {
  const hgmpm::SceneDev &sc = h->scenes[scene];
  const hgmpm::PtsRaw &p = h->buf[h->cur];
  const size_t T = h->tsize();
  ObsArgs a;
  a.x = p.x + sc.p_off;
  a.y = p.y + sc.p_off;
  a.vx = static_cast<const char *>(p.vx) + T * sc.p_off;
  a.vy = static_cast<const char *>(p.vy) + T * sc.p_off;
  a.m = p.m;  // indexed by point id
  a.id = p.id + sc.p_off;
  a.m_uniform = double(h->precision == HG_FP64 ? sc.m_uniform : double(float(sc.m_uniform)));
  a.n = sc.p_cnt;
  a.x0 = x0;
  a.inv_w = nbins ? 1.0 / bin_width : 0.0;
  a.nbins = nbins;
  a.s = nullptr;
  return a;
}

// ---- allocation-free variant ---------------------------------------------------
// Largest bin count hg_mpm_observables_noalloc accepts.
constexpr uint32_t kObsStaticBins = 8192;
constexpr int kObsMaxDevices = 64;

}  // namespace

// Scratch of hg_mpm_observables_noalloc: 90 KiB on every GPU the library runs
// on, held until the process ends.  In host code this name is a host-side
// stand-in, not a device address: host code uses the per-device address
// hipGetSymbolAddress returns (obs_static), never the name.
static __device__ __align__(256) uint64_t scratch[3 * kObsBlocks + 4 + kObsStaticBins];

static uint64_t *obs_static[kObsMaxDevices];
static std::mutex obs_static_mu;

// Looks up this device's copy of the scratch array once per device (the
// device is current: hg_ctx_create calls this after hipSetDevice).
hg_status hg_observe_init(hg_ctx *ctx) {
  if (ctx->device < 0 || ctx->device >= kObsMaxDevices) return HG_OK;
  std::lock_guard<std::mutex> lk(obs_static_mu);
  if (obs_static[ctx->device]) return HG_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return hg_fail(ctx, HG_ERR_HIP, std::string("observables scratch: hipGetSymbolAddress: ") +
                                        hipGetErrorString(e));
  obs_static[ctx->device] = static_cast<uint64_t *>(p);
  return HG_OK;
}

extern "C" hg_status hg_mpm_observables(hg_mpm *h, uint32_t scene, double x0, double bin_width,
                                        uint32_t nbins, double *out) {
  if (!h || scene >= h->nscenes) return HG_ERR_INVALID;
  HG_TRY_OBS(check_args(h->ctx, bin_width, nbins, out));
  if (h->phase_next != 0) return hg_fail(h->ctx, HG_ERR_INVALID, "a phased step is in progress");
  HG_TRY_OBS(grow(h->ctx, h->allocs, h->bytes, h->obs, h->obs_words, nbins));
  ObsArgs a = mpm_args(h, scene, x0, bin_width, nbins);
  a.s = h->obs;
  return h->precision == HG_FP64 ? obs_run<double>(h->ctx, h->stream, a, out)
                                 : obs_run<float>(h->ctx, h->stream, a, out);
}

extern "C" hg_status hg_dem_observables(hg_dem *h, uint32_t scene, double x0, double bin_width,
                                        uint32_t nbins, double *out) {
  if (!h || scene >= h->nscenes) return HG_ERR_INVALID;
  HG_TRY_OBS(check_args(h->ctx, bin_width, nbins, out));
  HG_TRY_OBS(grow(h->ctx, h->allocs, h->bytes, h->obs, h->obs_words, nbins));
  const hgdem::SceneDev &sc = h->scenes[scene];
  const size_t T = h->tsize();
  ObsArgs a;
  a.x = h->b.x + sc.b_off;
  a.y = h->b.y + sc.b_off;
  a.vx = static_cast<const char *>(h->b.vx) + T * sc.b_off;
  a.vy = static_cast<const char *>(h->b.vy) + T * sc.b_off;
  a.m = static_cast<const char *>(h->b.m) + T * sc.b_off;
  a.id = nullptr;
  a.m_uniform = 0.0;
  a.n = sc.b_cnt;
  a.x0 = x0;
  a.inv_w = nbins ? 1.0 / bin_width : 0.0;
  a.nbins = nbins;
  a.s = h->obs;
  return h->precision == HG_FP64 ? obs_run<double>(h->ctx, h->stream, a, out)
                                 : obs_run<float>(h->ctx, h->stream, a, out);
}

// hg_mpm_observables without any allocation: the same kernels on the
// device's static scratch array.  Calls that overlap on one GPU (handles on
// different streams or threads) share that array and race.
extern "C" hg_status hg_mpm_observables_noalloc(hg_mpm *h, uint32_t scene, double x0,
                                                double bin_width, uint32_t nbins, double *out) {
  if (!h || scene >= h->nscenes) return HG_ERR_INVALID;
  HG_TRY_OBS(check_args(h->ctx, bin_width, nbins, out));
  if (nbins > kObsStaticBins)
    return hg_fail(h->ctx, HG_ERR_INVALID,
                   "at most " + std::to_string(kObsStaticBins) + " bins without allocation (asked " +
                       std::to_string(nbins) + "); hg_mpm_observables takes more");
  if (h->phase_next != 0) return hg_fail(h->ctx, HG_ERR_INVALID, "a phased step is in progress");
  const int dev = h->ctx->device;
  uint64_t *scr = (dev >= 0 && dev < kObsMaxDevices) ? obs_static[dev] : nullptr;
  if (!scr) return hg_fail(h->ctx, HG_ERR_STATE, "observables scratch not set up on this device");
  ObsArgs a = mpm_args(h, scene, x0, bin_width, nbins);
  a.s = scr;
  return h->precision == HG_FP64 ? obs_run<double>(h->ctx, h->stream, a, out)
                                 : obs_run<float>(h->ctx, h->stream, a, out);
}
