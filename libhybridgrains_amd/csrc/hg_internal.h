// hg_internal.h -- shared host/device internals of libhg (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hg.h"

struct hg_ctx {
  int device = 0;
  std::string last_error;
};

hg_status hg_fail(hg_ctx *ctx, hg_status code, const std::string &msg);

// Per-device set-up of the observables' static scratch (hg_observe.hip);
// hg_ctx_create calls it with the context's device current.
hg_status hg_observe_init(hg_ctx *ctx);

#define HG_HIP(ctx, expr)                                                        \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess)                                                        \
      return hg_fail((ctx), HG_ERR_HIP,                                          \
                     std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)

// Per-kernel device timing with HIP events on one stream.  Off by default;
// when on, each launch is bracketed by two events and the totals are read on
// demand (which synchronises the stream).
struct KernelProfiler {
  bool on = false;
  struct Rec {
    int kernel;
    hipEvent_t a, b;
  };
  std::vector<std::string> names;
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  std::vector<double> total_ms;
  std::vector<int64_t> launches;

  int kernel_id(const char *name);
  hipEvent_t take();
  // Returns the record index or -1 when profiling is off.
  int begin(const char *name, hipStream_t s);
  void end(int rec, hipStream_t s);
  hipError_t collect();
  void reset();
  ~KernelProfiler();
};

struct ProfScope {
  KernelProfiler &p;
  hipStream_t s;
  int rec;
  ProfScope(KernelProfiler &prof, const char *name, hipStream_t st)
      : p(prof), s(st), rec(prof.begin(name, st)) {}
  ~ProfScope() { p.end(rec, s); }
};

// Enqueues a kernel that keeps the stream busy for about us microseconds
// (at most one second), measured on the device's constant-rate clock.
hg_status hg_stream_hold(hg_ctx *ctx, hipStream_t s, double us);

hg_status hg_profiler_report(KernelProfiler &prof, hg_ctx *ctx, int32_t max_kernels,
                             char *names, double *ms, int64_t *launches,
                             int32_t *nkernels);

// ---- device helpers ---------------------------------------------------------

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X; this
// maps workgroup b of a launch padded to a multiple of 8 so that each XCD gets
// one contiguous range of work items (tiles, particle chunks), which keeps
// neighbouring items in the same L2.
__device__ __forceinline__ uint32_t hg_xcd_remap(uint32_t b, uint32_t nblocks_padded) {
  const uint32_t per = nblocks_padded >> 3;
  return (b & 7u) * per + (b >> 3);
}
static inline uint32_t hg_pad8(uint32_t n) { return (n + 7u) & ~7u; }

// Slot of this lane's item among the items of its key: one atomicAdd per
// distinct key of the wavefront (by the key's lowest lane, for the whole
// group), instead of one per lane on the same few counters.  Lanes of a key
// get consecutive slots in lane order.  Call with the wavefront converged;
// inactive lanes take no part.
__device__ __forceinline__ uint32_t hg_wave_slot(uint32_t *counts, uint32_t key) {
  const int lane = threadIdx.x & 63;
  uint64_t pending = __ballot(1);
  int leader = lane;
  uint32_t rank = 0, group = 0;
  while (pending) {
    const int l = __ffsll((unsigned long long)pending) - 1;
    const uint32_t kl = __shfl(key, l, 64);
    const uint64_t same = __ballot(key == kl) & pending;
    if ((same >> lane) & 1ull) {
      leader = l;
      rank = uint32_t(__popcll(same & ((1ull << lane) - 1ull)));
    }
    if (lane == l) group = uint32_t(__popcll(same));
    pending &= ~same;
  }
  uint32_t base = 0;
  if (leader == lane) base = atomicAdd(&counts[key], group);
  base = __shfl(base, leader, 64);
  return base + rank;
}

// As hg_wave_slot, for items whose equal keys mostly sit in runs of
// neighbouring lanes (cell-sorted items): one atomicAdd per run of equal keys
// (by the run's first lane, for the whole run) instead of one pass per
// distinct key.  Lanes of a run get consecutive slots in lane order; a key
// that appears in several runs takes one range per run.  Inactive lanes take
// no part and end runs.
//
// In two halves, so that a caller can leave the atomic's return in flight
// while it works on: hg_run_slot_issue files the run's count and returns the
// pieces, hg_run_slot_take (called later by the same lanes, nothing else
// active) gives the slot.
struct HgRunSlot {
  uint32_t base;  // the run's first slot (valid in the run's first lane)
  uint32_t fr;    // the run's first lane | this lane's rank in the run << 8
};
__device__ __forceinline__ HgRunSlot hg_run_slot_issue(uint32_t *counts, uint32_t key) {
  const int lane = int(threadIdx.x & 63u);
  const uint64_t active = __ballot(1);
  const uint32_t prev = uint32_t(__shfl_up(int(key), 1, 64));
  const bool prev_active = lane > 0 && ((active >> (lane - 1)) & 1ull);
  const bool head = !prev_active || prev != key;
  const uint64_t heads = __ballot(head) & active;
  // the run's first lane: the highest head at or below this lane
  const uint64_t upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  const int first = 63 - __clzll((long long)(heads & upto));
  // the next run's first lane (64 if none)
  const uint64_t after = heads & ~upto;
  const int next = after ? (__ffsll((unsigned long long)after) - 1) : 64;
  const uint64_t below_first = (first == 0) ? 0ull : ((1ull << first) - 1ull);
  const uint64_t run = (next == 64 ? ~0ull : ((1ull << next) - 1ull)) & ~below_first;
  HgRunSlot r;
  r.base = 0;
  if (lane == first) r.base = atomicAdd(&counts[key], uint32_t(__popcll(run & active)));
  const uint64_t before = ((lane == 0) ? 0ull : ((1ull << lane) - 1ull)) & ~below_first;
  r.fr = uint32_t(first) | (uint32_t(__popcll(before & active)) << 8);
  return r;
}
__device__ __forceinline__ uint32_t hg_run_slot_take(const HgRunSlot &r) {
  return uint32_t(__shfl(int(r.base), int(r.fr & 0xffu), 64)) + (r.fr >> 8);
}
__device__ __forceinline__ uint32_t hg_run_slot(uint32_t *counts, uint32_t key) {
  return hg_run_slot_take(hg_run_slot_issue(counts, key));
}

// Wavefront (64-lane) inclusive prefix sum of 32-bit values.
__device__ __forceinline__ uint32_t hg_wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}
