// hg_sort.hip -- counting sort of items by cell key: a single-pass exclusive
// scan of the per-cell counts (block offsets from the producer's block
// totals) and a scatter of item indices.
#include "hg_sort.h"

namespace {

// ---------------------------------------------------------------------------
// Single-pass exclusive scan of the per-cell counts.  Block b of the scan
// covers cells [b kScanChunk, (b + 1) kScanChunk); the producer of the keys
// already summed each block's items, spread over kScanSubTotals counters per
// block (totals[b * kScanSubTotals + k]), so block b's offset is the sum of
// the first b * kScanSubTotals counters.  Every thread reads a few of them
// (all loads in flight together, along with the block's own counts, 16 bytes
// per thread) and the block reduces them with its scan.  The block writes
// cell_start, zeroes its counts for the next round of slot atomics and its
// counters of the other totals buffer for the next producer.
// ---------------------------------------------------------------------------

// Exclusive scan of v over the block, the block's total of v, and the block's
// sum of u.
__device__ __forceinline__ uint32_t block_excl_scan_1024(uint32_t v, uint32_t u, uint32_t &total,
                                                         uint32_t &usum) {
  __shared__ uint32_t s_wave[16], s_wu[16];
  __shared__ uint32_t s_total, s_usum;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t inc = hg_wave_incl_scan(v);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) u += __shfl_down(u, off, 64);
  if (lane == 63) s_wave[wave] = inc;
  if (lane == 0) s_wu[wave] = u;
  __syncthreads();
  if (wave == 0) {
    const uint32_t wv = lane < 16 ? s_wave[lane] : 0u;
    const uint32_t winc = hg_wave_incl_scan(wv);
    uint32_t wu = lane < 16 ? s_wu[lane] : 0u;
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) wu += __shfl_down(wu, off, 64);
    if (lane < 16) s_wave[lane] = winc - wv;  // exclusive wave offsets
    if (lane == 15) s_total = winc;
    if (lane == 0) s_usum = wu;
  }
  __syncthreads();
  total = s_total;
  usum = s_usum;
  return s_wave[wave] + inc - v;
}

__global__ __launch_bounds__(kScanThreads) void scan_kernel(uint32_t *counts, uint32_t n,
                                                            const uint32_t *totals,
                                                            uint32_t *totals_clear,
                                                            uint32_t *cell_start, uint32_t nitems,
                                                            int32_t *err) {
  const uint32_t bid = blockIdx.x;
  const uint32_t base = bid * uint32_t(kScanChunk) + threadIdx.x * uint32_t(kScanItems);
  const bool full = base + uint32_t(kScanItems) <= n;
  uint32_t v[kScanItems];
  if (full) {
    const uint4 q = *reinterpret_cast<const uint4 *>(counts + base);
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
    v[3] = q.w;
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) v[k] = (base + k < n) ? counts[base + k] : 0u;
  }
  // this thread's share of the counters of the blocks before this one
  const uint32_t nbefore = bid * uint32_t(kScanSubTotals);
  uint32_t before = 0;
  for (uint32_t j = threadIdx.x; j < nbefore; j += uint32_t(kScanThreads)) before += totals[j];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) s += v[k];
  uint32_t total = 0, offset = 0;
  const uint32_t excl = block_excl_scan_1024(s, before, total, offset);
  if (threadIdx.x < uint32_t(kScanSubTotals)) {
    // the block's own counters must add up to its counts
    uint32_t mine = totals[nbefore + threadIdx.x];
#pragma unroll
    for (int off = kScanSubTotals / 2; off > 0; off >>= 1) mine += __shfl_down(mine, off, 64);
    if (threadIdx.x == 0 && mine != total) atomicOr(err, kSortTotalsBit);
    totals_clear[nbefore + threadIdx.x] = 0u;
  }

  // Starts are clamped to the item count (a no-op unless the totals were
  // wrong), so every later index through cell_start stays in bounds.
  uint32_t run = offset + excl;
  if (full) {
    uint4 q;
    q.x = min(run, nitems);
    q.y = min(run + v[0], nitems);
    q.z = min(run + v[0] + v[1], nitems);
    q.w = min(run + v[0] + v[1] + v[2], nitems);
    *reinterpret_cast<uint4 *>(cell_start + base) = q;
    // counts that are already zero (most cells are empty) are not rewritten
    if ((v[0] | v[1] | v[2] | v[3]) != 0u)
      *reinterpret_cast<uint4 *>(counts + base) = make_uint4(0u, 0u, 0u, 0u);
    if (base + uint32_t(kScanItems) == n) cell_start[n] = min(run + s, nitems);
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      if (base + k < n) {
        cell_start[base + k] = min(run, nitems);
        if (v[k]) counts[base + k] = 0u;
        if (base + k == n - 1) cell_start[n] = min(run + v[k], nitems);
      }
      run += v[k];
    }
  }
}

__global__ __launch_bounds__(256) void zero_kernel(uint32_t *p, uint32_t n) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__global__ __launch_bounds__(256) void scatter_kernel(const uint32_t *key, const uint32_t *slot,
                                                      const uint32_t *cell_start, uint32_t n,
                                                      uint32_t *perm, uint32_t *pinv) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) hg_scatter_item(key, slot, cell_start, n, perm, pinv, i);
}

}  // namespace

void hg_counting_scan(KernelProfiler &prof, hipStream_t stream, uint32_t *counts, uint32_t ncell,
                      const uint32_t *totals_in, uint32_t *totals_clear, uint32_t nscan,
                      uint32_t *cell_start, uint32_t n, int32_t *err) {
  if (nscan) {
    ProfScope ps(prof, "scan", stream);
    hipLaunchKernelGGL(scan_kernel, dim3(nscan), dim3(kScanThreads), 0, stream, counts, ncell,
                       totals_in, totals_clear, cell_start, n, err);
  }
}

void hg_counting_sort(KernelProfiler &prof, hipStream_t stream, uint32_t *counts, uint32_t ncell,
                      const uint32_t *totals_in, uint32_t *totals_clear, uint32_t nscan,
                      uint32_t *cell_start, const uint32_t *key, const uint32_t *slot, uint32_t n,
                      uint32_t *perm, int32_t *err, uint32_t *pinv) {
  hg_counting_scan(prof, stream, counts, ncell, totals_in, totals_clear, nscan, cell_start, n, err);
  if (n) {
    ProfScope ps(prof, "scatter", stream);
    hipLaunchKernelGGL(scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, key, slot,
                       cell_start, n, perm, pinv);
  }
}

void hg_zero_words(hipStream_t stream, uint32_t *p, uint32_t n) {
  if (n) hipLaunchKernelGGL(zero_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
}
