// hg_sort.hip -- counting sort of items by cell key: a three-pass exclusive
// scan of the per-cell counts and a scatter of item indices.
#include "hg_sort.h"

namespace {

// ---------------------------------------------------------------------------
// Counting sort: exclusive scan of the per-cell counts (two passes: block
// totals, then block scans offset by the totals before them), then a scatter
// of point indices.  scan_apply also zeroes the counts for the next step's
// G2P.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t block_excl_scan_1024(uint32_t v, uint32_t &total) {
  __shared__ uint32_t s_wave[16];
  __shared__ uint32_t s_total;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t inc = hg_wave_incl_scan(v);
  if (lane == 63) s_wave[wave] = inc;
  __syncthreads();
  if (wave == 0) {
    const uint32_t wv = lane < 16 ? s_wave[lane] : 0u;
    const uint32_t winc = hg_wave_incl_scan(wv);
    if (lane < 16) s_wave[lane] = winc - wv;  // exclusive wave offsets
    if (lane == 15) s_total = winc;
  }
  __syncthreads();
  total = s_total;
  return s_wave[wave] + inc - v;
}

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t *counts,
                                                                   uint32_t n,
                                                                   uint32_t *block_sums) {
  const uint32_t base = blockIdx.x * kScanChunk + threadIdx.x * kScanItems;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) s += counts[base + k];
  uint32_t total = 0;
  (void)block_excl_scan_1024(s, total);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Each block first sums the totals of the blocks before it (at most
// kScanChunk of them, one per thread), so no separate pass over the block
// totals is needed.
__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(uint32_t *counts, uint32_t n,
                                                                  const uint32_t *block_sums,
                                                                  uint32_t *cell_start) {
  __shared__ uint32_t s_part[kScanThreads / 64];
  uint32_t before = 0;
  for (uint32_t b = threadIdx.x; b < blockIdx.x; b += kScanThreads) before += block_sums[b];
  {
    // block sum of 'before'
    uint32_t v = before;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) tot += s_part[w];
    before = tot;
    __syncthreads();
  }
  const uint32_t base = blockIdx.x * kScanChunk + threadIdx.x * kScanItems;
  uint32_t vals[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    vals[k] = (base + k < n) ? counts[base + k] : 0u;
    s += vals[k];
  }
  uint32_t total = 0;
  uint32_t run = block_excl_scan_1024(s, total) + before;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) {
      cell_start[base + k] = run;
      counts[base + k] = 0u;
      if (base + k == n - 1) cell_start[n] = run + vals[k];
    }
    run += vals[k];
  }
}

__global__ __launch_bounds__(256) void scatter_kernel(const uint32_t *key, const uint32_t *slot,
                                                      const uint32_t *cell_start, uint32_t n,
                                                      uint32_t *perm, uint32_t *pinv) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t pos = cell_start[key[i]] + slot[i];
  perm[pos] = i;
  if (pinv) pinv[i] = pos;
}

}  // namespace

void hg_counting_sort(KernelProfiler &prof, hipStream_t stream, uint32_t *counts, uint32_t ncell,
                      uint32_t *block_sums, uint32_t nscan, uint32_t *cell_start,
                      const uint32_t *key, const uint32_t *slot, uint32_t n, uint32_t *perm,
                      uint32_t *pinv) {
  {
    ProfScope ps(prof, "scan_reduce", stream);
    hipLaunchKernelGGL(scan_reduce_kernel, dim3(nscan), dim3(kScanThreads), 0, stream, counts,
                       ncell, block_sums);
  }
  {
    ProfScope ps(prof, "scan_apply", stream);
    hipLaunchKernelGGL(scan_apply_kernel, dim3(nscan), dim3(kScanThreads), 0, stream, counts,
                       ncell, block_sums, cell_start);
  }
  if (n) {
    ProfScope ps(prof, "scatter", stream);
    hipLaunchKernelGGL(scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, key, slot,
                       cell_start, n, perm, pinv);
  }
}
