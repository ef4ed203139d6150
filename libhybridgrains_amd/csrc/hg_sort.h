// hg_sort.h -- counting sort of items by cell key (shared by MPM and DEM).
//
// Items carry a key (their cell, offset by their scene's cell_off) and a slot
// (the value atomicAdd on counts[key] returned when the key was produced).
// The producer of the keys also adds each item to the total of its scan
// block (hg_scan_block_count: key / kScanChunk), so that the sort's single
// scan pass finds every block's offset without waiting for other blocks.
// hg_counting_sort scans the counts into cell_start (ncell + 1 entries),
// zeroes the counts for the next round and scatters item indices into perm,
// so that perm[cell_start[c] .. cell_start[c+1]) lists the items of cell c;
// pinv (optional) receives each item's position in that order.
//
// Block totals (kScanSubTotals words per scan block) are double-buffered: the producer adds into totals_in, which
// the sort reads, and the sort zeroes totals_clear (the buffer the previous
// sort read) for the producer after it.  Each owner alternates the two
// buffers with a parity that flips once per produced-and-sorted round.
#pragma once

#include "hg_internal.h"

constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanChunk = kScanThreads * kScanItems;

// Each scan block's item total is kept in this many counters, which the
// producer's workgroups use in turn, so that the waves adding to one block at
// the same time rarely hit the same word.
constexpr int kScanSubTotals = 8;

// Raised in err[0] if a scan block's counts do not add up to the total its
// producer reported (a producer that skipped hg_scan_block_count); the
// sort's results are then wrong.  Never expected.
constexpr int32_t kSortTotalsBit = 0x100;

// Kernels: scan, scatter (2 launches).  hg_counting_scan is the first of the
// two alone, for a caller that scatters with hg_scatter_item inside a kernel
// of its own.
void hg_counting_scan(KernelProfiler &prof, hipStream_t stream, uint32_t *counts, uint32_t ncell,
                      const uint32_t *totals_in, uint32_t *totals_clear, uint32_t nscan,
                      uint32_t *cell_start, uint32_t n, int32_t *err);
void hg_counting_sort(KernelProfiler &prof, hipStream_t stream, uint32_t *counts, uint32_t ncell,
                      const uint32_t *totals_in, uint32_t *totals_clear, uint32_t nscan,
                      uint32_t *cell_start, const uint32_t *key, const uint32_t *slot, uint32_t n,
                      uint32_t *perm, int32_t *err, uint32_t *pinv = nullptr);

// Zeroes n words on stream (a kernel, so that it may be captured in a graph).
void hg_zero_words(hipStream_t stream, uint32_t *p, uint32_t n);

static inline uint32_t hg_scan_blocks(uint32_t ncell) {
  return (ncell + kScanChunk - 1) / kScanChunk;
}
// words of one totals buffer
static inline uint32_t hg_scan_totals_words(uint32_t nscan) {
  return nscan * uint32_t(kScanSubTotals);
}

// Adds each active lane's item to the total of its scan block
// (key / kScanChunk), in that block's counter sub (any value; taken modulo
// kScanSubTotals, normally the producer's workgroup index): one atomicAdd
// per distinct block of the wavefront, by the block's lowest lane.  Call
// with every lane that produces a key active; other lanes take no part.
__device__ __forceinline__ void hg_scan_block_count(uint32_t *totals, uint32_t key, uint32_t sub) {
  const int lane = int(threadIdx.x & 63u);
  const uint32_t blk = key / uint32_t(kScanChunk) * uint32_t(kScanSubTotals) +
                       sub % uint32_t(kScanSubTotals);
  uint64_t pending = __ballot(1);
  while (pending) {
    const int l = __ffsll((unsigned long long)pending) - 1;
    const uint32_t bl = uint32_t(__shfl(int(blk), l, 64));
    const uint64_t same = __ballot(blk == bl) & pending;
    if (lane == l) atomicAdd(&totals[bl], uint32_t(__popcll(same)));
    pending &= ~same;
  }
}

// Places item i (< n) at cell_start[key[i]] + slot[i] of perm, and that
// position in pinv (if given).  A position past n arises only from wrong
// totals (flagged by the scan) and is dropped.
__device__ __forceinline__ void hg_scatter_item(const uint32_t *key, const uint32_t *slot,
                                                const uint32_t *cell_start, uint32_t n,
                                                uint32_t *perm, uint32_t *pinv, uint32_t i) {
  const uint32_t pos = cell_start[key[i]] + slot[i];
  if (pos < n) perm[pos] = i;
  if (pinv) pinv[i] = pos < n ? pos : i;
}
