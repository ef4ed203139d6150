// hg_sort.h -- counting sort of items by cell key (shared by MPM and DEM).
//
// Items carry a key (their cell, offset by their scene's cell_off) and a slot
// (the value atomicAdd on counts[key] returned when the key was produced).
// hg_counting_sort scans the counts into cell_start (ncell + 1 entries),
// zeroes the counts for the next round and scatters item indices into perm,
// so that perm[cell_start[c] .. cell_start[c+1]) lists the items of cell c;
// pinv (optional) receives each item's position in that order.
#pragma once

#include "hg_internal.h"

constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanChunk = kScanThreads * kScanItems;

// Kernels: scan_reduce, scan_apply, scatter (3 launches).
void hg_counting_sort(KernelProfiler &prof, hipStream_t stream, uint32_t *counts, uint32_t ncell,
                      uint32_t *block_sums, uint32_t nscan, uint32_t *cell_start,
                      const uint32_t *key, const uint32_t *slot, uint32_t n, uint32_t *perm,
                      uint32_t *pinv = nullptr);

static inline uint32_t hg_scan_blocks(uint32_t ncell) {
  return (ncell + kScanChunk - 1) / kScanChunk;
}
