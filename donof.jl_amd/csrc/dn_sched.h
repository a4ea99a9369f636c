// Persistent schedule of the grouped GEMM launches (FP64 and split precision).
#pragma once
#include <hip/hip_runtime.h>

#include "dn_internal.h"

namespace dn {

// Problem i of a launch owns tiles [0, n_i), n_i = tiles_m * tiles_n * nsplit, and the
// descriptors carry the prefix sums tile_begin (of n_i) and slot_begin (of ceil(n_i / 8)).
// With a grid that is a multiple of the 8 XCDs, workgroup b runs on XCD x = b % 8 and
// takes slots b / 8, b / 8 + grid / 8, ... of that XCD's list; slot s belongs to the
// problem i with slot_begin_i <= s < slot_begin_{i+1} and is its tile
// x * ceil(n_i / 8) + (s - slot_begin_i).  Every XCD thus works through one contiguous
// eighth of every problem: the M-tiles that share a B block (and consecutive B blocks)
// meet in the same L2, and a large problem is spread over all XCDs instead of landing on
// the one or two that a contiguous split of the whole list would give it.  Without the
// XCD multiple, workgroup b takes list positions b, b + grid, ... in tile_begin order.
// The schedule is static, so every workgroup's loop ends.  tile(desc, mt, nt, z) runs one
// output tile; the workgroup synchronises before the next one reuses LDS.
template <typename Desc, typename Tile>
__device__ __forceinline__ void grouped_schedule(const Desc* __restrict__ desc, int nprob,
                                                 long total_tiles, long total_slots,
                                                 Tile&& tile) {
  const bool chunk = (gridDim.x % 8u) == 0;
  const long x = blockIdx.x % 8u;
  const long first = chunk ? blockIdx.x / 8u : blockIdx.x;
  const long step = chunk ? gridDim.x / 8u : gridDim.x;
  const long span = chunk ? total_slots : total_tiles;
  for (long pos = first; pos < span; pos += step) {
    int lo = 0, hi = nprob - 1;  // the last problem starting at or before pos
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if ((chunk ? desc[mid].slot_begin : desc[mid].tile_begin) <= pos) lo = mid;
      else hi = mid - 1;
    }
    // lo is uniform over the workgroup; saying so makes the descriptor loads scalar (SGPRs)
    // instead of per-lane vector loads.  A private copy: the tile's stores cannot alias
    // it, so it is not reloaded after every store.
    lo = __builtin_amdgcn_readfirstlane(lo);
    const Desc d = desc[lo];
    const long n = (long)d.tiles_m * d.tiles_n * d.nsplit;
    const long local = chunk ? x * ((n + 7) / 8) + (pos - d.slot_begin) : pos - d.tile_begin;
    if (local >= n) continue;  // the last XCD's share of a problem can be short
    const int mt = local % d.tiles_m;
    const int nt = (local / d.tiles_m) % d.tiles_n;
    const int z = local / ((long)d.tiles_m * d.tiles_n);
    tile(d, mt, nt, z);
    __syncthreads();
  }
}

// Tile counts and prefix sums of desc[0..n) (any descriptor with the fields above) for
// tiles of bm x bn; nsplit is set by the caller.  Returns the total tile count.
template <typename Desc>
long grouped_offsets(Desc* desc, int n, int bm, int bn) {
  long begin = 0, slot = 0;
  for (int i = 0; i < n; ++i) {
    desc[i].tiles_m = (desc[i].g.M + bm - 1) / bm;
    desc[i].tiles_n = (desc[i].g.N + bn - 1) / bn;
    desc[i].tile_begin = begin;
    desc[i].slot_begin = slot;
    const long tiles = (long)desc[i].tiles_m * desc[i].tiles_n * desc[i].nsplit;
    begin += tiles;
    slot += (tiles + 7) / 8;
  }
  return begin;
}

// Slots of one XCD's list for descriptors planned by grouped_offsets.
template <typename Desc>
long grouped_slots(const Desc* desc, int n) {
  const Desc& last = desc[n - 1];
  return last.slot_begin + ((long)last.tiles_m * last.tiles_n * last.nsplit + 7) / 8;
}

// Persistent launch of a grouped kernel over `total` tiles (`slots` per XCD list): grid
// <= 0 means -grid times as many workgroups as stay resident (0 = once), a multiple of the
// 8 XCDs, at most one per tile (rounded up to the XCD multiple).
template <typename Desc, typename Kern>
int launch_persistent(hipStream_t s, Kern kern, int threads, const Desc* d, int n, long total,
                      long slots, int grid) {
  if (grid <= 0) {
    int dev = 0, cus = 0, occ = 0;
    DN_HIP(hipGetDevice(&dev));
    DN_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    DN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, threads, 0));
    const long full = (long)std::max(occ, 1) * cus * std::max(-grid, 1);
    grid = (int)std::min(std::min(full, (total + 7) / 8 * 8), 0x7ffffff8L);
  }
  hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), 0, s, d, n, total, slots);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

}  // namespace dn
