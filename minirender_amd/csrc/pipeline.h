// pipeline.h -- launchers of the rasterization kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"

namespace mr {

// Device buffers and sizes of one batch.
struct BatchArgs {
  BinItem* items;            // [nregions * kRegion] survivors, per wavefront
  uint32_t* region_count;    // [nregions] survivors of each wavefront
  int64_t nregions;
  uint32_t* tile_count;      // [ntiles] (tile, triangle) pairs per tile
  uint32_t* tile_off;        // [ntiles] exclusive scan of tile_count
  uint32_t* tile_cursor;     // [ntiles] scatter cursors
  uint32_t* tile_heavy;      // [ntiles] merge slot of a split tile or kNoSlot
  uint32_t* tile_done;       // [ntiles] finished parts of a split tile
  uint2* parts;              // [parts_cap] (tile, part) of split tiles
  uint32_t parts_cap;
  uint32_t* bins;            // [bins_cap] triangle ids grouped by tile
  uint64_t bins_cap;
  unsigned long long* heavy_keys;  // [heavy_cap * tile pixels] merge buffers
  uint32_t heavy_cap;
  uint32_t frag_size;        // max (tile, triangle) pairs per raster work item
  unsigned long long* scan_status;  // [scan_blocks] decoupled look-back words
  unsigned long long* counters;     // [kNumCounters]
  const DrawDesc* draws;
  const int64_t* draw_begin;  // [ndraws] tri_begin of each draw
  int32_t ndraws;
  const JobDesc* jobs;
  const int32_t* job_tile_base;   // [njobs]
  const int32_t* job_shade_base;  // [njobs]
  int32_t njobs;
  int64_t total_tris;
  int32_t ntiles;
  int32_t tile_w, tile_h;    // each 8, 16 or 32
  int32_t scan_blocks;
  int32_t shade_blocks;
  int32_t num_cus;
};

inline int32_t scan_blocks_for(int32_t tiles) {
  return (tiles + kScanBlock * kScanItems - 1) / (kScanBlock * kScanItems);
}

hipError_t launch_clear(const BatchArgs& a, hipStream_t s);
hipError_t launch_geometry(const BatchArgs& a, hipStream_t s);
hipError_t launch_scan(const BatchArgs& a, hipStream_t s);
hipError_t launch_bin(const BatchArgs& a, hipStream_t s);
hipError_t launch_tile(const BatchArgs& a, hipStream_t s);
hipError_t launch_shade(const BatchArgs& a, hipStream_t s);
bool tile_supported(int tile_w, int tile_h);

}  // namespace mr
