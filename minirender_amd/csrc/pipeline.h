// pipeline.h -- launchers of the rasterization kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"

namespace mr {

// Device buffers and sizes of one batch.
struct BatchArgs {
  TriRec* recs;             // capacity 2 * total_tris
  uint32_t* tile_count;     // [ntiles]
  uint32_t* tile_off;       // [ntiles]
  uint32_t* tile_cursor;    // [ntiles]
  uint32_t* bins;           // [bins_cap]
  uint64_t bins_cap;
  unsigned long long* geom_status;  // [geom_chunks] decoupled look-back words
  unsigned long long* scan_status;  // [scan_blocks]
  unsigned long long* counters;  // [kNumCounters]
  const DrawDesc* draws;
  int32_t ndraws;
  const JobDesc* jobs;
  int32_t njobs;
  int64_t total_tris;
  int32_t ntiles;
  int32_t tile_w, tile_h;   // each 8, 16 or 32
  int32_t geom_chunks;
  int32_t scan_blocks;
  int32_t num_cus;
};

inline int32_t geom_chunks_for(int64_t tris) {
  return (int32_t)((tris + kGeomChunk - 1) / kGeomChunk);
}
inline int32_t scan_blocks_for(int32_t tiles) {
  return (tiles + kScanBlock * kScanItems - 1) / (kScanBlock * kScanItems);
}

hipError_t launch_clear(const BatchArgs& a, hipStream_t s);
hipError_t launch_geometry(const BatchArgs& a, hipStream_t s);
hipError_t launch_scan(const BatchArgs& a, hipStream_t s);
hipError_t launch_bin(const BatchArgs& a, hipStream_t s);
hipError_t launch_tile(const BatchArgs& a, hipStream_t s);
bool tile_supported(int tile_w, int tile_h);

}  // namespace mr
