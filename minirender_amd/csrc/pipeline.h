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
  // [ntiles * kCtrStride64] per tile: (walk work << 32) | (tile, triangle)
  // pairs, where a pair's walk work is its bbox pixels in the tile + 4 per
  // row + 16 (for the group leader's pair, times the group size)
  unsigned long long* tile_count;
  uint32_t* tile_off;        // [ntiles] exclusive scan of the pair counts
  uint32_t* tile_cursor;     // [ntiles] scatter cursors
  uint32_t* tile_heavy;      // [ntiles] merge slot of a split tile, kNoSlot, or kEmptyTile
  uint4* tile_list;          // [ntiles] {tile, first bin entry, entries, 0} of the tiles
                             // rastered whole (first counters[kListed])
  uint32_t* tile_done;       // [ntiles] finished parts of a split tile
  uint2* parts;              // [parts_cap] (tile, part) of split tiles
  uint32_t parts_cap;
  uint2* bins;               // [bins_cap] (triangle id, draw) grouped by tile
  uint64_t bins_cap;
  // Edge checkpoints: the edge values of each row of a triangle spanning
  // several tile columns where the row enters each of them after the first
  // (written by the row pre-pass, read by the raster and shading).
  // Checkpoint of (row, tile
  // column tx) of a survivor id: ckpt[ckpt_base[id].x + row *
  // ckpt_base[id].y + tx], with .x a signed offset (the entries of the first
  // column are allocated, never written); .y == 0: no checkpoints.
  int2* ckpt_base;           // [2 * total_tris]
  float2* ckpt;              // [ckpt_cap], kCkptSegs equal segments
  uint64_t ckpt_cap;
  unsigned long long* ckpt_ctr;  // [kCkptSegs * kCtrStride] segment fill counters
  // Row tasks of the row pre-pass, (item index << 32) | pixel row, in
  // kCkptSegs segments of (ckpt_cap / kCkptSegs) / 2 (a task per row of an
  // item with checkpoints, each row holding two entries or more); the
  // segment counters ckpt_ctr count them in their high 24 bits.
  unsigned long long* ckpt_rows;
  unsigned long long* heavy_keys;  // [heavy_cap * tile pixels] merge buffers
  uint32_t heavy_cap;
  uint32_t frag_size;        // max (tile, triangle) pairs per raster work item
  uint32_t band_work;        // a tile with more walk work is rastered in row bands
  unsigned long long* scan_status;  // [scan_blocks] decoupled look-back words
  unsigned long long* counters;     // [kNumCounters]
  // The counters' copy in pinned host memory (device address), written by
  // the shading kernel's first workgroup once they are final; null: the
  // host copies them back itself.
  unsigned long long* host_counters;
  // Per-workgroup statistics (no shared counter: same-address atomics
  // serialize): [0, kStatSlots) clipped triangles per geometry workgroup,
  // [kStatSlots, 2 kStatSlots) binned triangles per bin workgroup.
  unsigned long long* block_stats;
  float4* vscr;               // [total_verts] {px, py, zv, ez} of each projected vertex
  int64_t total_verts;
  const int64_t* vg_begin;    // [nvgroups] first record of each vertex group
  const int32_t* vg_draw;     // [nvgroups] a draw of the group (mesh, modelview, job)
  int32_t nvgroups;
  const DrawDesc* draws;
  const int64_t* draw_begin;  // [ndraws] tri_begin of each draw
  int32_t ndraws;
  const JobDesc* jobs;
  const int32_t* job_tile_base;   // [njobs]
  const int32_t* job_shade_base;  // [njobs]
  // Job hints: entry i = {job of tile (resp. shading block) i * kHintSpan,
  // first tile (block) of the next job or INT32_MAX}.
  const int2* tile_hint;          // [ceil(ntiles / kHintSpan)]
  const int2* shade_hint;         // [ceil(shade_blocks / kHintSpan)]
  int32_t njobs;
  int64_t total_tris;
  int32_t ntiles;
  int32_t tile_w, tile_h;    // each 8, 16 or 32
  int32_t raster_threads;    // workgroup size of the tile raster
  int32_t scan_blocks;
  int32_t shade_blocks;
  int32_t num_cus;
  int32_t geom_blocks;       // resident geometry workgroups (<= kStatSlots)
  // Where the edge checkpoints are taken (launched stages; the persistent
  // kernel always takes them in its binning stage): 1, by the scan's launch
  // in workgroups beside the scan, with the row pre-pass at the end of the
  // binning kernel (a batch of at most kCkptScanRegions regions, whose items
  // that pass reads while the one-workgroup scan runs); 0, by the binning
  // kernel, which reads the items anyway, with the row pre-pass a launch of
  // its own.
  int32_t ckpt_in_scan;
};

constexpr int64_t kCkptScanRegions = 65536;

constexpr int kStatSlots = 8192;  // >= the geometry and bin grids

inline int32_t scan_blocks_for(int32_t tiles) {
  return (tiles + kScanBlock * kScanItems - 1) / (kScanBlock * kScanItems);
}

// Grid sizes of the grid-stride kernels (also used to sum their statistics).
// Workgroups of the geometry kernel, each taking one contiguous range of
// triangles: as many as the GPU holds at once (geom_blocks, from the
// kernel's occupancy), so none waits for a second round; a batch of more than
// 4096 triangles per such workgroup gets ~1024-triangle ranges instead, whose
// rounds even out the uneven draws of many jobs.
__host__ __device__ inline int32_t geom_grid(const BatchArgs& a) {
  const int64_t blocks = (a.total_tris + 255) / 256;
  int64_t g = a.geom_blocks;
  if (a.total_tris > g * 4096) {
    g = (a.total_tris + 1023) / 1024;
    if (g > kStatSlots) g = kStatSlots;
  }
  return (int32_t)(blocks < g ? blocks : g);
}
__host__ __device__ inline int32_t bin_grid(const BatchArgs& a) {
  int64_t blocks = (a.nregions + 3) / 4;
  if (blocks < 1) blocks = 1;
  const int64_t cap = (int64_t)a.num_cus * 7;  // one resident round (7 per CU)
  return (int32_t)(blocks < cap ? blocks : cap);
}

hipError_t launch_clear(const BatchArgs& a, hipStream_t s);
hipError_t launch_vertex(const BatchArgs& a, hipStream_t s);
hipError_t launch_geometry(const BatchArgs& a, hipStream_t s);
hipError_t launch_scan(const BatchArgs& a, hipStream_t s);
hipError_t launch_bin(const BatchArgs& a, hipStream_t s);
hipError_t launch_rows(const BatchArgs& a, hipStream_t s);  // nothing when ckpt_in_scan
hipError_t launch_tile(const BatchArgs& a, hipStream_t s);
// Geometry workgroups one CU holds at once (from the kernel's resources).
int geom_blocks_per_cu();
hipError_t launch_shade(const BatchArgs& a, hipStream_t s);
// The no-alloc entry point's scratch: kNoAllocScratchBytes of device memory
// per device (a static array).  noalloc_scratch_init(device), with that
// device current, looks up its address once (0, or -1 on failure);
// noalloc_scratch(device) returns it (null before the init).
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;
int noalloc_scratch_init(int device);
void* noalloc_scratch(int device);
// ---- the persistent kernel: every stage of a batch in one launch ----------
// Stages in queue order.
enum PersistStage : int32_t {
  kPClear, kPVertex, kPGeom, kPScan, kPBin, kPRows, kPTile, kPShade, kPStages
};
constexpr int kPQueues = 8;      // work queues (one per XCD under round-robin dispatch)
constexpr int kCtlStride = 32;   // u64 words between two control words (256 B)
// Control words (each kCtlStride words apart): queue heads and completed
// items per (stage, queue), queues completed per stage, completion flags per
// (stage, queue), workgroups finished, error.  Zero between launches: the
// last workgroup of a launch clears them.
constexpr int kCtlHead = 0;
constexpr int kCtlDone = kCtlHead + kPStages * kPQueues;
constexpr int kCtlFin = kCtlDone + kPStages * kPQueues;
constexpr int kCtlFlag = kCtlFin + kPStages;
constexpr int kCtlExit = kCtlFlag + kPStages * kPQueues;
constexpr int kCtlErr = kCtlExit + 1;
constexpr int kCtlWords = kCtlErr + 1;
constexpr size_t kCtlBytes = (size_t)kCtlWords * kCtlStride * 8;

struct PersistPlan {
  int32_t nvb[kPStages];    // workgroups of the stage's own launch (0: not run)
  int32_t depth[kPStages];  // items per workgroup a stage is dealt out in (>= 1)
  int32_t wait[kPStages];   // 1: the stage waits until every earlier stage is complete
  unsigned long long* ctl;  // kCtlBytes of control words, zero before the launch
  unsigned long long time_limit;  // wall-clock ticks one wait may take (0: every wait gives up)
  // Profiling (null: off): per stage s, words kStampStride * (2 s) and
  // kStampStride * (2 s + 1) take the maximum over workgroups of ~(first
  // wall-clock tick of the stage) and of its last tick; zero before the launch.
  unsigned long long* stamps;
};
constexpr int kStampStride = 32;
constexpr size_t kStampBytes = (size_t)2 * kPStages * kStampStride * 8;
// Launch the persistent kernel for a batch (16 x 16 tiles), on a grid of the
// workgroups the GPU holds at once.  `clear`: run the clear stage.
// wait_us: how long one wait may take, in microseconds (0: every wait gives
// up at once, which tests of the failure path use).
hipError_t launch_persistent(const BatchArgs& a, bool clear, unsigned long long* ctl,
                             unsigned long long* stamps, uint64_t wait_us, hipStream_t s);
// Wall-clock ticks per second of the device's clock (wall_clock64).
unsigned long long wall_clock_hz();
bool persistent_supported(int tile_w, int tile_h);

bool tile_supported(int tile_w, int tile_h);
// Raster workgroup sizes built per tile size: 8x8 {64, 128}, 16x16 and
// 32x32 {64, 128, 256, 512}; the default is 64 for 8x8, else 256.
bool raster_threads_supported(int tile_w, int threads);
int default_raster_threads(int tile_w);

}  // namespace mr
