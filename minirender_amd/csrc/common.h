// common.h -- device-side tables of the rasterization pipeline (gfx950).
//
// A batch is a set of jobs (frames).  Every kernel works over the whole batch:
// the plain Renderer::render() path is a batch of one job.  All tables below
// live in device memory and are written by the host once per batch.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define MR_HD __host__ __device__
#else
#define MR_HD
#endif

namespace mr {

// One draw (Renderable + material) of a job.  The modelview and normal matrix
// rows 0..2 are what Renderer::paintMesh multiplies by (Renderer.cpp:336-348).
struct DrawDesc {
  float mv[12];  // modelview rows 0..2, row-major
  float nm[12];  // normal matrix rows 0..2
  float diffuse[3];
  float specular[3];
  float emissive[3];
  float shininess;
  const float* pos;    // positions, 3 floats each
  const float* nrm;    // normals, 3 floats each
  const float* uv;     // texcoords, 2 floats each (null: uv = 0)
  const int32_t* idx;  // 3 per triangle
  const int32_t* nidx;
  const int32_t* tidx;  // null when the mesh has no texcoords
  const void* tex;      // null unless texturing applies (Renderer.cpp:228)
  int32_t tex_format;
  int32_t tex_rows, tex_cols;
  int32_t ntri;
  int32_t job;
  int32_t tile_base;    // the job's first global tile and tiles per row
  int64_t tri_begin;    // first flat triangle index of this draw in the batch
  uint32_t order_base;  // 2 * triangles of earlier draws of the same job
  int32_t tiles_x;
  int64_t vbase;        // first projected-vertex record of this draw's mesh
  int32_t no_clip;      // 1: no near-plane test (paintTriangle, world = false)
  int32_t pad_;
};

// One frame of a batch.
struct JobDesc {
  float P[16];  // projection, row-major
  float znear;  // Renderer.cpp:324-326
  int32_t persp;
  int32_t W, H;
  int32_t tiles_x, tiles_y;
  int32_t tile_base;  // first global tile of this job
  int32_t draw_begin, draw_end;
  float bg[3];
  float light[3];
  int32_t light_is_point;
  float ambient;
  int32_t lighting;
  int32_t save_normals;
  int32_t shade_base;  // first shading block of this job
  uint32_t* win;       // device scratch: draw-order id of each pixel's winner
  uint8_t* rgb8;       // device outputs (any may be null)
  float* depth;
  uint32_t* winner;
  float* rgb;
  float* normals;
  float* points;       // range image (Renderer.cpp:388-415)
};

// Projected-vertex record written once per vertex by the vertex kernel:
// {px, py, zv, ez}, the projection (Renderer.cpp:191-203) and eye depth
// (:344-345) of the vertex, all that triangle setup reads for a triangle in
// front of the near plane.  The eye x, y that the near-plane clip needs are
// transformed again from the positions where it runs.
constexpr int kVrecFloats = 4;

// A surviving (sub-)triangle as the binner sees it: its id (2 * flat
// triangle index + clip sub-index), its draw and clamped pixel bbox.
struct BinItem {
  uint32_t id;
  uint32_t draw;
  uint16_t bx0, bx1, by0, by1;
};
static_assert(sizeof(BinItem) == 16, "BinItem must be 16 bytes");

// Device counters of a batch (64-bit, cleared by the clear kernel).
enum Counter : int {
  kItems = 0,        // surviving (sub-)triangles (statistics)
  kScanTicket = 1,
  kBinOverflow = 2,  // set when the bin list was too small
  kError = 3,        // set when a bounded wait hit its limit
  kClipped = 4,      // input triangles that crossed the near plane
  kTileTicket = 5,
  kBinEntries = 6,   // total (tile, triangle) pairs (written by the scan)
  kParts = 7,        // work items of split tiles (written by the scan)
  kHeavy = 8,        // tiles split over several work items
  kHeavyOverflow = 9,
  kListed = 10,      // tiles in the raster's tile list (written by the scan)
  kNumCounters = 16
};

// Workgroup index b in [0, n) -> work index, a bijection under which each of
// the 8 XCDs (which receive workgroups round-robin) gets one contiguous range
// of work: neighbouring tiles share vertices, and each XCD has its own L2.
// XCD x = b % 8 owns q + (x < n % 8) items starting at x * q + min(x, n % 8).
MR_HD inline int64_t xcd_remap(int64_t b, int64_t n) {
  constexpr int64_t kXcds = 8;
  const int64_t q = n / kXcds, rem = n % kXcds;
  const int64_t x = b % kXcds, k = b / kXcds;
  return x * q + (x < rem ? x : rem) + k;
}

// Per-tile counters are spread one per 64-byte line: the atomics of the
// binner then spread over all memory channels instead of a few.
constexpr int kCtrStride = 16;
constexpr int kCtrStride64 = kCtrStride / 2;  // the same spacing in 64-bit words

// The edge-checkpoint buffer is cut into this many segments, each with its
// own allocation counter (kCtrStride words apart): one shared counter would
// serialize every binning wavefront's atomic.
constexpr int kCkptSegs = 64;

constexpr int kRegion = 128;       // item slots per geometry wavefront (64 x 2)
constexpr int kShadeBlock = 256;   // pixels per deferred-shading workgroup
constexpr int kShadeW = 16;        // shading block: kShadeW x kShadeH pixels
constexpr int kShadeH = kShadeBlock / kShadeW;
constexpr int kMaxLdsDraws = 256;  // draw-table entries cached in LDS

// Bounded waits (decoupled look-back): polls before giving up with kError.
constexpr uint32_t kSpinLimit = 1u << 22;

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;    // elements per thread in the tile-count scan
constexpr uint32_t kNoWinner = 0xFFFFFFFFu;
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
constexpr uint32_t kEmptyTile = 0xFFFFFFFEu;  // tile_heavy of a tile with no bin entry
// A raster work item (parts list entry .y) that covers a band of the tile's
// rows with all of its entries: bit 31 set, band count in bits 8-15, band
// index in bits 0-7.  Other entries are entry-range parts (the part index).
constexpr uint32_t kBandPart = 0x80000000u;
constexpr float kClearDepth = 1e11f;  // Renderer.cpp:116
// Job lookup hints: one per kHintSpan consecutive tiles or shading blocks.
constexpr int kHintShift = 6;
constexpr int kHintSpan = 1 << kHintShift;
constexpr int kMaxDim = 16384;        // max width / height

}  // namespace mr
