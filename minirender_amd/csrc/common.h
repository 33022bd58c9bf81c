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
  int32_t pad0;
  int64_t tri_begin;    // first flat triangle index of this draw in the batch
  uint32_t order_base;  // 2 * triangles of earlier draws of the same job
  uint32_t pad1;
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
  int32_t pad0;
  uint8_t* rgb8;     // device outputs (any may be null)
  float* depth;
  uint32_t* winner;
  float* rgb;
  float* normals;
};

// Raster record of one surviving (sub-)triangle: 64 bytes, written by the
// geometry kernel in draw order (order-preserving compaction), so its index
// is a draw-order key for the depth tie-break.
struct TriRec {
  float n1x, n1y, n2x, n2y;  // edge normals, Renderer.cpp:215-216
  float p0x, p0y, p2x, p2y;  // pixel positions of vertices 0 and 2
  float z0, z1, z2;          // -1/z_eye (perspective) or NDC z (ortho)
  uint32_t tri_sub;          // tri << 2 | clipped << 1 | sub
  uint16_t bx0, bx1, by0, by1;  // clamped pixel bbox (inclusive), bytes 48..55
  uint32_t draw;             // index into the batch's draw table
  uint32_t job;              // bytes 60..63: the binner reads 48..63 as one uint4
};
static_assert(sizeof(TriRec) == 64, "TriRec must be 64 bytes");

// Device counters of a batch (64-bit, cleared by the clear kernel).
enum Counter : int {
  kGeomTicket = 0,
  kScanTicket = 1,
  kRecords = 2,      // total records (written by the last geometry chunk)
  kBinOverflow = 3,  // set when the bin list was too small
  kError = 4,        // set when a bounded wait hit its limit
  kClipped = 5,      // input triangles that crossed the near plane
  kTileTicket = 6,
  kBinEntries = 7,   // total (tile, record) pairs (written by the scan)
  kNumCounters = 16
};

// Bounded waits (decoupled look-back): polls before giving up with kError.
constexpr uint32_t kSpinLimit = 1u << 22;

constexpr int kGeomChunk = 256;  // triangles per geometry chunk (1 per thread)
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;    // elements per thread in the tile-count scan
constexpr uint32_t kNoWinner = 0xFFFFFFFFu;
constexpr float kClearDepth = 1e11f;  // Renderer.cpp:116
constexpr int kMaxDim = 16384;        // max width / height

}  // namespace mr
