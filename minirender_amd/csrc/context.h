// context.h -- host-side state behind the C-ABI (context.cpp, server.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "minirender_amd.h"
#include "common.h"
#include "pipeline.h"

namespace mr {

constexpr size_t kGraphCache = 32;
constexpr int kOutKinds = 6;  // rgb8, depth, winner, rgb, normals, points (mr_outputs order)
// Stages timed when profiling (mr_stats.stage_ms order): clear, geometry
// (vertex + geometry kernels), scan, bin, rows, raster, shade.
constexpr int kStages = 7;

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct HostBuf {  // pinned (hipHostMalloc)
  void* p = nullptr;
  size_t bytes = 0;
};

// A resident TriMesh (Scene.h:73-87).  Its attribute arrays (pos, nrm, uv)
// are its own, or,
// for a mesh uploaded with mr_mesh_upload_shared, those of mesh `owner`
// (then its own pos/nrm/uv stay empty).  An owner keeps its attribute
// arrays, and its slot, after its release until `sharers` drops to zero.
struct MeshRec {
  DevBuf pos, nrm, uv, idx, nidx, tidx;
  int32_t npos = 0, nnrm = 0, nuv = 0, ntri = 0;
  int32_t owner = -1;   // mesh holding the attribute arrays, -1: this one
  int32_t sharers = 0;  // meshes using this mesh's attribute arrays
  bool has_uv = false;
  bool nidx_alias = false;  // normal indices == position indices
  bool tidx_alias = false;  // uv indices == position indices
  bool alive = false;
};

struct TexRec {
  DevBuf texels;
  int32_t fmt = 0, rows = 0, cols = 0;
  bool alive = false;
};

// Device scratch of one batch in flight, and its pinned staging.
struct Scratch {
  DevBuf items, region_count, tile_count, tile_off, tile_cursor, tile_heavy, tile_list,
      tile_done, parts, bins, ckpt_base, ckpt, ckpt_ctr, heavy_keys, scan_status, counters,
      block_stats, tables, vrec, win, wide, outputs;
  // Control words of the persistent kernel (kCtlBytes, zero between
  // launches), then its stage stamps when profiling (kStampBytes).
  DevBuf ctl;
  // The batch tables (jobs, job tile/shade bases, draws, draw starts) sit in
  // one pinned staging buffer and one device buffer, uploaded by one copy.
  HostBuf h_tables, h_counters, h_outputs;
};

// Byte offsets of the tables inside Scratch::tables.
struct TableLayout {
  size_t jobs = 0, bases = 0, draws = 0, draw_begin = 0, vg_begin = 0, vg_draw = 0, hints = 0,
         bytes = 0;
};

// Jobs and draws of a batch, as the kernels see them.
struct BatchPlan {
  std::vector<JobDesc> jobs;
  std::vector<DrawDesc> draws;
  std::vector<int64_t> draw_begin;
  std::vector<int32_t> job_bases;  // tile bases of all jobs, then shade bases
  std::vector<int2> hints;          // job hints of the tiles, then of the shading blocks
  size_t tile_hints = 0;
  // Vertex groups: draws of one job with the same attribute arrays (one
  // mesh, or meshes sharing them) and modelview share their projected
  // vertices.
  std::vector<int64_t> vg_begin;
  std::vector<int32_t> vg_draw;
  int64_t total_verts = 0;
  int64_t total_tris = 0;
  int64_t ntiles = 0;
  int64_t npixels = 0;
  int64_t shade_blocks = 0;
  TableLayout layout;
};

// An execution slot: scratch, stream and events for one batch in flight.
struct Slot {
  Scratch sc;
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;
  hipEvent_t ev[kStages + 1] = {};  // stage boundaries (profiling)
  hipEvent_t done = nullptr;
  void* h_counters_dev = nullptr;  // device address of sc.h_counters ...
  void* h_counters_of = nullptr;   // ... for this host buffer
  uint32_t heavy_slots = 0;  // merge buffers of split tiles (grown on demand)
  size_t parts_want = 0;     // raster work items a batch needed (after an overflow)
  int32_t last_geom_grid = 0, last_bin_grid = 0;
  int32_t last_launches = 0;  // kernel launches of the last batch
  int64_t retries = 0;
  // Steady state: batches' uploads + stages + counter read-back captured as
  // instantiated graphs, keyed by the launch parameters, most recent first
  // (at most kGraphCache); a batch with a known key replays its graph.
  struct Graph {
    std::vector<unsigned char> key;   // launch parameters it was captured (or updated) with
    std::vector<unsigned char> topo;  // what fixes its nodes: kernels, template, copies
    hipGraphExec_t exec = nullptr;
  };
  std::vector<Graph> graphs;
  int64_t graph_replays = 0, graph_captures = 0;
  // Leading tile counters / look-back words known to be zero (the shading
  // kernel re-zeroes what a batch used), valid for the buffers in clean_ref.
  int64_t clean_tiles = 0, clean_scan = 0;
  const void* clean_ref[3] = {nullptr, nullptr, nullptr};
};

}  // namespace mr

struct mr_ctx {
  int32_t device = 0;
  uint64_t budget = 0;
  uint64_t resident_bytes = 0;
  uint64_t scratch_bytes = 0;
  std::vector<mr::MeshRec> meshes;
  std::vector<mr::TexRec> textures;
  mr::Slot main;  // synchronous calls (mr_render, mr_render_batch)
  int32_t tile_w = 16, tile_h = 16;
  int32_t raster_threads = 256;  // tile raster workgroup size
  float bin_factor = 3.0f;
  uint32_t frag_size = 1024;  // max (tile, triangle) pairs per raster item
  uint32_t band_work = 32768;  // walk work above which a tile is rastered in row bands
  bool profiling = false;
  bool graphs = false;  // capture steady-state batches as HIP graphs
  bool persistent = false;  // run batches as one persistent launch (16 x 16 tiles)
  uint64_t wait_us = 2000000;  // how long one of its waits may take
  // Graph use over all slots of the context (captured batches).
  int64_t graph_replays = 0, graph_updates = 0, graph_captures = 0;
  mr_stats stats{};
  int32_t num_cus = 256;
  std::recursive_mutex mu;
};

namespace mr {
int dev_alloc(mr_ctx* c, DevBuf& b, size_t bytes, bool scratch);
void dev_free(mr_ctx* c, DevBuf& b, bool scratch);
int host_alloc(HostBuf& b, size_t bytes);
void host_free(HostBuf& b);
int slot_init(Slot& s, hipStream_t external);
void slot_free(mr_ctx* c, Slot& s);
// Jobs and draws of n frames (validates handles and sizes).
int build_batch(mr_ctx* c, const mr_frame* frames, int nframes, BatchPlan& plan);
// Point the jobs' outputs at device memory: the caller's device pointers, or
// the slot's staging for host outputs (offsets per job and output kind).
int bind_outputs(mr_ctx* c, Slot& s, BatchPlan& p, const mr_outputs* outs,
                 std::vector<size_t>& staging_offs);
// Upload the batch tables and launch every stage on the slot's stream, then
// queue the counters' read-back and record s.done.  No host wait.
int enqueue_batch(mr_ctx* c, Slot& s, BatchPlan& p, bool profile);
// Wait for s.done, handle bin-list / merge-slot overflow by growing and
// re-running, and record statistics.
int finish_batch(mr_ctx* c, Slot& s, BatchPlan& p, bool profile);
// Copy host outputs back from the slot's staging (synchronous).
int copy_outputs(mr_ctx* c, Slot& s, const BatchPlan& p, const mr_outputs* outs,
                 const std::vector<size_t>& staging_offs);
int render_frames(mr_ctx* c, const mr_frame* frames, int n, const mr_outputs* outs);
}  // namespace mr
