// context.cpp -- C-ABI (include/minirender_amd.h) over the HIP pipeline.
//
// A context owns one device, the resident meshes and textures
// (Renderer::paintMesh's per-frame copies, Renderer.cpp:340-349, become one
// upload) and a main execution slot (scratch + stream) for synchronous calls.
// mr_render is Renderer::render() (Renderer.cpp:311-332) for one frame;
// mr_render_batch runs many frames through each kernel once.  The serving
// layer (server.cpp) drives further slots asynchronously.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "minirender_amd.h"
#include "common.h"
#include "exact.h"
#include "pipeline.h"
#include "context.h"

using namespace mr;

#define HIP_TRY(x)             \
  do {                         \
    hipError_t e_ = (x);       \
    if (e_ != hipSuccess) {    \
      (void)hipGetLastError(); \
      return MR_ERR_HIP;       \
    }                          \
  } while (0)

namespace mr {

int dev_alloc(mr_ctx* c, DevBuf& b, size_t bytes, bool scratch) {
  if (bytes == 0) bytes = 16;
  if (b.p && b.bytes >= bytes) return MR_OK;
  size_t want = b.p ? std::max(bytes, b.bytes + b.bytes / 2) : bytes;
  want = (want + 255) & ~(size_t)255;
  uint64_t& acct = scratch ? c->scratch_bytes : c->resident_bytes;
  if (scratch && c->budget && acct - b.bytes + want > c->budget) {
    if (acct - b.bytes + bytes > c->budget) return MR_ERR_BUDGET;
    want = bytes;
  }
  if (b.p) {
    hipFree(b.p);
    acct -= b.bytes;
    b.p = nullptr;
    b.bytes = 0;
  }
  if (hipMalloc(&b.p, want) != hipSuccess) {
    (void)hipGetLastError();
    b.p = nullptr;
    return MR_ERR_NOMEM;
  }
  b.bytes = want;
  acct += want;
  return MR_OK;
}

void dev_free(mr_ctx* c, DevBuf& b, bool scratch) {
  if (b.p) {
    hipFree(b.p);
    (scratch ? c->scratch_bytes : c->resident_bytes) -= b.bytes;
  }
  b.p = nullptr;
  b.bytes = 0;
}

int host_alloc(HostBuf& b, size_t bytes) {
  if (b.p && b.bytes >= bytes) return MR_OK;
  if (b.p) hipHostFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
  size_t want = std::max<size_t>(bytes, 4096);
  if (hipHostMalloc(&b.p, want, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    b.p = nullptr;
    return MR_ERR_NOMEM;
  }
  b.bytes = want;
  return MR_OK;
}

void host_free(HostBuf& b) {
  if (b.p) hipHostFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

int slot_init(Slot& s, hipStream_t external) {
  if (!external) {
    HIP_TRY(hipStreamCreateWithFlags(&s.own_stream, hipStreamNonBlocking));
    s.stream = s.own_stream;
  } else {
    s.stream = external;
  }
  for (auto& e : s.ev) HIP_TRY(hipEventCreate(&e));
  HIP_TRY(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
  return MR_OK;
}

void slot_free(mr_ctx* c, Slot& s) {
  if (s.stream) hipStreamSynchronize(s.stream);
  Scratch& x = s.sc;
  DevBuf* bufs[] = {&x.items, &x.region_count, &x.tile_count, &x.tile_off,
                    &x.tile_cursor, &x.tile_heavy, &x.tile_list, &x.tile_done, &x.parts, &x.bins,
                    &x.ckpt_base, &x.ckpt, &x.ckpt_ctr, &x.heavy_keys, &x.scan_status, &x.counters, &x.block_stats,
                    &x.tables, &x.vrec, &x.win, &x.wide, &x.outputs, &x.ctl};
  for (DevBuf* b : bufs) dev_free(c, *b, true);
  HostBuf* hbufs[] = {&x.h_tables, &x.h_counters, &x.h_outputs};
  for (HostBuf* b : hbufs) host_free(*b);
  for (auto& g : s.graphs)
    if (g.exec) hipGraphExecDestroy(g.exec);
  for (auto& e : s.ev)
    if (e) hipEventDestroy(e);
  if (s.done) hipEventDestroy(s.done);
  if (s.own_stream) hipStreamDestroy(s.own_stream);
  s = Slot{};
}

// Describe frames as jobs + draws; sizes the batch.
int build_batch(mr_ctx* c, const mr_frame* frames, int nframes, BatchPlan& plan) {
  plan.jobs.assign(nframes, JobDesc{});
  plan.draws.clear();
  plan.draw_begin.clear();
  plan.job_bases.assign(2 * (size_t)nframes, 0);
  plan.total_tris = 0;
  plan.vg_begin.clear();
  plan.vg_draw.clear();
  plan.total_verts = 0;
  plan.ntiles = 0;
  plan.npixels = 0;
  plan.shade_blocks = 0;
  const int tw = c->tile_w, th = c->tile_h;
  for (int j = 0; j < nframes; j++) {
    const mr_frame* f = &frames[j];
    if (f->width < 1 || f->height < 1 || f->width > kMaxDim ||
        f->height > kMaxDim || f->ndraws < 0 || (f->ndraws > 0 && !f->draws))
      return MR_ERR_ARG;
    JobDesc& J = plan.jobs[j];
    memcpy(J.P, f->projection, sizeof(J.P));
    J.znear = znear_of(f->projection);
    J.persp = f->projection[15] == 0.0f;
    J.W = f->width;
    J.H = f->height;
    J.tiles_x = (f->width + tw - 1) / tw;
    J.tiles_y = (f->height + th - 1) / th;
    J.tile_base = (int32_t)plan.ntiles;
    if ((int64_t)plan.ntiles + (int64_t)J.tiles_x * J.tiles_y > 0x7FFFFFFF)
      return MR_ERR_ARG;
    plan.ntiles += J.tiles_x * J.tiles_y;
    plan.job_bases[j] = J.tile_base;
    const int64_t npix = (int64_t)f->width * f->height;
    J.shade_base = (int32_t)plan.shade_blocks;
    plan.job_bases[nframes + j] = J.shade_base;
    plan.shade_blocks += (int64_t)((f->width + kShadeW - 1) / kShadeW) *
                         ((f->height + kShadeH - 1) / kShadeH);
    plan.npixels += npix;
    if (plan.shade_blocks > 0x7FFFFFFF) return MR_ERR_ARG;
    J.draw_begin = (int32_t)plan.draws.size();
    for (int i = 0; i < 3; i++) {
      J.bg[i] = f->background[i];
      J.light[i] = f->light[i];
    }
    J.light_is_point = f->light_is_point != 0;
    J.ambient = f->ambient;
    J.lighting = f->lighting != 0;
    J.save_normals = f->save_normals != 0;
    uint64_t order = 0;
    // (attribute owner, modelview) -> first vertex record
    std::map<std::string, int64_t> groups;
    for (int d = 0; d < f->ndraws; d++) {
      const mr_draw& s = f->draws[d];
      if (s.mesh < 0 || s.mesh >= (int)c->meshes.size() || !c->meshes[s.mesh].alive)
        return MR_ERR_HANDLE;
      if (s.texture >= 0 && (s.texture >= (int)c->textures.size() ||
                             !c->textures[s.texture].alive))
        return MR_ERR_HANDLE;
      const MeshRec& m = c->meshes[s.mesh];
      const int32_t vowner = m.owner >= 0 ? m.owner : s.mesh;
      const MeshRec& va = c->meshes[vowner];  // attribute arrays
      DrawDesc D;
      memset(&D, 0, sizeof(D));
      memcpy(D.mv, s.modelview, sizeof(D.mv));
      memcpy(D.nm, s.normalmat, sizeof(D.nm));
      for (int i = 0; i < 3; i++) {
        D.diffuse[i] = s.diffuse[i];
        D.specular[i] = s.specular[i];
        D.emissive[i] = s.emissive[i];
      }
      D.shininess = s.shininess;
      D.no_clip = (s.flags & MR_DRAW_NO_NEAR_CLIP) ? 1 : 0;
      D.pos = (const float*)va.pos.p;
      D.nrm = (const float*)va.nrm.p;
      D.idx = (const int32_t*)m.idx.p;
      D.nidx = (const int32_t*)(m.nidx_alias ? m.idx.p : m.nidx.p);
      if (m.has_uv) {
        D.uv = (const float*)va.uv.p;
        D.tidx = (const int32_t*)(m.tidx_alias ? m.idx.p : m.tidx.p);
      }
      if (f->texturing && s.texture >= 0) {
        const TexRec& t = c->textures[s.texture];
        if (t.rows > 0 && t.cols > 0) {
          D.tex = t.texels.p;
          D.tex_format = t.fmt;
          D.tex_rows = t.rows;
          D.tex_cols = t.cols;
        }
      }
      D.ntri = m.ntri;
      D.job = j;
      D.tile_base = J.tile_base;
      D.tiles_x = J.tiles_x;
      {
        std::string key((const char*)&vowner, sizeof(vowner));
        key.append((const char*)s.modelview, 12 * sizeof(float));
        auto it = groups.find(key);
        if (it != groups.end()) {
          D.vbase = it->second;
        } else {
          D.vbase = plan.total_verts;
          groups.emplace(key, D.vbase);
          plan.vg_begin.push_back(plan.total_verts);
          plan.vg_draw.push_back((int32_t)plan.draws.size());
          plan.total_verts += m.npos;
        }
      }
      D.tri_begin = plan.total_tris;
      D.order_base = (uint32_t)order;
      order += 2ull * (uint64_t)m.ntri;
      plan.total_tris += m.ntri;
      plan.draws.push_back(D);
      plan.draw_begin.push_back(D.tri_begin);
    }
    if (order >= 0xFFFFFFFFull) return MR_ERR_ARG;
    J.draw_end = (int32_t)plan.draws.size();
  }
  if (2 * plan.total_tris >= 0xFFFFFFFFll || plan.total_verts >= 0x7FFFFFFFll) return MR_ERR_ARG;
  // Job hints for the raster's and shading's job lookup: every kHintSpan-th
  // tile (block), its job and the next job's first tile (block).
  plan.hints.clear();
  for (int k = 0; k < 2; k++) {
    const int32_t* base = plan.job_bases.data() + (size_t)k * nframes;
    const int64_t n = k == 0 ? plan.ntiles : plan.shade_blocks;
    int j = 0;
    for (int64_t x = 0; x < n; x += kHintSpan) {
      while (j + 1 < nframes && base[j + 1] <= x) j++;
      plan.hints.push_back(make_int2(j, j + 1 < nframes ? base[j + 1] : 0x7FFFFFFF));
    }
    if (k == 0) plan.tile_hints = plan.hints.size();
  }
  auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
  TableLayout& L = plan.layout;
  const size_t nd = std::max<size_t>(plan.draws.size(), 1);
  L.jobs = 0;
  L.bases = up(plan.jobs.size() * sizeof(JobDesc));
  L.draws = L.bases + up(plan.job_bases.size() * 4);
  L.draw_begin = L.draws + up(nd * sizeof(DrawDesc));
  const size_t ng = std::max<size_t>(plan.vg_begin.size(), 1);
  L.vg_begin = L.draw_begin + up(nd * 8);
  L.vg_draw = L.vg_begin + up(ng * 8);
  L.hints = L.vg_draw + up(ng * 4);
  L.bytes = L.hints + up(std::max<size_t>(plan.hints.size(), 1) * 8);
  return MR_OK;
}

// Edge checkpoints of a batch: two per pixel covers every frame here with
// room; a batch that needs more leaves the extra rows to shading's own walk.
static size_t ckpt_entries(int64_t npixels) { return (size_t)npixels * 2 + 65536; }

// Device scratch of a batch, buffer by buffer (bytes), in the order of
// scratch_bufs().
constexpr int kScratchBufs = 21;

static void scratch_sizes(mr_ctx* c, const BatchPlan& p, uint64_t bins_cap, uint32_t heavy_slots,
                          size_t parts_want, size_t out[kScratchBufs]) {
  const int64_t T = p.total_tris;
  const int scb = scan_blocks_for((int32_t)p.ntiles);
  const size_t nt = (size_t)p.ntiles * 4;
  const int64_t nregions = std::max<int64_t>((T + 63) / 64, 1);
  const size_t npt = (size_t)c->tile_w * c->tile_h * 8;
  // entry parts of split tiles, and row bands of heavy tiles (one per tile
  // on average; a batch that wants more re-runs with the measured need)
  const size_t parts = std::max<size_t>((size_t)(bins_cap / c->frag_size) + heavy_slots + 16 +
                                            (size_t)p.ntiles,
                                        parts_want);
  const size_t v[kScratchBufs] = {
      (size_t)nregions * kRegion * sizeof(BinItem),  // items
      (size_t)nregions * 4,                          // region_count
      nt * kCtrStride,                               // tile_count
      nt,                                            // tile_off
      nt * kCtrStride,                               // tile_cursor
      nt,                                            // tile_heavy
      nt * 4,                                        // tile_list
      nt,                                            // tile_done
      parts * 8,                                     // parts
      (size_t)bins_cap * 8,                          // bins
      (size_t)std::max<int64_t>(2 * T, 1) * 8,       // ckpt_base
      ckpt_entries(p.npixels) * 8,                   // ckpt
      (size_t)kCkptSegs * kCtrStride * 8,            // ckpt_ctr
      (size_t)heavy_slots * npt,                     // heavy_keys
      (size_t)std::max(scb, 1) * 8,                  // scan_status
      (size_t)kNumCounters * 8,                      // counters
      (size_t)2 * kStatSlots * 8,                    // block_stats
      p.layout.bytes,                                // tables
      (size_t)p.total_verts * kVrecFloats * 4,       // vrec
      (size_t)p.npixels * 4,                         // win
      (ckpt_entries(p.npixels) / 2 + 64) * 8,        // row tasks of the row pre-pass
  };
  for (int k = 0; k < kScratchBufs; k++) out[k] = v[k];
}

static void scratch_bufs(Scratch& x, DevBuf* b[kScratchBufs]) {
  DevBuf* v[kScratchBufs] = {&x.items, &x.region_count, &x.tile_count, &x.tile_off,
                             &x.tile_cursor, &x.tile_heavy, &x.tile_list, &x.tile_done,
                             &x.parts, &x.bins, &x.ckpt_base, &x.ckpt, &x.ckpt_ctr,
                             &x.heavy_keys, &x.scan_status, &x.counters, &x.block_stats,
                             &x.tables, &x.vrec, &x.win, &x.wide};
  for (int k = 0; k < kScratchBufs; k++) b[k] = v[k];
}

static int ensure_scratch(mr_ctx* c, Slot& sl, const BatchPlan& p, uint64_t bins_cap,
                          uint32_t heavy_slots) {
  Scratch& s = sl.sc;
  size_t bytes[kScratchBufs];
  DevBuf* bufs[kScratchBufs];
  scratch_sizes(c, p, bins_cap, heavy_slots, sl.parts_want, bytes);
  scratch_bufs(s, bufs);
  int rc;
  for (int k = 0; k < kScratchBufs; k++)
    if ((rc = dev_alloc(c, *bufs[k], bytes[k], true))) return rc;
  if ((rc = host_alloc(s.h_tables, p.layout.bytes))) return rc;
  if ((rc = host_alloc(s.h_counters, kNumCounters * 8))) return rc;
  if (c->persistent && &sl == &c->main && !s.ctl.p) {
    if ((rc = dev_alloc(c, s.ctl, kCtlBytes + kStampBytes, true))) return rc;
    HIP_TRY(hipMemsetAsync(s.ctl.p, 0, kCtlBytes + kStampBytes, sl.stream));
  }
  return MR_OK;
}

static BatchArgs make_args(mr_ctx* c, Slot& sl, const BatchPlan& p) {
  Scratch& s = sl.sc;
  BatchArgs a;
  memset(&a, 0, sizeof(a));
  a.items = (BinItem*)s.items.p;
  a.region_count = (uint32_t*)s.region_count.p;
  a.nregions = (p.total_tris + 63) / 64;
  a.ckpt_in_scan = a.nregions <= kCkptScanRegions ? 1 : 0;
  a.tile_count = (unsigned long long*)s.tile_count.p;
  a.tile_off = (uint32_t*)s.tile_off.p;
  a.tile_cursor = (uint32_t*)s.tile_cursor.p;
  a.tile_heavy = (uint32_t*)s.tile_heavy.p;
  a.tile_list = (uint4*)s.tile_list.p;
  a.tile_done = (uint32_t*)s.tile_done.p;
  a.parts = (uint2*)s.parts.p;
  // Raster work items of split tiles and row bands this batch may use: its
  // own bound, not the buffer's capacity (which only grows, e.g. over an
  // autotune's candidates): the raster grid is ntiles + parts_cap, and a
  // batch that wants more re-runs with the measured need.
  const uint64_t batch_bins = (uint64_t)((double)p.total_tris * c->bin_factor) + (uint64_t)p.ntiles;
  const uint64_t batch_parts = std::max<uint64_t>(
      batch_bins / c->frag_size + (uint64_t)sl.heavy_slots + 16 + 2 * (uint64_t)p.ntiles,
      (uint64_t)sl.parts_want);
  a.parts_cap = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(s.parts.bytes / 8, batch_parts),
                                             0xFFFFFFF0u);
  a.bins = (uint2*)s.bins.p;
  a.bins_cap = s.bins.bytes / 8;
  if (a.bins_cap > 0xFFFFFFF0ull) a.bins_cap = 0xFFFFFFF0ull;
  a.ckpt_base = (int2*)s.ckpt_base.p;
  a.ckpt = (float2*)s.ckpt.p;
  // Checkpoint capacity, held to twice the row-task entries: the row tasks
  // of every item given checkpoints then fit in their segment.
  a.ckpt_cap = std::min<uint64_t>(std::min<uint64_t>(s.ckpt.bytes / 8, 2 * (s.wide.bytes / 8)),
                                  0xFFFFFFF0ull);
  a.ckpt_ctr = (unsigned long long*)s.ckpt_ctr.p;
  a.ckpt_rows = (unsigned long long*)s.wide.p;
  a.heavy_keys = (unsigned long long*)s.heavy_keys.p;
  a.heavy_cap = (uint32_t)(s.heavy_keys.bytes / ((size_t)c->tile_w * c->tile_h * 8));
  a.frag_size = c->frag_size;
  a.band_work = c->band_work;
  a.scan_status = (unsigned long long*)s.scan_status.p;
  a.counters = (unsigned long long*)s.counters.p;
  a.block_stats = (unsigned long long*)s.block_stats.p;
  char* tb = (char*)s.tables.p;
  a.vscr = (float4*)s.vrec.p;
  a.total_verts = p.total_verts;
  a.vg_begin = (const int64_t*)(tb + p.layout.vg_begin);
  a.vg_draw = (const int32_t*)(tb + p.layout.vg_draw);
  a.nvgroups = (int32_t)p.vg_begin.size();
  a.draws = (const DrawDesc*)(tb + p.layout.draws);
  a.draw_begin = (const int64_t*)(tb + p.layout.draw_begin);
  a.ndraws = (int32_t)p.draws.size();
  a.jobs = (const JobDesc*)(tb + p.layout.jobs);
  a.job_tile_base = (const int32_t*)(tb + p.layout.bases);
  a.job_shade_base = (const int32_t*)(tb + p.layout.bases) + p.jobs.size();
  a.tile_hint = (const int2*)(tb + p.layout.hints);
  a.shade_hint = a.tile_hint + p.tile_hints;
  a.njobs = (int32_t)p.jobs.size();
  a.total_tris = p.total_tris;
  a.ntiles = (int32_t)p.ntiles;
  a.tile_w = c->tile_w;
  a.tile_h = c->tile_h;
  a.raster_threads = c->raster_threads;
  a.scan_blocks = scan_blocks_for((int32_t)p.ntiles);
  a.shade_blocks = (int32_t)p.shade_blocks;
  a.num_cus = c->num_cus;
  a.geom_blocks = std::min(std::max(geom_blocks_per_cu(), 1) * std::max(c->num_cus, 1),
                           kStatSlots);
  return a;
}

// persistent: every stage in one launch of the persistent kernel (profiling
// then times the whole launch as stage 0).
static int launch_stages(const BatchArgs& a, Slot& s, bool profile, bool clear,
                         bool persistent, uint64_t wait_us) {
  hipStream_t st = s.stream;
  if (persistent) {
    unsigned long long* ctl = (unsigned long long*)s.sc.ctl.p;
    unsigned long long* stamps = profile ? ctl + kCtlBytes / 8 : nullptr;
    if (profile) HIP_TRY(hipMemsetAsync(stamps, 0, kStampBytes, st));
    if (profile) HIP_TRY(hipEventRecord(s.ev[0], st));
    HIP_TRY(launch_persistent(a, clear, ctl, stamps, wait_us, st));
    if (profile)
      for (int i = 1; i <= kStages; i++) HIP_TRY(hipEventRecord(s.ev[i], st));
    s.last_launches = 1;
    return MR_OK;
  }
  s.last_launches = (clear ? 1 : 0) + (a.total_verts ? 1 : 0) + (a.total_tris ? 1 : 0) +
                    (a.ckpt_in_scan ? 4 : 5);
  if (profile) HIP_TRY(hipEventRecord(s.ev[0], st));
  if (clear) HIP_TRY(launch_clear(a, st));
  if (profile) HIP_TRY(hipEventRecord(s.ev[1], st));
  HIP_TRY(launch_vertex(a, st));
  HIP_TRY(launch_geometry(a, st));
  if (profile) HIP_TRY(hipEventRecord(s.ev[2], st));
  HIP_TRY(launch_scan(a, st));
  if (profile) HIP_TRY(hipEventRecord(s.ev[3], st));
  HIP_TRY(launch_bin(a, st));
  if (profile) HIP_TRY(hipEventRecord(s.ev[4], st));
  HIP_TRY(launch_rows(a, st));
  if (profile) HIP_TRY(hipEventRecord(s.ev[5], st));
  HIP_TRY(launch_tile(a, st));
  if (profile) HIP_TRY(hipEventRecord(s.ev[6], st));
  HIP_TRY(launch_shade(a, st));
  if (profile) HIP_TRY(hipEventRecord(s.ev[7], st));
  return MR_OK;
}

static uint64_t bins_wanted(mr_ctx* c, const Slot& s, const BatchPlan& p) {
  uint64_t want = std::max<uint64_t>(
      (uint64_t)((double)p.total_tris * c->bin_factor) + (uint64_t)p.ntiles, 1u << 16);
  return std::max<uint64_t>(want, s.sc.bins.bytes / 8);
}

int enqueue_batch(mr_ctx* c, Slot& s, BatchPlan& p, bool profile) {
  const uint64_t want_bins = bins_wanted(c, s, p);
  const uint32_t want_heavy = std::max<uint32_t>(
      s.heavy_slots, (uint32_t)std::min<uint64_t>(want_bins / c->frag_size + 16, 1u << 20));
  int rc = ensure_scratch(c, s, p, want_bins, want_heavy);
  if (rc) return rc;
  Scratch& x = s.sc;
  uint32_t* win = (uint32_t*)x.win.p;
  for (auto& J : p.jobs) {
    J.win = win;
    win += (size_t)J.W * J.H;
  }
  // Host tables go to pinned staging first; the uploads read them from there
  // (also on a graph replay, which re-reads the same staging).
  char* ht = (char*)x.h_tables.p;
  memcpy(ht + p.layout.jobs, p.jobs.data(), p.jobs.size() * sizeof(JobDesc));
  memcpy(ht + p.layout.bases, p.job_bases.data(), p.job_bases.size() * 4);
  if (!p.draws.empty()) {
    memcpy(ht + p.layout.draws, p.draws.data(), p.draws.size() * sizeof(DrawDesc));
    memcpy(ht + p.layout.draw_begin, p.draw_begin.data(), p.draws.size() * 8);
    memcpy(ht + p.layout.vg_begin, p.vg_begin.data(), p.vg_begin.size() * 8);
    memcpy(ht + p.layout.vg_draw, p.vg_draw.data(), p.vg_draw.size() * 4);
  }
  memcpy(ht + p.layout.hints, p.hints.data(), p.hints.size() * 8);
  // The persistent kernel runs for the synchronous calls only: two of them
  // in flight at once (the service's slots) would compete for residency.
  const bool persistent = c->persistent && &s == &c->main && x.ctl.p &&
                          persistent_supported(c->tile_w, c->tile_h);
  BatchArgs a = make_args(c, s, p);
  // The shading kernel writes the counters into the pinned read-back buffer
  // (the persistent kernel sets its error word after shading: it keeps the
  // copy).
  if (!persistent) {
    if (s.h_counters_of != x.h_counters.p) {
      void* d = nullptr;
      s.h_counters_dev = nullptr;
      if (hipHostGetDevicePointer(&d, x.h_counters.p, 0) == hipSuccess) s.h_counters_dev = d;
      else (void)hipGetLastError();
      s.h_counters_of = x.h_counters.p;
    }
    a.host_counters = (unsigned long long*)s.h_counters_dev;
  }
  s.last_geom_grid = p.total_tris ? geom_grid(a) : 0;
  s.last_bin_grid = bin_grid(a);
  // The clear kernel runs only when the counters it zeroes are not known to
  // be zero: new buffers, more tiles than cleared before, or no geometry
  // kernel (which clears the batch counters).
  const bool same_bufs = s.clean_ref[0] == x.tile_count.p && s.clean_ref[1] == x.tile_done.p &&
                         s.clean_ref[2] == x.scan_status.p;
  const bool clear = !same_bufs || p.total_tris == 0 || p.ntiles > s.clean_tiles ||
                     a.scan_blocks > s.clean_scan;
  // Zero prefixes after this batch: the clear kernel (if run) zeroes this
  // batch's range, the shading kernel re-zeroes it, entries past it keep
  // their state.
  const int64_t next_tiles = std::max<int64_t>(p.ntiles, same_bufs ? s.clean_tiles : 0);
  const int64_t next_scan = std::max<int64_t>(a.scan_blocks, same_bufs ? s.clean_scan : 0);
  s.clean_tiles = s.clean_scan = 0;  // until the launches below succeed
  auto issue = [&]() -> int {
    HIP_TRY(hipMemcpyAsync(x.tables.p, x.h_tables.p, p.layout.bytes, hipMemcpyHostToDevice,
                           s.stream));
    int r = launch_stages(a, s, profile, clear, persistent, c->wait_us);
    if (r) return r;
    if (!a.host_counters)
      HIP_TRY(hipMemcpyAsync(x.h_counters.p, a.counters, kNumCounters * 8,
                             hipMemcpyDeviceToHost, s.stream));
    return MR_OK;
  };
  // Graphs only on the slot's own stream (a caller's stream may be the null
  // stream, which cannot be captured) and without per-stage events.
  if (c->graphs && !profile && s.own_stream && s.stream == s.own_stream) {
    const void* hp[5] = {x.h_tables.p, x.h_counters.p, (const void*)p.layout.bytes,
                         (const void*)(uintptr_t)clear, persistent ? x.ctl.p : nullptr};
    std::vector<unsigned char> key(sizeof(BatchArgs) + sizeof(hp));
    memcpy(key.data(), &a, sizeof(BatchArgs));
    memcpy(key.data() + sizeof(BatchArgs), hp, sizeof(hp));
    size_t hit = s.graphs.size();
    for (size_t i = 0; i < s.graphs.size(); i++)
      if (s.graphs[i].key == key) {
        hit = i;
        break;
      }
    // A batch of other sizes with the same kernels (same clear, vertex and
    // geometry launches, raster template, persistent or not) updates the
    // parameters of an instantiated graph in place instead of instantiating
    // a new one, so a service of mixed batches keeps replaying a few graphs.
    const int32_t topo_v[7] = {clear ? 1 : 0, a.total_verts ? 1 : 0, a.total_tris ? 1 : 0,
                               a.tile_w, a.raster_threads, persistent ? 1 : 0,
                               persistent ? 0 : a.ckpt_in_scan};
    std::vector<unsigned char> topo((const unsigned char*)topo_v,
                                    (const unsigned char*)topo_v + sizeof(topo_v));
    if (hit == s.graphs.size()) {
      HIP_TRY(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
      const int r = issue();
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(s.stream, &g);
      if (r) {
        if (g) hipGraphDestroy(g);
        return r;
      }
      if (e != hipSuccess) return MR_ERR_HIP;
      size_t upd = s.graphs.size();
      for (size_t i = 0; i < s.graphs.size(); i++) {
        if (s.graphs[i].topo != topo) continue;
        hipGraphNode_t err_node = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        if (hipGraphExecUpdate(s.graphs[i].exec, g, &err_node, &res) == hipSuccess &&
            res == hipGraphExecUpdateSuccess) {
          upd = i;
          break;
        }
        (void)hipGetLastError();
      }
      if (upd < s.graphs.size()) {
        hipGraphDestroy(g);
        s.graphs[upd].key = std::move(key);
        if (upd) std::rotate(s.graphs.begin(), s.graphs.begin() + upd, s.graphs.begin() + upd + 1);
        c->graph_updates++;
      } else {
        Slot::Graph ng;
        const hipError_t ei = hipGraphInstantiate(&ng.exec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (ei != hipSuccess) return MR_ERR_HIP;
        ng.key = std::move(key);
        ng.topo = std::move(topo);
        if (s.graphs.size() >= kGraphCache) {
          hipGraphExecDestroy(s.graphs.back().exec);
          s.graphs.pop_back();
        }
        s.graphs.insert(s.graphs.begin(), std::move(ng));
        s.graph_captures++;
        c->graph_captures++;
      }
    } else {
      if (hit) std::rotate(s.graphs.begin(), s.graphs.begin() + hit, s.graphs.begin() + hit + 1);
      s.graph_replays++;
      c->graph_replays++;
    }
    HIP_TRY(hipGraphLaunch(s.graphs.front().exec, s.stream));
  } else {
    if ((rc = issue())) return rc;
  }
  HIP_TRY(hipEventRecord(s.done, s.stream));
  s.clean_ref[0] = x.tile_count.p;
  s.clean_ref[1] = x.tile_done.p;
  s.clean_ref[2] = x.scan_status.p;
  s.clean_tiles = next_tiles;
  s.clean_scan = next_scan;
  return MR_OK;
}

int finish_batch(mr_ctx* c, Slot& s, BatchPlan& p, bool profile) {
  Scratch& x = s.sc;
  const unsigned long long* hc = (const unsigned long long*)x.h_counters.p;
  s.retries = 0;
  for (int attempt = 0;; attempt++) {
    HIP_TRY(hipEventSynchronize(s.done));
    if (hc[kError]) {
      // A bounded wait gave up: stages after it did not run, so the tile
      // counters and look-back words shading re-zeroes may hold this batch's
      // values; the next batch clears them first.  The persistent kernel's
      // last workgroup clears its control words in any case; cleared again
      // to be sure.
      s.clean_tiles = s.clean_scan = 0;
      if (x.ctl.p) HIP_TRY(hipMemset(x.ctl.p, 0, kCtlBytes));
      return MR_ERR_DEVICE;
    }
    if (!hc[kBinOverflow] && !hc[kHeavyOverflow]) break;
    if (attempt >= 3 || hc[kBinEntries] >= 0xFFFFFFF0ull) return MR_ERR_NOMEM;
    s.retries++;
    int rc;
    if (hc[kBinOverflow]) {
      const uint64_t need = hc[kBinEntries] + hc[kBinEntries] / 4 + 1024;
      if ((rc = dev_alloc(c, x.bins, (size_t)need * 8, true))) return rc;
    }
    if (hc[kHeavyOverflow]) {
      s.parts_want = std::max<size_t>(s.parts_want, (size_t)(hc[kParts] + hc[kParts] / 4 + 16));
      s.heavy_slots = (uint32_t)(hc[kHeavy] + hc[kHeavy] / 4 + 16);
      if ((rc = dev_alloc(c, x.heavy_keys,
                          (size_t)s.heavy_slots * c->tile_w * c->tile_h * 8, true)))
        return rc;
    }
    if ((rc = enqueue_batch(c, s, p, profile))) return rc;
  }
  s.heavy_slots = std::max<uint32_t>(s.heavy_slots, (uint32_t)hc[kHeavy]);
  if (&s == &c->main) {
    mr_stats& st = c->stats;
    st.triangles_in = p.total_tris;
    st.triangles_clipped = (int64_t)hc[kClipped];
    st.records = (int64_t)hc[kItems];
    st.bin_entries = (int64_t)hc[kBinEntries];
    st.tiles = p.ntiles;
    st.retries = s.retries;
    st.launches = s.last_launches;
    st.split_tiles = (int64_t)hc[kHeavy];
    st.raster_parts = (int64_t)hc[kParts];
    if (profile && s.last_launches == 1 && x.ctl.p) {
      // One persistent launch: each stage's span from its workgroups' stamps
      // (first start to last end; the stages of the vertex and geometry
      // kernels are reported together, as the separate launches are).
      unsigned long long w[2 * kPStages * kStampStride];
      HIP_TRY(hipMemcpy(w, (const char*)x.ctl.p + kCtlBytes, kStampBytes, hipMemcpyDeviceToHost));
      const double hz = (double)wall_clock_hz();
      const int slot[kPStages] = {0, 1, 1, 2, 3, 4, 5, 6};
      unsigned long long b[kStages], e[kStages];
      for (int i = 0; i < kStages; i++) b[i] = ~0ull, e[i] = 0;
      for (int k = 0; k < kPStages; k++) {
        const unsigned long long nb = w[(2 * k) * kStampStride], ne = w[(2 * k + 1) * kStampStride];
        if (!nb || !ne) continue;  // the stage had no items
        b[slot[k]] = std::min(b[slot[k]], ~nb);
        e[slot[k]] = std::max(e[slot[k]], ne);
      }
      for (int i = 0; i < kStages; i++)
        st.stage_ms[i] = e[i] > b[i] ? (double)(e[i] - b[i]) / hz * 1e3 : 0.0;
      float ms = 0;
      hipEventElapsedTime(&ms, s.ev[0], s.ev[kStages]);
      st.frame_ms = ms;
    } else if (profile) {
      float ms = 0;
      for (int i = 0; i < kStages; i++) {
        hipEventElapsedTime(&ms, s.ev[i], s.ev[i + 1]);
        st.stage_ms[i] = ms;
      }
      hipEventElapsedTime(&ms, s.ev[0], s.ev[kStages]);
      st.frame_ms = ms;
    }
  }
  return MR_OK;
}

static const size_t kOutBytes[kOutKinds] = {3, 4, 4, 12, 12, 12};  // rgb8 depth winner rgb normals points

int bind_outputs(mr_ctx* c, Slot& s, BatchPlan& p, const mr_outputs* outs,
                 std::vector<size_t>& offs) {
  const int n = (int)p.jobs.size();
  size_t total = 0;
  offs.assign((size_t)n * kOutKinds, 0);
  for (int j = 0; j < n; j++) {
    const mr_outputs& o = outs[j];
    const size_t npix = (size_t)p.jobs[j].W * p.jobs[j].H;
    const void* ptrs[kOutKinds] = {o.rgb8, o.depth, o.winner, o.rgb, o.normals, o.points};
    for (int k = 0; k < kOutKinds; k++) {
      offs[(size_t)j * kOutKinds + k] = total;
      if (ptrs[k] && !o.on_device) total += (npix * kOutBytes[k] + 255) & ~(size_t)255;
    }
  }
  int rc;
  if (total && (rc = dev_alloc(c, s.sc.outputs, total, true))) return rc;
  char* base = (char*)s.sc.outputs.p;
  for (int j = 0; j < n; j++) {
    const mr_outputs& o = outs[j];
    JobDesc& J = p.jobs[j];
    if (o.on_device) {
      J.rgb8 = o.rgb8; J.depth = o.depth; J.winner = o.winner; J.rgb = o.rgb; J.normals = o.normals;
      J.points = o.points;
    } else {
      const size_t* of = &offs[(size_t)j * kOutKinds];
      J.rgb8 = o.rgb8 ? (uint8_t*)(base + of[0]) : nullptr;
      J.depth = o.depth ? (float*)(base + of[1]) : nullptr;
      J.winner = o.winner ? (uint32_t*)(base + of[2]) : nullptr;
      J.rgb = o.rgb ? (float*)(base + of[3]) : nullptr;
      J.normals = o.normals ? (float*)(base + of[4]) : nullptr;
      J.points = o.points ? (float*)(base + of[5]) : nullptr;
    }
  }
  return MR_OK;
}

int copy_outputs(mr_ctx* c, Slot& s, const BatchPlan& p, const mr_outputs* outs,
                 const std::vector<size_t>& offs) {
  (void)c;
  bool copied = false;
  for (size_t j = 0; j < p.jobs.size(); j++) {
    const mr_outputs& o = outs[j];
    if (o.on_device) continue;
    const size_t npix = (size_t)p.jobs[j].W * p.jobs[j].H;
    void* dst[kOutKinds] = {o.rgb8, o.depth, o.winner, o.rgb, o.normals, o.points};
    for (int k = 0; k < kOutKinds; k++) {
      if (!dst[k]) continue;
      HIP_TRY(hipMemcpyAsync(dst[k], (char*)s.sc.outputs.p + offs[j * kOutKinds + k], npix * kOutBytes[k],
                             hipMemcpyDeviceToHost, s.stream));
      copied = true;
    }
  }
  if (copied) HIP_TRY(hipStreamSynchronize(s.stream));
  return MR_OK;
}

// ---------------------------------------------------------------------------
// mr_render without allocating: every scratch buffer is carved from the
// static device array of kernels.hip (kNoAllocScratchBytes per device), the
// batch tables and counters pass through fixed host arrays per device, and
// outputs must be device buffers.  Stale contents of the array from earlier
// calls do not matter: the clear kernel and the geometry kernel zero what
// must start at zero (counters, look-back words, tile counters, checkpoint
// segment counters), and every other buffer is written before it is read.
constexpr size_t kNoAllocTableBytes = (size_t)1 << 20;
static char g_na_tables[kMaxDevices][kNoAllocTableBytes];
static unsigned long long g_na_counters[kMaxDevices][kNumCounters];

static int render_noalloc(mr_ctx* c, const mr_frame* frame, const mr_outputs* out) {
  if (!out->on_device) return MR_ERR_ARG;
  if (c->device < 0 || c->device >= kMaxDevices) return MR_ERR_ARG;
  char* base = (char*)noalloc_scratch(c->device);
  if (!base) return MR_ERR_HIP;
  BatchPlan p;
  int rc;
  if ((rc = build_batch(c, frame, 1, p))) return rc;
  if (p.layout.bytes > kNoAllocTableBytes) return MR_ERR_LIMIT;
  JobDesc& J = p.jobs[0];
  J.rgb8 = out->rgb8; J.depth = out->depth; J.winner = out->winner; J.rgb = out->rgb;
  J.normals = out->normals; J.points = out->points;
  // Bin list and merge slots sized as the allocating path sizes them first;
  // they cannot grow here, so an overflow is an error.
  const uint64_t bins_cap = std::max<uint64_t>(
      (uint64_t)((double)p.total_tris * c->bin_factor) + (uint64_t)p.ntiles, 1u << 16);
  const uint32_t heavy = (uint32_t)std::min<uint64_t>(bins_cap / c->frag_size + 16, 1u << 20);
  size_t bytes[kScratchBufs];
  DevBuf* bufs[kScratchBufs];
  Slot ns;
  scratch_sizes(c, p, bins_cap, heavy, 0, bytes);
  scratch_bufs(ns.sc, bufs);
  size_t off = 0;
  for (int k = 0; k < kScratchBufs; k++) {
    const size_t b = (bytes[k] + 255) & ~(size_t)255;
    if (b > kNoAllocScratchBytes - off) return MR_ERR_LIMIT;
    bufs[k]->p = base + off;
    bufs[k]->bytes = bytes[k];
    off += b;
  }
  J.win = (uint32_t*)ns.sc.win.p;
  ns.stream = c->main.stream;
  char* ht = g_na_tables[c->device];
  memcpy(ht + p.layout.jobs, p.jobs.data(), p.jobs.size() * sizeof(JobDesc));
  memcpy(ht + p.layout.bases, p.job_bases.data(), p.job_bases.size() * 4);
  if (!p.draws.empty()) {
    memcpy(ht + p.layout.draws, p.draws.data(), p.draws.size() * sizeof(DrawDesc));
    memcpy(ht + p.layout.draw_begin, p.draw_begin.data(), p.draws.size() * 8);
    memcpy(ht + p.layout.vg_begin, p.vg_begin.data(), p.vg_begin.size() * 8);
    memcpy(ht + p.layout.vg_draw, p.vg_draw.data(), p.vg_draw.size() * 4);
  }
  memcpy(ht + p.layout.hints, p.hints.data(), p.hints.size() * 8);
  const BatchArgs a = make_args(c, ns, p);
  unsigned long long* hc = g_na_counters[c->device];
  HIP_TRY(hipMemcpyAsync(ns.sc.tables.p, ht, p.layout.bytes, hipMemcpyHostToDevice, ns.stream));
  if ((rc = launch_stages(a, ns, false, true, false, 0))) return rc;
  HIP_TRY(hipMemcpyAsync(hc, a.counters, kNumCounters * 8, hipMemcpyDeviceToHost, ns.stream));
  HIP_TRY(hipStreamSynchronize(ns.stream));
  if (hc[kError]) return MR_ERR_DEVICE;
  if (hc[kBinOverflow] || hc[kHeavyOverflow]) return MR_ERR_NOMEM;
  mr_stats& st = c->stats;
  st.triangles_in = p.total_tris;
  st.triangles_clipped = (int64_t)hc[kClipped];
  st.records = (int64_t)hc[kItems];
  st.bin_entries = (int64_t)hc[kBinEntries];
  st.tiles = p.ntiles;
  st.retries = 0;
  return MR_OK;
}

int render_frames(mr_ctx* c, const mr_frame* frames, int n, const mr_outputs* outs) {
  BatchPlan p;
  std::vector<size_t> offs;
  int rc;
  if ((rc = build_batch(c, frames, n, p))) return rc;
  if ((rc = bind_outputs(c, c->main, p, outs, offs))) return rc;
  if ((rc = enqueue_batch(c, c->main, p, c->profiling))) return rc;
  if ((rc = finish_batch(c, c->main, p, c->profiling))) return rc;
  return copy_outputs(c, c->main, p, outs, offs);
}

}  // namespace mr

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

int mr_device_count(int32_t* count) {
  if (!count) return MR_ERR_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    n = 0;
  }
  *count = n;
  return MR_OK;
}

int mr_ctx_create(int32_t device, uint64_t hbm_budget, mr_ctx** out) {
  if (!out) return MR_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return MR_ERR_NOGPU;
  }
  if (device < 0 || device >= n) return MR_ERR_ARG;
  HIP_TRY(hipSetDevice(device));
  mr_ctx* c = new (std::nothrow) mr_ctx();
  if (!c) return MR_ERR_NOMEM;
  c->device = device;
  c->budget = hbm_budget;
  if (slot_init(c->main, nullptr) != MR_OK) {
    slot_free(c, c->main);
    delete c;
    return MR_ERR_HIP;
  }
  // Address of this device's copy of the no-alloc scratch (once per device;
  // on failure mr_render_noalloc reports MR_ERR_HIP).
  (void)noalloc_scratch_init(device);
  int cus = 0;
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  c->num_cus = cus > 0 ? cus : 256;
  *out = c;
  return MR_OK;
}

int mr_ctx_destroy(mr_ctx* c) {
  if (!c) return MR_ERR_ARG;
  hipSetDevice(c->device);
  hipDeviceSynchronize();
  for (auto& m : c->meshes) {
    dev_free(c, m.pos, false); dev_free(c, m.nrm, false); dev_free(c, m.uv, false);
    dev_free(c, m.idx, false); dev_free(c, m.nidx, false); dev_free(c, m.tidx, false);
  }
  for (auto& t : c->textures) dev_free(c, t.texels, false);
  slot_free(c, c->main);
  delete c;
  return MR_OK;
}

int mr_ctx_set_stream(mr_ctx* c, void* stream) {
  if (!c) return MR_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  c->main.stream = stream ? (hipStream_t)stream : c->main.own_stream;
  return MR_OK;
}

int mr_ctx_get_stream(mr_ctx* c, void** stream) {
  if (!c || !stream) return MR_ERR_ARG;
  *stream = (void*)c->main.stream;
  return MR_OK;
}

int mr_ctx_set_tile(mr_ctx* c, int32_t tw, int32_t th) {
  if (!c || !tile_supported(tw, th)) return MR_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  if (c->tile_w != tw || c->tile_h != th) {
    // merge buffers are sized per tile: drop them
    hipStreamSynchronize(c->main.stream);
    dev_free(c, c->main.sc.heavy_keys, true);
    c->main.heavy_slots = 0;
  }
  c->tile_w = tw;
  c->tile_h = th;
  c->raster_threads = default_raster_threads(tw);
  return MR_OK;
}

int mr_ctx_set_frag_size(mr_ctx* c, int32_t pairs) {
  if (!c || pairs < 64 || pairs > (1 << 20)) return MR_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  c->frag_size = (uint32_t)pairs;
  return MR_OK;
}

int mr_ctx_set_raster_threads(mr_ctx* c, int32_t threads) {
  if (!c) return MR_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  if (threads == 0) threads = default_raster_threads(c->tile_w);
  if (!raster_threads_supported(c->tile_w, threads)) return MR_ERR_ARG;
  c->raster_threads = threads;
  return MR_OK;
}

int mr_ctx_set_band_work(mr_ctx* c, int32_t units) {
  if (!c || units < 0) return MR_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  c->band_work = (uint32_t)units;
  return MR_OK;
}

int mr_ctx_set_bin_factor(mr_ctx* c, float f) {
  if (!c || !(f >= 1.0f) || f > 1024.0f) return MR_ERR_ARG;
  c->bin_factor = f;
  return MR_OK;
}

int mr_ctx_set_profiling(mr_ctx* c, int32_t on) {
  if (!c) return MR_ERR_ARG;
  c->profiling = on != 0;
  return MR_OK;
}

int mr_ctx_set_wait_limit(mr_ctx* c, int64_t us) {
  if (!c) return MR_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  c->wait_us = us < 0 ? 2000000ull : (uint64_t)us;
  return MR_OK;
}

int mr_ctx_set_persistent(mr_ctx* c, int32_t on) {
  if (!c) return MR_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  if (on && !persistent_supported(c->tile_w, c->tile_h)) return MR_ERR_ARG;
  c->persistent = on != 0;
  return MR_OK;
}

int mr_ctx_set_graphs(mr_ctx* c, int32_t on) {
  if (!c) return MR_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  c->graphs = on != 0;
  return MR_OK;
}

int mr_ctx_get_stats(mr_ctx* c, mr_stats* st) {
  if (!c || !st) return MR_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  *st = c->stats;
  st->graph_replays = c->graph_replays;
  st->graph_updates = c->graph_updates;
  st->graph_captures = c->graph_captures;
  return MR_OK;
}

int mr_ctx_memory(mr_ctx* c, uint64_t* resident, uint64_t* scratch) {
  if (!c) return MR_ERR_ARG;
  if (resident) *resident = c->resident_bytes;
  if (scratch) *scratch = c->scratch_bytes;
  return MR_OK;
}

// Store a new mesh record in a free slot (one whose attribute arrays no
// other mesh still uses); returns its handle.
static int32_t add_mesh(mr_ctx* c, const MeshRec& m) {
  MeshRec rec = m;
  rec.alive = true;
  for (size_t i = 0; i < c->meshes.size(); i++)
    if (!c->meshes[i].alive && c->meshes[i].sharers == 0) {
      c->meshes[i] = rec;
      return (int32_t)i;
    }
  c->meshes.push_back(rec);
  return (int32_t)c->meshes.size() - 1;
}

static int upload(mr_ctx* c, DevBuf& b, const void* src, size_t bytes) {
  int rc = dev_alloc(c, b, bytes, false);
  if (rc) return rc;
  if (bytes) HIP_TRY(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
  return MR_OK;
}

static bool indices_ok(const int32_t* ix, int64_t n, int32_t limit) {
  for (int64_t i = 0; i < n; i++)
    if (ix[i] < 0 || ix[i] >= limit) return false;
  return true;
}

int mr_mesh_upload(mr_ctx* c, const float* positions, int32_t npos,
                   const float* normals, int32_t nnrm, const float* uvs,
                   int32_t nuv, const int32_t* indices,
                   const int32_t* normal_indices, const int32_t* uv_indices,
                   int32_t ntri, int32_t* handle) {
  if (!c || !handle || npos < 0 || nnrm < 0 || nuv < 0 || ntri < 0) return MR_ERR_ARG;
  if (ntri > 0 && (!positions || !normals || !indices || !normal_indices)) return MR_ERR_ARG;
  if (ntri >= (1 << 30)) return MR_ERR_ARG;
  const bool has_uv = uvs && uv_indices && nuv > 0;
  const int64_t ni = 3 * (int64_t)ntri;
  if (!indices_ok(indices, ni, npos) || !indices_ok(normal_indices, ni, nnrm) ||
      (has_uv && !indices_ok(uv_indices, ni, nuv)))
    return MR_ERR_INDEX;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  MeshRec m;
  m.npos = npos; m.nnrm = nnrm; m.nuv = has_uv ? nuv : 0; m.ntri = ntri;
  m.has_uv = has_uv;
  m.nidx_alias = normal_indices == indices;
  m.tidx_alias = has_uv && uv_indices == indices;
  int rc;
  if ((rc = upload(c, m.pos, positions, (size_t)npos * 12)) ||
      (rc = upload(c, m.nrm, normals, (size_t)nnrm * 12)) ||
      (rc = upload(c, m.idx, indices, (size_t)ni * 4)) ||
      (!m.nidx_alias && (rc = upload(c, m.nidx, normal_indices, (size_t)ni * 4))) ||
      (has_uv && (rc = upload(c, m.uv, uvs, (size_t)nuv * 8))) ||
      (has_uv && !m.tidx_alias && (rc = upload(c, m.tidx, uv_indices, (size_t)ni * 4)))) {
    dev_free(c, m.pos, false); dev_free(c, m.nrm, false); dev_free(c, m.uv, false);
    dev_free(c, m.idx, false); dev_free(c, m.nidx, false); dev_free(c, m.tidx, false);
    return rc;
  }
  *handle = add_mesh(c, m);
  return MR_OK;
}

int mr_mesh_upload_shared(mr_ctx* c, int32_t base, const int32_t* indices,
                          const int32_t* normal_indices, const int32_t* uv_indices,
                          int32_t ntri, int32_t* handle) {
  if (!c || !handle || ntri < 0) return MR_ERR_ARG;
  if (ntri > 0 && (!indices || !normal_indices)) return MR_ERR_ARG;
  if (ntri >= (1 << 30)) return MR_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  if (base < 0 || base >= (int)c->meshes.size() || !c->meshes[base].alive) return MR_ERR_HANDLE;
  const int32_t owner = c->meshes[base].owner >= 0 ? c->meshes[base].owner : base;
  const MeshRec& o = c->meshes[owner];
  const bool has_uv = o.has_uv && uv_indices;
  const int64_t ni = 3 * (int64_t)ntri;
  if (!indices_ok(indices, ni, o.npos) || !indices_ok(normal_indices, ni, o.nnrm) ||
      (has_uv && !indices_ok(uv_indices, ni, o.nuv)))
    return MR_ERR_INDEX;
  HIP_TRY(hipSetDevice(c->device));
  MeshRec m;
  m.npos = o.npos; m.nnrm = o.nnrm; m.nuv = has_uv ? o.nuv : 0; m.ntri = ntri;
  m.has_uv = has_uv;
  m.owner = owner;
  m.nidx_alias = normal_indices == indices;
  m.tidx_alias = has_uv && uv_indices == indices;
  int rc;
  if ((rc = upload(c, m.idx, indices, (size_t)ni * 4)) ||
      (!m.nidx_alias && (rc = upload(c, m.nidx, normal_indices, (size_t)ni * 4))) ||
      (has_uv && !m.tidx_alias && (rc = upload(c, m.tidx, uv_indices, (size_t)ni * 4)))) {
    dev_free(c, m.idx, false); dev_free(c, m.nidx, false); dev_free(c, m.tidx, false);
    return rc;
  }
  c->meshes[owner].sharers++;
  *handle = add_mesh(c, m);
  return MR_OK;
}

int mr_mesh_release(mr_ctx* c, int32_t h) {
  if (!c) return MR_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  if (h < 0 || h >= (int)c->meshes.size() || !c->meshes[h].alive) return MR_ERR_HANDLE;
  hipDeviceSynchronize();  // no batch in flight may still read the mesh
  MeshRec& m = c->meshes[h];
  dev_free(c, m.idx, false); dev_free(c, m.nidx, false); dev_free(c, m.tidx, false);
  m.alive = false;
  // Attribute arrays go with the last mesh that uses them.
  const int32_t owner = m.owner >= 0 ? m.owner : h;
  if (m.owner >= 0) c->meshes[owner].sharers--;
  MeshRec& o = c->meshes[owner];
  if (!o.alive && o.sharers == 0) {
    dev_free(c, o.pos, false); dev_free(c, o.nrm, false); dev_free(c, o.uv, false);
  }
  return MR_OK;
}

int mr_texture_upload(mr_ctx* c, const void* texels, int32_t format,
                      int32_t rows, int32_t cols, int32_t* handle) {
  if (!c || !handle || rows < 0 || cols < 0 || (rows * (int64_t)cols > 0 && !texels))
    return MR_ERR_ARG;
  if (format != MR_TEX_RGB8 && format != MR_TEX_RGB32F) return MR_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  TexRec t;
  t.fmt = format; t.rows = rows; t.cols = cols;
  const size_t bytes = (size_t)rows * cols * (format == MR_TEX_RGB8 ? 3 : 12);
  int rc = upload(c, t.texels, texels, bytes);
  if (rc) return rc;
  t.alive = true;
  for (size_t i = 0; i < c->textures.size(); i++)
    if (!c->textures[i].alive) {
      c->textures[i] = t;
      *handle = (int32_t)i;
      return MR_OK;
    }
  c->textures.push_back(t);
  *handle = (int32_t)c->textures.size() - 1;
  return MR_OK;
}

int mr_texture_release(mr_ctx* c, int32_t h) {
  if (!c) return MR_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  if (h < 0 || h >= (int)c->textures.size() || !c->textures[h].alive) return MR_ERR_HANDLE;
  hipDeviceSynchronize();
  dev_free(c, c->textures[h].texels, false);
  c->textures[h].alive = false;
  return MR_OK;
}

int mr_render(mr_ctx* c, const mr_frame* frame, const mr_outputs* out) {
  if (!c || !frame || !out) return MR_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  return render_frames(c, frame, 1, out);
}

int mr_render_noalloc(mr_ctx* c, const mr_frame* frame, const mr_outputs* out) {
  if (!c || !frame || !out) return MR_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  return render_noalloc(c, frame, out);
}

int mr_render_batch(mr_ctx* c, const mr_frame* frames, int32_t n, const mr_outputs* outs) {
  if (!c || !frames || !outs || n < 1) return MR_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  return render_frames(c, frames, n, outs);
}

}  // extern "C"
