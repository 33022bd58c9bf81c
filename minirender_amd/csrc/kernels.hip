// kernels.hip -- the rasterization pipeline of minirender's Renderer::render()
// (src/Renderer.cpp:311-381) as HIP kernels for gfx950 (wave64).
//
//   k_clear   look-back words, tile counts and merge counters when the
//             scratch is new (afterwards the shading kernel leaves them
//             zero); batch and checkpoint counters when the batch has no
//             triangles (else geometry zeroes them)
//   k_vertex  (a) eye-space transform and projection once per vertex of each
//             (attribute arrays, modelview) group of a job (Renderer.cpp:340-349,
//             191-203): the per-material meshes of a model share one group
//   k_geom    (a)+(b) index gather, near clip, setup and cull of every input
//             triangle (Renderer.cpp:351-380, 163-236).  Each wavefront
//             compacts its survivors into its own region of the item stream
//             (ballot + popcount prefix, no global atomic) and counts its
//             (tile, triangle) pairs with one atomic per distinct tile
//   k_scan    (c) exclusive scan of the tile counts (decoupled look-back);
//             tiles with more pairs than one work item are split into parts,
//             the others with pairs listed for the raster, empty ones marked.
//             For a batch of at most kCkptScanRegions regions, the launch's
//             other workgroups, beside the one-workgroup scan, allocate the
//             edge checkpoints and row tasks of triangles at least 64 pixels
//             wide spanning several tile columns
//   k_bin     (c) scatter of (triangle id, draw) into the tiles' bin lists:
//             per wavefront step, one atomic instruction for all distinct
//             tiles.  For larger batches binning allocates the checkpoints
//             itself (it reads the items anyway)
//   k_rows    (c) row pre-pass: one lane per row of such a triangle walks it
//             once and stores the edge values where it enters each tile
//             column; for the smaller batches the tail of the binning
//             kernel (k_bin_rows)
//   k_tile    (d) per-tile raster (listed tiles, parts of split tiles, row
//             bands) with the tile's 64-bit (depth key, draw-order id) buffer
//             in LDS and ds_min_u64; parts of a split tile merge through
//             global 64-bit atomicMin.  Writes each pixel's winning id
//   k_shade   (e) deferred Blinn-Phong + texture, one lane per pixel in image
//             order, writing RGB8 / depth / winner / float colour / normals /
//             range image (background for the pixels of empty tiles)
//   k_persist every stage above in one launch of a resident grid fed by
//             device work queues (optional; the same stage functions)
//
// A triangle's id is 2 * (flat triangle index in the batch) + clip sub-index,
// which increases with draw order inside a job.  The minimum of
// (depth key, id) is therefore the first-drawn triangle of minimum depth --
// the winner of the reference's sequential strict `z < depth`
// (Renderer.cpp:267).  Triangle setup is recomputed from the projected
// vertices wherever it is needed (geometry, raster, shading) instead of being
// stored: the same fp32 operations on the same inputs give the same bits, and
// ALU is cheaper than writing and re-reading a record per triangle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "common.h"
#include "exact.h"
#include "pipeline.h"

namespace mr {

using u64 = unsigned long long;

constexpr u64 kFlagAgg = 1ull << 62;
constexpr u64 kFlagPre = 2ull << 62;
constexpr u64 kValMask = (1ull << 62) - 1;

// Decoupled look-back (single thread): publish this block's aggregate, sum
// predecessors back to the first inclusive prefix, publish ours.  Block ids
// come from a ticket counter, so every predecessor is already running.
__device__ u64 lookback(u64* status, int idx, u64 aggregate, u64* counters) {
  if (idx == 0) {
    __hip_atomic_store(&status[0], kFlagPre | aggregate, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    return 0;
  }
  __hip_atomic_store(&status[idx], kFlagAgg | aggregate, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  u64 excl = 0;
  int j = idx - 1;
  uint32_t spins = 0;
  while (j >= 0) {
    const u64 w = __hip_atomic_load(&status[j], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
    const u64 flag = w >> 62;
    if (flag == 0) {
      if (++spins > kSpinLimit) {
        atomicOr(&counters[kError], 1ull);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    excl += w & kValMask;
    if (flag == 2) break;
    --j;
  }
  __hip_atomic_store(&status[idx], kFlagPre | (excl + aggregate),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return excl;
}

// Each stage is a device function of its workgroup index vb in a grid of
// nvb workgroups of 256 threads: run by its own kernel (vb = blockIdx.x),
// or by the persistent kernel, whose workgroups take stage work from queues.
__device__ __forceinline__ void stage_clear(const BatchArgs& a, int vb, int nvb) {
  // The checkpoint segment counters too: a batch without triangles (no
  // geometry, which zeroes them otherwise) must not leave the row pre-pass
  // the last batch's row tasks.
  if (vb == 0 && threadIdx.x < kCkptSegs) a.ckpt_ctr[threadIdx.x * kCtrStride] = 0;
  const int n = max(max(a.scan_blocks, a.ntiles), (int)kNumCounters);
  for (int i = vb * 256 + (int)threadIdx.x; i < n; i += nvb * 256) {
    if (i < kNumCounters) a.counters[i] = 0;
    if (i < a.scan_blocks) a.scan_status[i] = 0;
    if (i < a.ntiles) {
      a.tile_count[(size_t)i * kCtrStride64] = 0;
      a.tile_done[i] = 0;
    }
  }
}

__global__ __launch_bounds__(256) void k_clear(BatchArgs a) {
  stage_clear(a, blockIdx.x, gridDim.x);
}

// Last index in [lo, hi] with begin[index] <= x (begin ascending).
template <typename T>
__device__ inline int search_last(const T* begin, int lo, int hi, int64_t x) {
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((int64_t)begin[mid] <= x) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// search_last by a whole wavefront (every lane active, x uniform): the 64
// lanes probe 64 evenly spaced entries per step, so n entries take
// ceil(log64 n) dependent loads instead of log2 n (the first workgroup
// step of the vertex and geometry kernels waits on this search).
// begin[0] <= x.
__device__ inline int wave_search_last(const int64_t* begin, int n, int64_t x) {
  const int lane = threadIdx.x & 63;
  int lo = 0, span = n;  // the answer lies in [lo, lo + span)
  while (span > 1) {
    const int step = (span + 63) >> 6;
    const int idx = lo + lane * step;
    const bool le = idx < lo + span && begin[idx] <= x;
    const u64 m = __ballot(le);  // lane 0 (idx = lo) is always set
    const int k = 63 - __builtin_clzll(m);
    const int nlo = lo + k * step;
    span = min(step, lo + span - nlo);
    lo = __builtin_amdgcn_readfirstlane(nlo);
    span = __builtin_amdgcn_readfirstlane(span);
  }
  return lo;
}

// Job of tile (shading block) x from the job hints: the job of x's span of
// kHintSpan, stepped past job boundaries inside the span (at most one unless
// jobs are smaller than the span).  One scalar load where a binary search
// over the jobs' first tiles takes log2(njobs) dependent ones.
__device__ inline int job_lookup(const int2* hint, const int32_t* base, int njobs, int64_t x) {
  if (njobs == 1) return 0;
  const int2 h = hint[x >> kHintShift];
  int j = h.x;
  if (x >= h.y) {
    j++;
    while (j + 1 < njobs && (int64_t)base[j + 1] <= x) j++;
  }
  return j;
}

// Last index i in [0, N) with rs[i] <= q, for non-decreasing rs with
// rs[0] <= q: a 4-ary search (three independent LDS reads per level), half
// the dependent steps of a binary search.
template <int N>
__device__ inline int owner_search(const int* rs, int q) {
  int lo = 0;
  int n = N;
#pragma unroll
  for (int it = 0; it < 8; it++) {
    if (n >= 4) {
      const int s = n >> 2;
      const int a1 = rs[lo + s], a2 = rs[lo + 2 * s], a3 = rs[lo + 3 * s];
      lo += (a1 <= q ? s : 0) + (a2 <= q ? s : 0) + (a3 <= q ? s : 0);
      n = s;
    } else if (n == 2) {
      if (rs[lo + 1] <= q) lo += 1;
      n = 1;
    }
  }
  return lo;
}

// Draw of flat triangle g in job J: the job's only draw (ncached < 0), the
// job's draw table cached in LDS when it fits (ncached > 0), else searched in
// global memory.
__device__ inline int lookup_draw(const int64_t* dcache, int ncached,
                                  const BatchArgs& a, const JobDesc& J, int64_t g) {
  if (ncached < 0) return J.draw_begin;  // the job has one draw
  if (ncached > 0) return J.draw_begin + search_last(dcache, 0, ncached - 1, g);
  return search_last(a.draw_begin, J.draw_begin, J.draw_end - 1, g);
}

// Fill the LDS draw cache for job J; returns the number of cached entries.
__device__ inline int fill_draw_cache(int64_t* dcache, const BatchArgs& a,
                                      const JobDesc& J, int tid, int nt) {
  const int nd = J.draw_end - J.draw_begin;
  const int n = nd <= kMaxLdsDraws ? nd : 0;
  for (int i = tid; i < n; i += nt) dcache[i] = a.draw_begin[J.draw_begin + i];
  return n;
}

// A draw descriptor read through the constant address space: for a
// wavefront-uniform index its words are fetched with scalar loads (into
// SGPRs) instead of one vector load per field.
__device__ inline DrawDesc uniform_draw(const BatchArgs& a, int d) {
  typedef const __attribute__((address_space(4))) uint32_t* ConstWords;
  const ConstWords src = (ConstWords)(a.draws + d);
  DrawDesc out;
  uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(DrawDesc) / 4); i++) dst[i] = src[i];
  return out;
}

// The same for a job descriptor indexed by a workgroup-uniform value.
__device__ inline JobDesc uniform_job(const BatchArgs& a, int j) {
  typedef const __attribute__((address_space(4))) uint32_t* ConstWords;
  const ConstWords src = (ConstWords)(a.jobs + j);
  JobDesc out;
  uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(JobDesc) / 4); i++) dst[i] = src[i];
  return out;
}

// Global-memory view of a device pointer read from a descriptor.  The
// compiler cannot tell where such a pointer points and accesses it with
// flat instructions, which also count against the scalar-load counter: every
// wait for a descriptor field then waits for them too.  Through this view
// they are global loads and stores (vector memory only), with a scalar base
// and 32-bit offsets where the index allows.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* G(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

// A per-pixel output element at a 32-bit byte offset from its base: a frame
// has at most kMaxDim^2 = 2^28 pixels, so offsets of up to 12 B per pixel
// fit, and the access takes its (wavefront-uniform) base from scalar
// registers instead of forming a 64-bit address per lane.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* px_at(T* base, uint32_t off) {
  typedef __attribute__((address_space(1))) char GChar;
  return (__attribute__((address_space(1))) T*)((GChar*)G(base) + off);
}
static_assert((uint64_t)kMaxDim * kMaxDim * 12 <= 0xFFFFFFFFull, "pixel offsets are 32-bit");

__device__ inline f3 load3(const float* p, int i) {
  const auto* q = G(p);
  return f3{q[3 * i], q[3 * i + 1], q[3 * i + 2]};
}

// Corner indices of triangle t, their projections and eye depths
// (Renderer.cpp:344-345, 364-366): one 16-byte load per corner.  The eye
// x, y are left 0: corner_eyes computes them where the near-plane clip
// needs them.
__device__ inline void corners(const BatchArgs& a, const DrawDesc& D, int t, int vi[3],
                               f3 e[3], ScreenVtx v[3]) {
  const auto* ip = G(D.idx) + 3 * t;
  vi[0] = ip[0];
  vi[1] = ip[1];
  vi[2] = ip[2];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float4 r = a.vscr[D.vbase + vi[k]];
    v[k] = ScreenVtx{r.x, r.y, r.z};
    e[k] = f3{0.0f, 0.0f, r.w};
  }
}

// The corners' eye x, y, transformed again from the positions (the vertex
// kernel's arithmetic on the same modelview: the same bits, Renderer.cpp:
// 344-345); only a triangle across the near plane or of a draw without the
// near test needs them.
__device__ inline void corner_eyes(const DrawDesc& D, const int vi[3], f3 e[3]) {
  f3 p[3];
#pragma unroll
  for (int k = 0; k < 3; k++) p[k] = load3(D.pos, vi[k]);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const f3 q = xform(D.mv, p[k]);
    e[k].x = q.x;
    e[k].y = q.y;
  }
}

// Near test, clip and setup of triangle t (Renderer.cpp:169-236).  Returns
// the number of surviving sub-triangles (0..2) and their sub-indices.
// Unclipped triangles use the per-vertex projections; clipped ones project
// their clipped corners here.
__device__ inline int tri_setup(const BatchArgs& a, const DrawDesc& D, const JobDesc& J, int t,
                                TriSetup& s0, int& sub0, TriSetup& s1,
                                int& sub1, bool& clipped) {
  int vi[3];
  f3 e[3];
  ScreenVtx v[3];
  corners(a, D, t, vi, e, v);
  const float zn = J.znear;
  const bool persp = J.persp != 0;
  const float w = (float)J.W, h = (float)J.H;
  const bool b0 = e[0].z > zn, b1 = e[1].z > zn, b2 = e[2].z > zn;
  clipped = false;
  if (!D.no_clip && (b0 || b1 || b2)) {
    if (b0 && b1 && b2) return 0;
    clipped = true;
    corner_eyes(D, vi, e);
    f3 a0, a1, a2, c0, c1, c2;
    const int ns = clip_positions(zn, e[0], e[1], e[2], a0, a1, a2, c0, c1, c2);
    int n = 0;
    if (setup_triangle(J.P, persp, w, h, a0, a1, a2, s0)) {
      sub0 = 0;
      n = 1;
    }
    if (ns == 2) {
      TriSetup tmp;
      if (setup_triangle(J.P, persp, w, h, c0, c1, c2, tmp)) {
        if (n == 0) { s0 = tmp; sub0 = 1; } else { s1 = tmp; sub1 = 1; }
        ++n;
      }
    }
    return n;
  }
  if (setup_screen(w, h, v[0], v[1], v[2], s0)) {
    sub0 = 0;
    return 1;
  }
  return 0;
}

// Whether a (sub-)triangle covers any pixel centre of its bounding box.
// For a box of at most kCoverProbe x kCoverProbe pixels this is decided with
// the raster's own arithmetic: each row starts at the box's left column
// (Renderer.cpp:240-242) and steps e += n.x (:243), the chain the raster
// continues from there, so a triangle found to cover nothing here covers
// nothing in the raster either and is not binned.  Most triangles smaller
// than a pixel cover no centre.  Larger boxes are assumed to cover one.
constexpr int kCoverProbe = 3;
__device__ inline bool covers_any(const TriSetup& s) {
  if (s.bx1 - s.bx0 >= kCoverProbe || s.by1 - s.by0 >= kCoverProbe) return true;
  for (int row = s.by0; row <= s.by1; row++) {
    float e1, e2;
    edge_start(s.n1x, s.n1y, s.n2x, s.n2y, s.p0x, s.p0y, s.p2x, s.p2y,
               (float)s.bx0 + 0.5f, (float)row + 0.5f, e1, e2);
    for (int c = s.bx0; c <= s.bx1; c++) {
      if (covered(e1, e2)) return true;
      e1 = e1 + s.n1x;
      e2 = e2 + s.n2x;
    }
  }
  return false;
}

// Setup of sub-triangle `sub` of a binned triangle from its corners.
__device__ inline void tri_setup_corners(const JobDesc& J, const f3 e[3], const ScreenVtx v[3],
                                         int sub, bool no_clip, TriSetup& s) {
  const float zn = J.znear;
  const float w = (float)J.W, h = (float)J.H;
  if (!no_clip && (e[0].z > zn || e[1].z > zn || e[2].z > zn)) {
    const bool persp = J.persp != 0;
    f3 a0, a1, a2, c0, c1, c2;
    clip_positions(zn, e[0], e[1], e[2], a0, a1, a2, c0, c1, c2);
    if (sub) setup_triangle(J.P, persp, w, h, c0, c1, c2, s);
    else setup_triangle(J.P, persp, w, h, a0, a1, a2, s);
    return;
  }
  setup_screen(w, h, v[0], v[1], v[2], s);
}

// Setup of sub-triangle `sub` of a binned triangle (which survived setup).
__device__ inline void tri_setup_sub(const BatchArgs& a, const DrawDesc& D, const JobDesc& J,
                                     int t, int sub, TriSetup& s) {
  int vi[3];
  f3 e[3];
  ScreenVtx v[3];
  corners(a, D, t, vi, e, v);
  if (!D.no_clip && (e[0].z > J.znear || e[1].z > J.znear || e[2].z > J.znear))
    corner_eyes(D, vi, e);
  tri_setup_corners(J, e, v, sub, D.no_clip != 0, s);
}

// (a): kVertsPerLane vertices per lane of every vertex group: eye-space
// position and its projection (Renderer.cpp:340-349, 191-203), once per
// vertex instead of once per triangle corner in each later kernel.  A
// wavefront takes 64 * kVertsPerLane consecutive vertices per step and loads
// all of a lane's positions before transforming any: the kernel is bound by
// loads in flight.
constexpr int kVertsPerLane = 2;
constexpr int kVertexStep = 256 * kVertsPerLane;  // vertices per workgroup step

__device__ inline void vertex_record(const BatchArgs& a, const DrawDesc& D, const JobDesc& J,
                                     int64_t i, f3 p) {
  const f3 e = xform(D.mv, p);
  const ScreenVtx v = project_vertex(J.P, J.persp != 0, (float)J.W, (float)J.H, e);
  a.vscr[i] = make_float4(v.px, v.py, v.zv, e.z);
}

__device__ __forceinline__ void stage_vertex(const BatchArgs& a, int vb, int nvb) {
  // Contiguous range per workgroup (a multiple of kVertexStep): one group
  // search, then stepping.
  const int64_t chunk = ((a.total_verts + (int64_t)nvb * kVertexStep - 1) /
                         ((int64_t)nvb * kVertexStep)) * kVertexStep;
  const int64_t r_begin = (int64_t)vb * chunk;
  const int64_t r_end = min(r_begin + chunk, a.total_verts);
  const int lane = threadIdx.x & 63;
  int gw = wave_search_last(a.vg_begin, a.nvgroups, min(r_begin, a.total_verts));
  gw = __builtin_amdgcn_readfirstlane(gw);
  for (int64_t base = r_begin; base < r_end; base += kVertexStep) {
    const int64_t wfirst = base + kVertsPerLane * (threadIdx.x & ~63);
    if (wfirst >= r_end) continue;
    while (gw + 1 < a.nvgroups && a.vg_begin[gw + 1] <= wfirst) gw++;
    gw = __builtin_amdgcn_readfirstlane(gw);
    const int64_t gw_end = gw + 1 < a.nvgroups ? a.vg_begin[gw + 1] : 0x7FFFFFFFFFFFFFFFll;
    if (min(wfirst + 64 * kVertsPerLane - 1, r_end - 1) < gw_end) {
      // The wavefront's vertices are all in group gw: descriptors through
      // scalar loads.
      const DrawDesc Du = uniform_draw(a, __builtin_amdgcn_readfirstlane(a.vg_draw[gw]));
      const JobDesc Ju = uniform_job(a, __builtin_amdgcn_readfirstlane(Du.job));
      const int64_t gb = a.vg_begin[gw];
      f3 p[kVertsPerLane];
#pragma unroll
      for (int h = 0; h < kVertsPerLane; h++) {
        const int64_t i = wfirst + 64 * h + lane;
        p[h] = f3{0.0f, 0.0f, 0.0f};
        if (i < r_end) p[h] = load3(Du.pos, (int)(i - gb));
      }
#pragma unroll
      for (int h = 0; h < kVertsPerLane; h++) {
        const int64_t i = wfirst + 64 * h + lane;
        if (i < r_end) vertex_record(a, Du, Ju, i, p[h]);
      }
    } else {
      for (int h = 0; h < kVertsPerLane; h++) {
        const int64_t i = wfirst + 64 * h + lane;
        if (i >= r_end) continue;
        int g = gw;
        while (g + 1 < a.nvgroups && a.vg_begin[g + 1] <= i) g++;
        const DrawDesc& D = a.draws[a.vg_draw[g]];
        vertex_record(a, D, a.jobs[D.job], i, load3(D.pos, (int)(i - a.vg_begin[g])));
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_vertex(BatchArgs a) {
  stage_vertex(a, blockIdx.x, gridDim.x);
}

// Pairs of one wavefront step: lanes with ok set hold a pair (tile key, id,
// draw).  Lanes with equal keys form a group (at most kMaxGroups groups per
// step; further lanes are groups of one) whose leader calls op(key, group
// size) once.  kScatter: op returns the first bin slot of the group and
// every lane writes its (id, draw) pair behind it.
template <bool kScatter, typename Op>
__device__ inline void wave_pairs_apply(const BatchArgs& a, bool ok, int key, uint32_t id,
                                        uint32_t draw, uint32_t work, Op&& op) {
  const int lane = threadIdx.x & 63;
  // Group equal keys (bounded: steps of a wide triangle have ~64 distinct
  // tiles, where grouping only costs).
  u64 pend = __ballot(ok);
  int leader = lane;
  uint32_t rank = 0, gsize = 1;
  constexpr int kMaxGroups = 8;
  for (int it = 0; it < kMaxGroups && pend; it++) {
    const int l = __builtin_ctzll(pend);
    const int kk = __builtin_amdgcn_readlane(key, l);
    const u64 grp = pend & __ballot(key == kk);
    if ((grp >> lane) & 1ull) {
      leader = l;
      rank = (uint32_t)__popcll(grp & ((1ull << lane) - 1ull));
    }
    if (lane == l) gsize = (uint32_t)__popcll(grp);
    pend &= ~grp;
  }
  uint32_t old = 0;
  if (ok && lane == leader) old = op(key, gsize, work * gsize);
  if (kScatter) {
    old = __shfl(old, leader);
    if (ok) {
      const uint32_t slot = old + rank;
      if ((u64)slot < a.bins_cap) a.bins[slot] = make_uint2(id, draw);
      else atomicOr(&a.counters[kBinOverflow], 1ull);
    }
  }
}

// Walk work of a (triangle, tile) pair: the triangle's bbox pixels inside the
// tile, 4 per bbox row there and 16 for the entry (a rough cost of the
// raster's setup, row starts and column steps).
__device__ inline uint32_t pair_work(int rows, int cols) {
  return (uint32_t)(rows * cols + 4 * rows + 16);
}

// For every active lane's pixel bbox, visit all (lane, tile) pairs with the
// whole wavefront (64 pairs per step, owners found by binary search of the
// wave prefix sum) and hand each step to wave_pairs_apply.  When every
// active bbox lies in a single tile (small triangles, the common case) the
// lanes' own tiles are the step.
template <bool kScatter, typename Op>
__device__ void wave_tile_pairs(const BatchArgs& a, bool active, uint32_t id, uint32_t draw,
                                int tbase, int tiles_x, int bx0, int bx1, int by0, int by1,
                                int sw, int sh, Op&& op) {
  const int lane = threadIdx.x & 63;
  if (!active) { bx0 = bx1 = by0 = by1 = 0; }
  const int tx0 = bx0 >> sw, tx1 = bx1 >> sw, ty0 = by0 >> sh, ty1 = by1 >> sh;
  const int cnt = active ? (tx1 - tx0 + 1) * (ty1 - ty0 + 1) : 0;
  if (__ballot(cnt > 1) == 0ull) {
    const bool ok = cnt == 1;
    const uint32_t work = kScatter ? 0u : pair_work(by1 - by0 + 1, bx1 - bx0 + 1);
    wave_pairs_apply<kScatter>(a, ok, ok ? tbase + ty0 * tiles_x + tx0 : -1, id, draw, work, op);
    return;
  }
  int P = cnt;  // inclusive prefix over the wave
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(P, o);
    if (lane >= o) P += y;
  }
  const int C = __shfl(P, 63);
  const int ex = P - cnt;
  for (int j = 0; j < C; j += 64) {
    const int q = j + lane;
    int lo = 0, hi = 63;
#pragma unroll
    for (int s = 0; s < 6; s++) {
      const int mid = (lo + hi) >> 1;
      if (__shfl(P, mid) > q) hi = mid; else lo = mid + 1;
    }
    const bool ok = q < C;
    const int k = q - __shfl(ex, lo);
    const int ox0 = __shfl(tx0, lo), ox1 = __shfl(tx1, lo), oy0 = __shfl(ty0, lo);
    const int ob = __shfl(tbase, lo), otx = __shfl(tiles_x, lo);
    const int w = max(ox1 - ox0 + 1, 1);
    const int r = k / w;
    const int ptx = ox0 + (k - r * w), pty = oy0 + r;
    const int key = ok ? ob + pty * otx + ptx : -1;
    uint32_t oid = 0, odraw = 0, work = 0;
    if (kScatter) {
      oid = __shfl(id, lo);
      odraw = __shfl(draw, lo);
    } else {
      const int obx0 = __shfl(bx0, lo), obx1 = __shfl(bx1, lo);
      const int oby0 = __shfl(by0, lo), oby1 = __shfl(by1, lo);
      const int rows = min(oby1, ((pty + 1) << sh) - 1) - max(oby0, pty << sh) + 1;
      const int cols = min(obx1, ((ptx + 1) << sw) - 1) - max(obx0, ptx << sw) + 1;
      work = pair_work(rows, cols);
    }
    wave_pairs_apply<kScatter>(a, ok, key, oid, odraw, work, op);
  }
}

// (a)+(b): one lane per input triangle; wavefront w writes its survivors to
// items[w * kRegion ...] and their number to region_count[w].
__device__ __forceinline__ void stage_geom(const BatchArgs& a, int vb, int nvb) {
  // The batch counters and checkpoint segment counters are first used by the
  // scan and the binner: clear them here.
  if (vb == 0 && threadIdx.x < kNumCounters) a.counters[threadIdx.x] = 0;
  if (vb == 0 && threadIdx.x < kCkptSegs) a.ckpt_ctr[threadIdx.x * kCtrStride] = 0;
  const int lane = threadIdx.x & 63;
  const int sw = __ffs(a.tile_w) - 1, sh = __ffs(a.tile_h) - 1;
  uint32_t nclipped = 0;
  // Each workgroup takes one contiguous range of triangles (a multiple of
  // 256), so a wavefront finds its draw once and then only steps over draw
  // boundaries; its draw and job descriptors are read with scalar loads.
  const int64_t chunk = ((a.total_tris + (int64_t)nvb * 256 - 1) /
                         ((int64_t)nvb * 256)) * 256;
  const int64_t r_begin = (int64_t)vb * chunk;
  const int64_t r_end = min(r_begin + chunk, a.total_tris);
  int dw = wave_search_last(a.draw_begin, a.ndraws, min(r_begin, a.total_tris));
  dw = __builtin_amdgcn_readfirstlane(dw);
  for (int64_t base = r_begin; base < r_end; base += 256) {
    const int64_t wfirst = base + (threadIdx.x & ~63);
    if (wfirst >= r_end) continue;
    const int64_t g = wfirst + lane;
    TriSetup s0, s1;
    s0.bx0 = s0.bx1 = s0.by0 = s0.by1 = 0;
    s1.bx0 = s1.bx1 = s1.by0 = s1.by1 = 0;
    int sub0 = 0, sub1 = 0, ns = 0;
    bool clipped = false;
    uint32_t draw = 0;
    int tbase = 0, tiles_x = 1;
    // The wave's draw: step the uniform draw index to the wave's first
    // triangle; lanes past the next draw's start step further (draw
    // boundaries inside a wavefront are rare).
    while (dw + 1 < a.ndraws && a.draw_begin[dw + 1] <= wfirst) dw++;
    dw = __builtin_amdgcn_readfirstlane(dw);
    const int64_t next_begin = dw + 1 < a.ndraws ? a.draw_begin[dw + 1] : 0x7FFFFFFFFFFFFFFFll;
    if (g < r_end) {
      int d = dw;
      if (g >= next_begin)
        while (d + 1 < a.ndraws && a.draw_begin[d + 1] <= g) d++;
      auto setup = [&](const DrawDesc& D, const JobDesc& J) {
        ns = tri_setup(a, D, J, (int)(g - D.tri_begin), s0, sub0, s1, sub1, clipped);
        tbase = D.tile_base;
        tiles_x = D.tiles_x;
      };
      if (__ballot(d == dw) == __ballot(1)) {
        const DrawDesc Du = uniform_draw(a, dw);
        setup(Du, uniform_job(a, __builtin_amdgcn_readfirstlane(Du.job)));
      } else {
        setup(a.draws[d], a.jobs[a.draws[d].job]);
      }
      draw = (uint32_t)d;
      // A triangle that covers no pixel centre is not binned (clipped ones,
      // rare, are binned as they are).
      if (!clipped && ns == 1 && !covers_any(s0)) ns = 0;
    }
    nclipped += (uint32_t)__popcll(__ballot(clipped));
    const u64 b1 = __ballot(ns >= 1), b2 = __ballot(ns == 2);
    const int64_t region = wfirst >> 6;
    if (lane == 0) a.region_count[region] = (uint32_t)(__popcll(b1) + __popcll(b2));
    if (!b1) continue;
    const u64 lt = (1ull << lane) - 1ull;
    const int64_t pos = region * kRegion + __popcll(b1 & lt) + __popcll(b2 & lt);
    const uint32_t id0 = (uint32_t)(2 * g + sub0), id1 = (uint32_t)(2 * g + sub1);
    if (ns >= 1)
      a.items[pos] = BinItem{id0, draw, (uint16_t)s0.bx0, (uint16_t)s0.bx1,
                             (uint16_t)s0.by0, (uint16_t)s0.by1};
    if (ns == 2)
      a.items[pos + 1] = BinItem{id1, draw, (uint16_t)s1.bx0, (uint16_t)s1.bx1,
                                 (uint16_t)s1.by0, (uint16_t)s1.by1};
    // Pairs and walk work per tile, in one 64-bit atomic per group.
    auto count = [&](int key, uint32_t n, uint32_t work) -> uint32_t {
      atomicAdd(&a.tile_count[(size_t)key * kCtrStride64], ((u64)work << 32) | (u64)n);
      return 0u;
    };
    wave_tile_pairs<false>(a, ns >= 1, id0, draw, tbase, tiles_x, s0.bx0, s0.bx1, s0.by0,
                           s0.by1, sw, sh, count);
    if (b2)
      wave_tile_pairs<false>(a, ns == 2, id1, draw, tbase, tiles_x, s1.bx0, s1.bx1, s1.by0,
                             s1.by1, sw, sh, count);
  }
  __shared__ uint32_t sh_clipped;
  if (threadIdx.x == 0) sh_clipped = 0;
  __syncthreads();
  if (lane == 0 && nclipped) atomicAdd(&sh_clipped, nclipped);
  __syncthreads();
  if (threadIdx.x == 0 && vb < kStatSlots) a.block_stats[vb] = sh_clipped;
}

__global__ __launch_bounds__(256, 7) void k_geom(BatchArgs a) {
  stage_geom(a, blockIdx.x, gridDim.x);
}

// Row bands of a tile rastered by several workgroups (each with all of the
// tile's entries, rows of its band only): tiles with at most frag_size pairs
// but more walk work than band_work, about band_work per band, whose entries
// walk kBandMinEntryWork or more on average (every band sets up every entry:
// tiles of many small triangles are split by entries instead).  0: none.
constexpr uint32_t kBandMinEntryWork = 64;
__device__ inline uint32_t band_count(const BatchArgs& a, uint32_t cnt, uint32_t work) {
  if (cnt == 0 || cnt > a.frag_size || a.band_work == 0 || work <= a.band_work) return 0;
  if (work < kBandMinEntryWork * cnt) return 0;
  uint32_t nb = (work + a.band_work - 1) / a.band_work;
  if (nb > (uint32_t)a.tile_h) nb = (uint32_t)a.tile_h;
  const uint32_t bh = ((uint32_t)a.tile_h + nb - 1) / nb;
  nb = ((uint32_t)a.tile_h + bh - 1) / bh;
  return nb >= 2 ? nb : 0;
}

// (c) exclusive scan of (tile, triangle) pairs per tile -> offsets and
// cursors; tiles with more than frag_size pairs get a merge slot and their
// parts are listed for the raster, tiles with much walk work their row
// bands.
__device__ __forceinline__ void stage_scan(const BatchArgs& a) {
  __shared__ int sh_blk;
  __shared__ u64 sh_w[kScanBlock / 64];
  __shared__ u64 sh_excl;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // (A single workgroup needs no ticket: it is workgroup 0.)
  if (tid == 0) sh_blk = a.scan_blocks == 1 ? 0 : (int)atomicAdd(&a.counters[kScanTicket], 1ull);
  __syncthreads();
  const int blk = sh_blk;
  const int base = blk * (kScanBlock * kScanItems) + tid * kScanItems;
  uint32_t v[kScanItems], nbv[kScanItems];
  u64 sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    const u64 tc = (base + i < a.ntiles) ? a.tile_count[(size_t)(base + i) * kCtrStride64] : 0ull;
    v[i] = (uint32_t)tc;
    nbv[i] = band_count(a, v[i], (uint32_t)(tc >> 32));
    sum += v[i];
  }
  u64 x = sum;  // wave inclusive scan
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const u64 y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh_w[wave] = x;
  __syncthreads();
  u64 wpre = 0, btot = 0;
#pragma unroll
  for (int w = 0; w < kScanBlock / 64; w++) {
    if (w < wave) wpre += sh_w[w];
    btot += sh_w[w];
  }
  if (tid == 0) {
    const u64 excl = lookback(a.scan_status, blk, btot, a.counters);
    sh_excl = excl;
    if (blk == a.scan_blocks - 1) {
      a.counters[kBinEntries] = excl + btot;
      // More pairs than the bin list holds (32-bit offsets and cursors could
      // then wrap into it): flagged here, whatever the binner's checks see.
      if (excl + btot > a.bins_cap) atomicOr(&a.counters[kBinOverflow], 1ull);
    }
  }
  // Split (heavy) tiles: the workgroup reserves its merge slots and part
  // list entries with one atomic each (same-address atomics serialise); the
  // same for its entries of the raster's tile list (tiles rastered whole:
  // with entries, not split).
  uint32_t nh = 0, npt = 0, nl = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++)
    if (base + i < a.ntiles && v[i] > a.frag_size) {
      nh++;
      npt += (v[i] + a.frag_size - 1) / a.frag_size;
    } else if (base + i < a.ntiles && nbv[i]) {
      npt += nbv[i];
    } else if (base + i < a.ntiles && v[i] > 0) {
      nl++;
    }
  uint32_t hx = nh, px = npt, lx = nl;  // wave inclusive scans
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t yh = __shfl_up(hx, o), yp = __shfl_up(px, o), yl = __shfl_up(lx, o);
    if (lane >= o) {
      hx += yh;
      px += yp;
      lx += yl;
    }
  }
  __shared__ uint32_t sh_h[kScanBlock / 64], sh_p[kScanBlock / 64], sh_l[kScanBlock / 64];
  __shared__ u64 sh_hbase, sh_pbase, sh_lbase;
  if (lane == 63) {
    sh_h[wave] = hx;
    sh_p[wave] = px;
    sh_l[wave] = lx;
  }
  __syncthreads();
  uint32_t hpre = 0, ppre = 0, lpre = 0, htot = 0, ptot = 0, ltot = 0;
#pragma unroll
  for (int w = 0; w < kScanBlock / 64; w++) {
    if (w < wave) {
      hpre += sh_h[w];
      ppre += sh_p[w];
      lpre += sh_l[w];
    }
    htot += sh_h[w];
    ptot += sh_p[w];
    ltot += sh_l[w];
  }
  if (tid == 0 && a.scan_blocks == 1) {  // the only workgroup: bases 0, totals stored
    a.counters[kHeavy] = htot;
    a.counters[kParts] = ptot;
    a.counters[kListed] = ltot;
    sh_hbase = sh_pbase = sh_lbase = 0;
  } else if (tid == 0) {
    sh_hbase = htot ? atomicAdd(&a.counters[kHeavy], (u64)htot) : 0;
    sh_pbase = ptot ? atomicAdd(&a.counters[kParts], (u64)ptot) : 0;
    sh_lbase = ltot ? atomicAdd(&a.counters[kListed], (u64)ltot) : 0;
  }
  __syncthreads();
  u64 hs = sh_hbase + hpre + (hx - nh), ps = sh_pbase + ppre + (px - npt);
  u64 ls = sh_lbase + lpre + (lx - nl);
  u64 run = sh_excl + wpre + (x - sum);
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    const int t = base + i;
    if (t < a.ntiles) {
      const uint32_t o = run > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)run;
      a.tile_off[t] = o;
      a.tile_cursor[(size_t)t * kCtrStride] = o;
      uint32_t slot = kNoSlot;
      if (v[i] > a.frag_size) {
        const uint32_t nparts = (v[i] + a.frag_size - 1) / a.frag_size;
        if (hs < a.heavy_cap && ps + nparts <= a.parts_cap) {
          slot = (uint32_t)hs;
          for (uint32_t p = 0; p < nparts; p++) a.parts[ps + p] = make_uint2((uint32_t)t, p);
        } else {
          atomicOr(&a.counters[kHeavyOverflow], 1ull);
        }
        hs++;
        ps += nparts;
      } else if (nbv[i]) {
        if (ps + nbv[i] <= a.parts_cap) {
          for (uint32_t p = 0; p < nbv[i]; p++)
            a.parts[ps + p] = make_uint2((uint32_t)t, kBandPart | (nbv[i] << 8) | p);
        } else {
          atomicOr(&a.counters[kHeavyOverflow], 1ull);
        }
        ps += nbv[i];
      } else if (v[i] > 0) {
        a.tile_list[ls++] = make_uint4((uint32_t)t, o, v[i], 0u);
      } else {
        slot = kEmptyTile;  // shading writes its pixels' background
      }
      a.tile_heavy[t] = slot;
    }
    run += v[i];
  }
}


// Edge checkpoints of the wavefront's items: an item at least
// kCkptMinSpan pixels wide that spans ntx > 1 tile columns gets ntx entries
// per pixel row (one per column, where the row enters it) and one row task
// per pixel row for the row pre-pass, all taken from the checkpoint segment
// of the wavefront (wavefronts spread over kCkptSegs segments) by one 64-bit
// atomic: entries in the low 40 bits of the segment counter, row tasks in
// the high 24.  Every such row holds at least two entries, so a segment's
// row tasks fit in half its entry count.  The base is stored pre-offset by
// the item's first row and column, so that a reader needs only the row and
// the tile column.  Items narrower than kCkptMinSpan pixels get no entry
// (their readers test the width first); wide ones that found no room get
// {0, 0}.  Either way the raster and shading walk such rows from the bbox's
// left column (fewer than kCkptMinSpan steps beyond a tile for the narrow).
constexpr int kCkptMinSpan = 64;
constexpr u64 kCkptEntryMask = (1ull << 40) - 1;

__device__ inline void ckpt_alloc(const BatchArgs& a, bool act, const BinItem& it, int sw,
                                  uint32_t item, int vb) {
  const int lane = threadIdx.x & 63;
  const int btx = it.bx0 >> sw;
  const int ntx = act ? (it.bx1 >> sw) - btx + 1 : 0;
  const int rows = it.by1 - it.by0 + 1;
  const bool want = ntx > 1 && it.bx1 - it.bx0 + 1 >= kCkptMinSpan;
  if (__ballot(want) == 0ull) return;  // the common case: no wide triangle here
  const u64 need = want ? (u64)ntx * (u64)rows : 0ull;
  const uint32_t nrow = want ? (uint32_t)rows : 0u;
  u64 P = need;  // inclusive prefixes over the wave
  uint32_t R = nrow;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const u64 y = __shfl_up(P, o);
    const uint32_t yr = __shfl_up(R, o);
    if (lane >= o) {
      P += y;
      R += yr;
    }
  }
  const u64 total = __shfl(P, 63);
  if (total == 0ull) {
    if (want) a.ckpt_base[it.id] = make_int2(0, 0);
    return;
  }
  const uint32_t rtotal = __shfl(R, 63);
  const u64 seg_cap = a.ckpt_cap / kCkptSegs;
  const u64 seg_rows = seg_cap / 2;
  const int seg = (int)(((uint32_t)vb * 4u + (threadIdx.x >> 6)) % kCkptSegs);
  u64 old = 0;
  if (lane == 63) old = atomicAdd(&a.ckpt_ctr[seg * kCtrStride], ((u64)rtotal << 40) | total);
  old = __shfl(old, 63);
  const u64 base = old & kCkptEntryMask, rbase = old >> 40;
  bool ok = false;
  if (want) {
    const u64 first = base + P - need;  // within the segment
    const int64_t off = (int64_t)(seg * seg_cap + first) - (int64_t)it.by0 * ntx - btx;
    ok = need && first + need <= seg_cap && off >= INT32_MIN && off <= INT32_MAX;
    a.ckpt_base[it.id] = ok ? make_int2((int)off, ntx) : make_int2(0, 0);
  }
  // Row tasks of the items that got checkpoints: (item << 32) | row, the
  // wavefront's rows spread over its lanes (owners by binary search of the
  // prefix).  An item that got no room reserved its tasks all the same:
  // they are written as "no task" (~0) and skipped.
  const uint32_t ex = R - nrow;
  for (uint32_t j = 0; j < rtotal; j += 64) {
    const uint32_t q = j + lane;
    int lo = 0, hi = 63;
#pragma unroll
    for (int st = 0; st < 6; st++) {
      const int mid = (lo + hi) >> 1;
      if (__shfl(R, mid) > q) hi = mid; else lo = mid + 1;
    }
    const uint32_t oex = __shfl(ex, lo);
    const int oby0 = __shfl((int)it.by0, lo);
    const uint32_t oitem = __shfl(item, lo);
    const bool ook = __shfl((int)ok, lo) != 0;
    const u64 slot = rbase + q;
    if (q < rtotal && slot < seg_rows)
      a.ckpt_rows[(u64)seg * seg_rows + slot] =
          ook ? ((u64)oitem << 32) | (uint32_t)(oby0 + (int)(q - oex)) : ~0ull;
  }
}

// (c) scatter triangle ids into the bin lists of the tiles they overlap;
// also resets the merge buffers of split tiles.  kAlloc: also take the
// edge checkpoints of the items (else the scan's launch has taken them).
template <bool kAlloc>
__device__ __forceinline__ void stage_bin(const BatchArgs& a, int vb, int nvb) {
  const int lane = threadIdx.x & 63;
  const u64 np = (u64)a.tile_w * a.tile_h;
  const u64 nheavy = min(a.counters[kHeavy], (u64)a.heavy_cap);
  for (u64 i = (u64)vb * 256 + threadIdx.x; i < nheavy * np; i += (u64)nvb * 256)
    a.heavy_keys[i] = ~0ull;
  const int sw = __ffs(a.tile_w) - 1, sh = __ffs(a.tile_h) - 1;
  const int64_t nw = (int64_t)nvb * 4;
  u64 nitems = 0;
  auto reserve = [&](int key, uint32_t n, uint32_t) -> uint32_t {
    return atomicAdd(&a.tile_cursor[(size_t)key * kCtrStride], n);
  };
  // The next region's count and first 64 items are loaded while the current
  // region is scattered (both loads are independent: every region slot is
  // allocated, slots past the count are read but not used).
  int64_t r = (int64_t)vb * 4 + (threadIdx.x >> 6);
  uint32_t cnt = 0;
  BinItem first = BinItem{0, 0, 0, 0, 0, 0};
  if (r < a.nregions) {
    cnt = a.region_count[r];
    first = a.items[r * kRegion + lane];
  }
  while (r < a.nregions) {
    const int64_t rn = r + nw;
    uint32_t cnt_n = 0;
    BinItem first_n = BinItem{0, 0, 0, 0, 0, 0};
    if (rn < a.nregions) {
      cnt_n = a.region_count[rn];
      first_n = a.items[rn * kRegion + lane];
    }
    nitems += cnt;
    for (uint32_t j = 0; j < cnt; j += 64) {
      const bool act = j + lane < cnt;
      BinItem it = first;
      if (j > 0) it = act ? a.items[r * kRegion + j + lane] : BinItem{0, 0, 0, 0, 0, 0};
      // The job's tile grid, through scalar loads when the wavefront's items
      // share one draw (the common case).
      // (Lane 0 is active in every step: j < cnt.)
      const uint32_t d0 = __builtin_amdgcn_readfirstlane(it.draw);
      const uint32_t dd = act ? it.draw : d0;
      int tbase, tiles_x;
      if (__ballot(dd != d0) == 0ull) {
        const DrawDesc Du = uniform_draw(a, (int)d0);
        tbase = Du.tile_base;
        tiles_x = Du.tiles_x;
      } else {
        tbase = a.draws[dd].tile_base;
        tiles_x = a.draws[dd].tiles_x;
      }
      if (kAlloc) ckpt_alloc(a, act, it, sw, (uint32_t)(r * kRegion + j + lane), vb);
      wave_tile_pairs<true>(a, act, it.id, it.draw, tbase, tiles_x, it.bx0, it.bx1, it.by0,
                            it.by1, sw, sh, reserve);
    }
    r = rn;
    cnt = cnt_n;
    first = first_n;
  }
  __shared__ unsigned long long sh_items;
  if (threadIdx.x == 0) sh_items = 0;
  __syncthreads();
  if (lane == 0 && nitems) atomicAdd(&sh_items, nitems);
  __syncthreads();
  if (threadIdx.x == 0 && vb < kStatSlots) a.block_stats[kStatSlots + vb] = sh_items;
}


// Row pre-pass: one lane per row task of an item with edge checkpoints (a
// triangle at least kCkptMinSpan pixels wide spanning several tile
// columns): walk the row once from the bbox's left column (Renderer.cpp:
// 240-243) and store the two edge values where it enters each later tile
// column.  The raster starts such rows from these checkpoints and shading
// from the one of its pixel's tile column: the same serial chain of
// additions as one walk from the left column, done once per row instead of
// once per (row, tile column).  Tasks sit in kCkptSegs segments; a
// wavefront takes 64 consecutive tasks of one segment.
__device__ __forceinline__ void stage_rows(const BatchArgs& a, int vb, int nvb) {
  const int lane = threadIdx.x & 63;
  const int sw = __ffs(a.tile_w) - 1;
  const u64 seg_rows = (a.ckpt_cap / kCkptSegs) / 2;
  // Tasks per segment (lane s holds segment s) and the largest count.
  const u64 mine = min(a.ckpt_ctr[lane * kCtrStride] >> 40, seg_rows);
  u64 most = mine;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) most = max(most, (u64)__shfl_xor(most, o));
  const u64 nw = (u64)nvb * 4;
  for (u64 u = (u64)vb * 4 + (threadIdx.x >> 6); (u / kCkptSegs) * 64 < most; u += nw) {
    const int seg = (int)(u % kCkptSegs);
    const u64 first = (u / kCkptSegs) * 64;
    const u64 count = __shfl(mine, seg);
    if (first >= count) continue;
    const u64 t = first + lane;
    const u64 task = t < count ? a.ckpt_rows[(u64)seg * seg_rows + t] : ~0ull;
    if (task == ~0ull) continue;
    const BinItem it = a.items[task >> 32];
    const int row = (int)(uint32_t)task;
    const int64_t g = it.id >> 1;
    const DrawDesc& D = a.draws[it.draw];
    TriSetup s;
    tri_setup_sub(a, D, a.jobs[D.job], (int)(g - D.tri_begin), (int)(it.id & 1u), s);
    const int2 cb = a.ckpt_base[it.id];
    float e1, e2;
    edge_start(s.n1x, s.n1y, s.n2x, s.n2y, s.p0x, s.p0y, s.p2x, s.p2y, (float)s.bx0 + 0.5f,
               (float)row + 0.5f, e1, e2);
    int c = s.bx0;
    float2* out = a.ckpt + ((int64_t)cb.x + (int64_t)row * cb.y);
    for (int tx = (s.bx0 >> sw) + 1; tx <= (s.bx1 >> sw); tx++) {
      advance2(e1, s.n1x, e2, s.n2x, (tx << sw) - c);
      c = tx << sw;
      out[tx] = make_float2(e1, e2);
    }
  }
}

// The edge checkpoints of every item (allocation and row tasks), over the
// regions geometry wrote (wavefront w of the grid takes regions w, w + 4 nvb,
// ...).
__device__ __forceinline__ void stage_ckpt(const BatchArgs& a, int vb, int nvb) {
  const int lane = threadIdx.x & 63;
  const int sw = __ffs(a.tile_w) - 1;
  const int64_t nw = (int64_t)nvb * 4;
  int64_t r = (int64_t)vb * 4 + (threadIdx.x >> 6);
  uint32_t cnt = r < a.nregions ? a.region_count[r] : 0u;
  while (r < a.nregions) {
    const int64_t rn = r + nw;
    const uint32_t cnt_n = rn < a.nregions ? a.region_count[rn] : 0u;
    for (uint32_t j = 0; j < cnt; j += 64) {
      const bool act = j + lane < cnt;
      const BinItem it = act ? a.items[r * kRegion + j + lane] : BinItem{0, 0, 0, 0, 0, 0};
      ckpt_alloc(a, act, it, sw, (uint32_t)(r * kRegion + j + lane), vb);
    }
    r = rn;
    cnt = cnt_n;
  }
}

// The scan's workgroups first, then (ckpt_in_scan) the checkpoint
// allocation's, which needs only geometry's items and so runs beside the
// scan: one launch for both.
__global__ __launch_bounds__(kScanBlock) void k_scan(BatchArgs a) {
  static_assert(kScanBlock == 256, "the checkpoint pass runs in workgroups of 256");
  if ((int)blockIdx.x < a.scan_blocks) stage_scan(a);
  else stage_ckpt(a, (int)blockIdx.x - a.scan_blocks, (int)gridDim.x - a.scan_blocks);
}

// Binning that takes the checkpoints (k_rows walks them next), and binning
// followed by the row pre-pass of the row tasks the scan's launch wrote
// (ckpt_in_scan): kernels of their own, each with only its own code.
__global__ __launch_bounds__(256, 7) void k_bin(BatchArgs a) {
  stage_bin<true>(a, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(256, 7) void k_bin_rows(BatchArgs a) {
  stage_bin<false>(a, blockIdx.x, gridDim.x);
  stage_rows(a, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(256) void k_rows(BatchArgs a) {
  stage_rows(a, blockIdx.x, gridDim.x);
}

// Depth test of one pixel into the tile's LDS key buffer.
__device__ inline void raster_pixel(u64* keys, int p, float e1, float e2, float z0, float z1,
                                    float z2, bool persp, uint32_t id) {
  if (!covered(e1, e2)) return;
  float k[3];
  const float z = depth_of(e1, e2, z0, z1, z2, persp, k);
  if (!(z < kClearDepth)) return;  // strict test against the clear value
  const u64 key = ((u64)depth_key(z) << 32) | id;
  if (key < keys[p]) atomicMin(&keys[p], key);
}

// (d): one workgroup per raster work item: first the parts of split tiles
// (listed by the scan), then every tile.  The grid has parts_cap + ntiles
// workgroups; workgroups past the actual number of items exit at once.
// remap: listed tiles in XCD-interleaved order (own launch, where
// consecutive workgroups go to different XCDs), else in list order.
template <int TW, int TH, int NT>
__device__ __forceinline__ void stage_tile(const BatchArgs& a, int64_t b, bool remap) {
  constexpr int NP = TW * TH;
  // Key rows padded by one entry: pixels of one column in different rows
  // fall in different LDS banks.
  constexpr int KS = TW + 1;
  __shared__ u64 keys[TH * KS];
  __shared__ TriSetup setups[NT];
  __shared__ uint32_t setup_id[NT];
  __shared__ int row_start[NT];
  __shared__ int wave_rows[NT / 64];
  __shared__ int sh_last;
  const int tid = threadIdx.x;
  // After a merge-slot or part-list overflow the part list is incomplete:
  // split tiles are then skipped (the frame is re-run with larger buffers).
  const int64_t nparts_all =
      a.counters[kHeavyOverflow] ? 0 : (int64_t)min(a.counters[kParts], (u64)a.parts_cap);
  int tile, part;
  uint32_t cnt, off;  // the tile's bin entries: number and first
  int band_lo = 0, band_hi = TH - 1;  // tile rows of this work item
  if (b < nparts_all) {
    const uint2 pp = a.parts[b];
    tile = (int)pp.x;
    part = (int)pp.y;
    if (pp.y & kBandPart) {  // one row band of the tile, all of its entries
      const int nb = (int)((pp.y >> 8) & 0xFFu), bi = (int)(pp.y & 0xFFu);
      const int bh = (TH + nb - 1) / nb;
      band_lo = bi * bh;
      band_hi = min(TH - 1, band_lo + bh - 1);
      part = 0;
    }
    cnt = (uint32_t)a.tile_count[(size_t)tile * kCtrStride64];
    off = a.tile_off[tile];
  } else {
    // Tiles rastered whole come from the scan's list, with their bin range
    // (one load fewer before the bin entries); tiles without bin entries
    // are not in it (shading writes their background).
    const int64_t t = b - nparts_all;
    const int64_t nlisted = (int64_t)min(a.counters[kListed], (u64)a.ntiles);
    if (t >= nlisted) return;
    const uint4 le = a.tile_list[remap ? xcd_remap(t, nlisted) : t];
    tile = (int)le.x;
    part = 0;
    off = le.y;
    cnt = le.z;
  }
  {
    const uint32_t nparts = cnt <= a.frag_size ? 1u : (cnt + a.frag_size - 1) / a.frag_size;
    if (b >= nparts_all && nparts > 1) return;  // rendered through its parts
    const int j = job_lookup(a.tile_hint, a.job_tile_base, a.njobs, tile);
    const JobDesc& J = a.jobs[j];
    const bool persp = J.persp != 0;
    const int local = tile - J.tile_base;
    const int tyi = local / J.tiles_x, txi = local - tyi * J.tiles_x;
    const int x0 = txi * TW, y0 = tyi * TH;
    for (int p = tid; p < TH * KS; p += NT) keys[p] = ~0ull;
    uint32_t avail = cnt;  // entries actually written (bin-list overflow)
    if ((u64)off >= a.bins_cap) avail = 0;
    else if ((u64)off + cnt > a.bins_cap) avail = (uint32_t)(a.bins_cap - off);
    const uint32_t e_begin = (uint32_t)part * a.frag_size;
    const uint32_t e_end = min(avail, e_begin + a.frag_size);
    __syncthreads();
    for (uint32_t base = e_begin; base < e_end; base += NT) {
      // One lane per entry: setup, and the tile rows it covers.
      int nrows = 0;
      if (base + tid < e_end) {
        const uint2 be = a.bins[off + base + tid];
        const uint32_t id = be.x;
        const int64_t g = id >> 1;
        TriSetup s;
        const uint32_t d0 = __builtin_amdgcn_readfirstlane(be.y);
        if (__ballot(be.y == d0) == __ballot(1)) {
          const DrawDesc Du = uniform_draw(a, (int)d0);
          tri_setup_sub(a, Du, J, (int)(g - Du.tri_begin), (int)(id & 1u), s);
        } else {
          const DrawDesc& D = a.draws[be.y];
          tri_setup_sub(a, D, J, (int)(g - D.tri_begin), (int)(id & 1u), s);
        }
        nrows = max(min(s.by1, y0 + band_hi) - max(s.by0, y0 + band_lo) + 1, 0);
        setups[tid] = s;
        setup_id[tid] = id;
      }
      // Exclusive scan of the row counts over the workgroup.
      int incl = nrows;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if ((tid & 63) >= o) incl += y;
      }
      if ((tid & 63) == 63) wave_rows[tid >> 6] = incl;
      __syncthreads();
      int before = 0, total = 0;
#pragma unroll
      for (int w = 0; w < NT / 64; w++) {
        if (w < (tid >> 6)) before += wave_rows[w];
        total += wave_rows[w];
      }
      row_start[tid] = before + incl - nrows;
      __syncthreads();
      // One lane per (triangle, tile row): start the row at the tile's left
      // edge (exact advance of the serial recurrence), then step across the
      // tile as the reference's column loop does (Renderer.cpp:238-269).
      for (int q = tid; q < total; q += NT) {
        const int i = owner_search<NT>(row_start, q);
        const TriSetup& s = setups[i];
        const int row = max(s.by0, y0 + band_lo) + (q - row_start[i]);
        const int cx0 = max(s.bx0, x0), cx1 = min(s.bx1, x0 + TW - 1);
        const uint32_t id = setup_id[i];
        // A row entering a tile column after its triangle's first starts
        // from the edge checkpoint the row pre-pass left there (the same
        // serial chain from the bbox's left column); without checkpoints
        // (none allocated) it walks from the left column itself.
        float e1 = 0.0f, e2 = 0.0f;
        bool from_ckpt = false;
        if (cx0 > s.bx0 && s.bx1 - s.bx0 + 1 >= kCkptMinSpan) {  // only these have any
          const int2 cb = a.ckpt_base[id];
          if (cb.y) {
            const float2 st = a.ckpt[(int64_t)cb.x + (int64_t)row * cb.y + x0 / TW];
            e1 = st.x;
            e2 = st.y;
            from_ckpt = true;
          }
        }
        if (!from_ckpt) {
          edge_start(s.n1x, s.n1y, s.n2x, s.n2y, s.p0x, s.p0y, s.p2x, s.p2y,
                     (float)s.bx0 + 0.5f, (float)row + 0.5f, e1, e2);
          advance2(e1, s.n1x, e2, s.n2x, cx0 - s.bx0);
        }
        // The triangle's steps and depths in registers for the row: read from
        // the workgroup's setup table once, not again after each key update.
        const float n1x = s.n1x, n2x = s.n2x, z0 = s.z0, z1 = s.z1, z2 = s.z2;
        const int prow = (row - y0) * KS - x0;
        for (int c = cx0; c <= cx1; c++) {
          raster_pixel(keys, prow + c, e1, e2, z0, z1, z2, persp, id);
          e1 = e1 + n1x;
          e2 = e2 + n2x;
        }
      }
      __syncthreads();
    }
    // The job's size and winner buffer in registers: the stores below could
    // alias the job table for all the compiler knows, which would reload
    // them for every pixel.
    const int jw = J.W, jh = J.H;
    uint32_t* const jwin = J.win;
    if (nparts == 1) {
      // (a band writes its own rows only)
      for (int p = tid + band_lo * TW; p < (band_hi + 1) * TW; p += NT) {
        const int c = x0 + (p % TW), row = y0 + (p / TW), lp = (p / TW) * KS + (p % TW);
        if (c < jw && row < jh) {
          const uint32_t pix = (uint32_t)row * (uint32_t)jw + (uint32_t)c;
          *px_at(jwin, 4u * pix) = keys[lp] == ~0ull ? kNoWinner : (uint32_t)keys[lp];
        }
      }
    } else {
      // Split tile: merge into the tile's global key buffer; the last part
      // to finish writes the winners.
      // No device-wide fence here: on this GPU each one writes back and
      // invalidates the XCD's whole L2.  The merges are returning
      // device-scope atomics whose results are consumed before the barrier,
      // so they are performed before the part counter is incremented; the
      // last part reads the merged keys with device-scope atomic loads.
      const uint32_t slot = a.tile_heavy[tile];
      if (slot != kNoSlot) {
        u64* gk = a.heavy_keys + (u64)slot * NP;
        u64 seen = ~0ull;
        for (int p = tid; p < NP; p += NT) {
          const u64 kp = keys[(p / TW) * KS + (p % TW)];
          if (kp != ~0ull)
            seen &= __hip_atomic_fetch_min(&gk[p], kp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (seen == 0ull) sh_last = 0;  // consumes the returned values
        __syncthreads();
        if (tid == 0)
          sh_last = __hip_atomic_fetch_add(&a.tile_done[tile], 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT) == nparts - 1;
        __syncthreads();
        if (sh_last) {
          for (int p = tid; p < NP; p += NT) {
            const int c = x0 + (p % TW), row = y0 + (p / TW);
            const u64 k = __hip_atomic_load(&gk[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (c < jw && row < jh) {
              const uint32_t pix = (uint32_t)row * (uint32_t)jw + (uint32_t)c;
              *px_at(jwin, 4u * pix) = k == ~0ull ? kNoWinner : (uint32_t)k;
            }
          }
        }
      }
    }
  }
}

template <int TW, int TH, int NT>
__global__ __launch_bounds__(NT, 6) void k_tile(BatchArgs a) {
  stage_tile<TW, TH, NT>(a, blockIdx.x, true);
}

// Range image of the pixel from its depth (Renderer.cpp:388-415).
__device__ inline void store_point(const JobDesc& J, uint32_t pix, int c, int row, float depth) {
  const f3 q = range_point(J.P, range_far_depth(J.P), (float)J.W, (float)J.H, row, c, depth);
  auto* o = px_at(J.points, 12u * pix);
  o[0] = q.x;
  o[1] = q.y;
  o[2] = q.z;
}

// Reciprocal length and normalisation with the hardware reciprocal square
// root (v_rsq_f32, about 1 ulp): for colour only.
__device__ inline float fast_rlen3(f3 a) { return __builtin_amdgcn_rsqf(dot3(a, a)); }
__device__ inline f3 fast_norm3(f3 a) { return scale3(fast_rlen3(a), a); }

// Eye-space point on the ray through the centre of pixel (c, row) at eye
// depth ze: the projection (project_vertex) solved for eye x, y, with the
// hardware reciprocal (colour only).
__device__ inline f3 unproject_pixel(const JobDesc& J, int c, int row, float ze) {
  const float* P = J.P;
  const float xn = ((float)c + 0.5f) * (2.0f * __builtin_amdgcn_rcpf((float)J.W)) - 1.0f;
  const float yn = 1.0f - ((float)row + 0.5f) * (2.0f * __builtin_amdgcn_rcpf((float)J.H));
  const float a11 = P[0] - xn * P[12], a12 = P[1] - xn * P[13];
  const float a21 = P[4] - yn * P[12], a22 = P[5] - yn * P[13];
  const float w = P[14] * ze + P[15];
  const float b1 = xn * w - (P[2] * ze + P[3]);
  const float b2 = yn * w - (P[6] * ze + P[7]);
  const float idet = __builtin_amdgcn_rcpf(a11 * a22 - a12 * a21);
  return f3{(b1 * a22 - b2 * a12) * idet, (a11 * b2 - a21 * b1) * idet, ze};
}

// Shading of a covered pixel whose winner `id` belongs to draw D.
__device__ __forceinline__ void shade_lit(const BatchArgs& a, const JobDesc& J, const DrawDesc& D,
                                          uint32_t id, int c, int row, uint32_t pix) {
  const int64_t g = id >> 1;
  const int sub = (int)(id & 1u);
  const int tri = (int)(g - D.tri_begin);
  // All index loads first, then all attribute loads: two dependent levels
  // of memory access instead of four.
  const auto* ip = G(D.idx) + 3 * tri;
  const auto* in = G(D.nidx) + 3 * tri;
  const int sw = __ffs(a.tile_w) - 1;
  const int ctx = c >> sw;
  const int i0 = ip[0], i1 = ip[1], i2 = ip[2];
  // Normal and texcoord indices that share the position index buffer are
  // not loaded again.
  int n0 = i0, n1 = i1, n2 = i2;
  if (D.nidx != D.idx) {
    n0 = in[0];
    n1 = in[1];
    n2 = in[2];
  }
  int t0 = 0, t1 = 0, t2 = 0;
  if (D.tidx == D.idx) {
    t0 = i0;
    t1 = i1;
    t2 = i2;
  } else if (D.tidx) {
    const auto* it = G(D.tidx) + 3 * tri;
    t0 = it[0];
    t1 = it[1];
    t2 = it[2];
  }
  // The corners' projections and eye depths; their eye x, y only for a
  // triangle re-clipped below or of a draw without the near test (whose
  // corners may lie behind the eye).
  const int vi[3] = {i0, i1, i2};
  f3 ec[3];
  ScreenVtx vc[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float4 r = a.vscr[D.vbase + vi[k]];
    vc[k] = ScreenVtx{r.x, r.y, r.z};
    ec[k] = f3{0.0f, 0.0f, r.w};
  }
  const f3 N0 = load3(D.nrm, n0), N1 = load3(D.nrm, n1), N2 = load3(D.nrm, n2);
  f2 U0 = f2{0.0f, 0.0f}, U1 = U0, U2 = U0;  // Renderer.cpp:371-379
  if (D.tidx) {
    const auto* uv = G(D.uv);
    U0 = f2{uv[2 * t0], uv[2 * t0 + 1]};
    U1 = f2{uv[2 * t1], uv[2 * t1 + 1]};
    U2 = f2{uv[2 * t2], uv[2 * t2 + 1]};
  }
  const bool clip = !D.no_clip && (ec[0].z > J.znear || ec[1].z > J.znear || ec[2].z > J.znear);
  const bool eyes = clip || D.no_clip;
  if (eyes) corner_eyes(D, vi, ec);
  // Eye-space vertices as paintMesh computes them (Renderer.cpp:344-349).
  Vtx v0, v1, v2;
  v0.p = ec[0];
  v1.p = ec[1];
  v2.p = ec[2];
  TriSetup s;
  if (clip)
    tri_setup_corners(J, ec, vc, sub, false, s);
  else
    setup_edges((float)J.W, vc[0], vc[1], vc[2], s);
  // Edge values at the pixel, from the same serial walk along the row as the
  // raster's: from the row pre-pass's checkpoint where the row enters the
  // pixel's tile column, or from the bbox's left column.  Only triangles at
  // least kCkptMinSpan pixels wide can have checkpoints (the raster's test):
  // the others read no checkpoint base.
  float e1 = 0.0f, e2 = 0.0f;
  bool from_ckpt = false;
  if (ctx > (s.bx0 >> sw) && s.bx1 - s.bx0 + 1 >= kCkptMinSpan) {
    const int2 cb = a.ckpt_base[id];
    if (cb.y) {
      const float2 st = a.ckpt[(int64_t)cb.x + (int64_t)row * cb.y + ctx];
      e1 = st.x;
      e2 = st.y;
      advance2(e1, s.n1x, e2, s.n2x, c - (ctx << sw));
      from_ckpt = true;
    }
  }
  if (!from_ckpt) {
    edge_start(s.n1x, s.n1y, s.n2x, s.n2y, s.p0x, s.p0y, s.p2x, s.p2y,
               (float)s.bx0 + 0.5f, (float)row + 0.5f, e1, e2);
    advance2(e1, s.n1x, e2, s.n2x, c - s.bx0);
  }
  float k[3];
  const float z = depth_of(e1, e2, s.z0, s.z1, s.z2, J.persp != 0, k);
  if (!J.rgb8 && !J.rgb && !(J.normals && J.save_normals)) {
    if (J.depth) *px_at(J.depth, 4u * pix) = z;
    if (J.winner) *px_at(J.winner, 4u * pix) = D.order_base + 2u * (uint32_t)tri + (uint32_t)sub;
    if (J.points) store_point(J, pix, c, row, z);
    return;
  }

  // Normals and uvs of the triangle, re-clipped for a clipped sub-triangle
  // (Renderer.cpp:131-161).  The reference moves each vertex normal into eye
  // space and interpolates (:347-348, 285-288); the saved normals image keeps
  // that order (bit-exact).  For colour alone the object-space normals are
  // interpolated and moved once: the same affine map of the same weights
  // (they sum to 1, and clipping interpolates linearly), equal up to
  // rounding, with two matrix products fewer per pixel.
  const bool save_n = J.normals && J.save_normals;
  if (save_n) {
    v0.n = xform(D.nm, N0);
    v1.n = xform(D.nm, N1);
    v2.n = xform(D.nm, N2);
  } else {
    v0.n = N0;
    v1.n = N1;
    v2.n = N2;
  }
  v0.uv = U0;
  v1.uv = U1;
  v2.uv = U2;
  if (clip) {
    Vtx a0, a1, a2, b0, b1, b2;
    clip_triangle(J.znear, v0, v1, v2, a0, a1, a2, b0, b1, b2);
    if (sub) { v0 = b0; v1 = b1; v2 = b2; }
    else { v0 = a0; v1 = a1; v2 = a2; }
  }

  f3 color = f3{D.diffuse[0], D.diffuse[1], D.diffuse[2]};
  if (D.tex) {  // :271-275
    const float u = (k[0] * v0.uv.x + k[1] * v1.uv.x) + k[2] * v2.uv.x;
    const float v = (k[0] * v0.uv.y + k[1] * v1.uv.y) + k[2] * v2.uv.y;
    color = texel(D.tex, D.tex_format, D.tex_cols, tex_index(v, D.tex_rows),
                  tex_index(u, D.tex_cols));
  }
  f3 value = f3{D.emissive[0], D.emissive[1], D.emissive[2]};
  if (J.lighting) {  // :279-304, FAST_LIGHT
    // Colour only (the contract is 1 LSB of RGB8): hardware reciprocal
    // square root and exp2/log2 (v_rsq, v_log, v_exp) replace the correctly
    // rounded sqrt, divide and powf; the error is a few ulp of fp32.
    // The pixel's eye position: the corners' eye positions interpolated
    // (:280).  For a triangle in front of the near plane (not re-clipped,
    // near test on) that is the point on the ray through the pixel centre
    // at the interpolated eye depth (whose arithmetic is the reference's),
    // found without the corners' eye x, y.
    const f3 position =
        eyes ? add3(add3(scale3(k[0], v0.p), scale3(k[1], v1.p)), scale3(k[2], v2.p))
             : unproject_pixel(J, c, row, (k[0] * v0.p.z + k[1] * v1.p.z) + k[2] * v2.p.z);
    const f3 L = f3{J.light[0], J.light[1], J.light[2]};
    const f3 lightdir = J.light_is_point ? fast_norm3(sub3(L, position)) : L;
    f3 normal = add3(add3(scale3(k[0], v0.n), scale3(k[1], v1.n)), scale3(k[2], v2.n));
    if (!save_n) normal = xform(D.nm, normal);
    const float inl = fast_rlen3(normal);
    const float diff = amax(0.0f, dot3(normal, lightdir)) * inl + J.ambient;
    value = add3(value, scale3(diff, color));
    if (D.shininess != 0.0f) {
      const f3 view = fast_norm3(position);
      const f3 hv = sub3(lightdir, view);
      const float cosa = amax(dot3(hv, normal), 0.0f) * (fast_rlen3(hv) * inl);
      // powf(cosa, shininess) as exp2(shininess * log2(cosa)): cosa = 0
      // gives 0 (or inf for a negative exponent), as powf does.
      const float sp = __builtin_amdgcn_exp2f(D.shininess * __builtin_amdgcn_logf(cosa));
      value = add3(value, scale3(sp, f3{D.specular[0], D.specular[1], D.specular[2]}));
    }
    if (save_n) {
      auto* o = px_at(J.normals, 12u * pix);
      o[0] = normal.x;
      o[1] = normal.y;
      o[2] = normal.z;
    }
  } else if (save_n) {
    auto* o = px_at(J.normals, 12u * pix);
    o[0] = 0.0f;
    o[1] = 0.0f;
    o[2] = 1.0f;
  }
  if (J.rgb8) {
    auto* o = px_at(J.rgb8, 3u * pix);
    o[0] = quantize(value.x);
    o[1] = quantize(value.y);
    o[2] = quantize(value.z);
  }
  if (J.rgb) {
    auto* o = px_at(J.rgb, 12u * pix);
    o[0] = value.x;
    o[1] = value.y;
    o[2] = value.z;
  }
  if (J.depth) *px_at(J.depth, 4u * pix) = z;
  if (J.winner) *px_at(J.winner, 4u * pix) = D.order_base + 2u * (uint32_t)tri + (uint32_t)sub;
  if (J.points) store_point(J, pix, c, row, z);
}

// A pixel nothing was drawn on: clear(), Renderer.cpp:113-119.
__device__ inline void shade_background(const JobDesc& J, uint32_t pix, int c, int row) {
  if (J.rgb8) {
    auto* o = px_at(J.rgb8, 3u * pix);
    o[0] = quantize(J.bg[0]);
    o[1] = quantize(J.bg[1]);
    o[2] = quantize(J.bg[2]);
  }
  if (J.depth) *px_at(J.depth, 4u * pix) = kClearDepth;
  if (J.winner) *px_at(J.winner, 4u * pix) = kNoWinner;
  if (J.rgb) {
    auto* o = px_at(J.rgb, 12u * pix);
    o[0] = J.bg[0];
    o[1] = J.bg[1];
    o[2] = J.bg[2];
  }
  if (J.normals && J.save_normals) {
    auto* o = px_at(J.normals, 12u * pix);
    o[0] = 0.0f;
    o[1] = 0.0f;
    o[2] = 1.0f;
  }
  if (J.points) store_point(J, pix, c, row, kClearDepth);
}

// Deferred shading of one pixel from its winner (Renderer.cpp:271-305).
__device__ void shade_pixel(const BatchArgs& a, const JobDesc& J,
                            const int64_t* dcache, int ncached, uint32_t id,
                            int c, int row) {
  const uint32_t pix = (uint32_t)row * (uint32_t)J.W + (uint32_t)c;
  if (id == kNoWinner) {
    shade_background(J, pix, c, row);
    return;
  }
  const int d = lookup_draw(dcache, ncached, a, J, (int64_t)(id >> 1));
  // A winner id names a triangle of one of the job's draws; anything else
  // (a winner buffer this batch did not write) is drawn as background.  The
  // check reads the draw's range from the descriptor the shading reads
  // anyway (through scalar loads when the whole wavefront has one draw).
  const int d0 = __builtin_amdgcn_readfirstlane(d);
  if (__ballot(d == d0) == __ballot(1)) {
    const DrawDesc Du = uniform_draw(a, d0);
    const int64_t tri = (int64_t)(id >> 1) - Du.tri_begin;
    if (tri < 0 || tri >= (int64_t)Du.ntri) {
      shade_background(J, pix, c, row);
      return;
    }
    shade_lit(a, J, Du, id, c, row, pix);
  } else {
    const DrawDesc& Dc = a.draws[d];
    const int64_t tri = (int64_t)(id >> 1) - Dc.tri_begin;
    if (tri < 0 || tri >= (int64_t)Dc.ntri) {
      shade_background(J, pix, c, row);
      return;
    }
    shade_lit(a, J, Dc, id, c, row, pix);
  }
}

// Statistics of the frame: sum the per-workgroup counts of the geometry and
// bin kernels (finished before shading starts) into the counters read back
// with the batch.  Run by one shading workgroup; uses one LDS word per wave.
__device__ void reduce_stats(const BatchArgs& a, int tid) {
  __shared__ unsigned long long part[2][kShadeBlock / 64];
  const int ng = a.total_tris ? geom_grid(a) : 0, nb = bin_grid(a);
  unsigned long long c = 0, r = 0;
  for (int i = tid; i < ng; i += kShadeBlock) c += a.block_stats[i];
  for (int i = tid; i < nb; i += kShadeBlock) r += a.block_stats[kStatSlots + i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    c += __shfl_xor(c, o);
    r += __shfl_xor(r, o);
  }
  if ((tid & 63) == 0) {
    part[0][tid >> 6] = c;
    part[1][tid >> 6] = r;
  }
  __syncthreads();
  if (tid == 0) {
    unsigned long long sc = 0, sr = 0;
    for (int w = 0; w < kShadeBlock / 64; w++) {
      sc += part[0][w];
      sr += part[1][w];
    }
    a.counters[kClipped] = sc;
    a.counters[kItems] = sr;
    // Every counter is final here (the kernels that add to them are done):
    // hand them to the host with the batch, no copy of its own.
    if (a.host_counters)
      for (int i = 0; i < kNumCounters; i++)
        a.host_counters[i] = i == kClipped ? sc : (i == kItems ? sr : a.counters[i]);
  }
}

// (e): one lane per pixel; a workgroup shades a kShadeW x kShadeH block of
// one job (16 x 16: each wavefront a 16 x 4 patch, whose pixels share more
// triangles and tiles than a 64-pixel row segment does: 10 % less shading
// time on a batch of 1M-triangle frames).  Held to 7 waves
// per SIMD (72 VGPRs, a few bytes of spill): the kernel is bound by its
// chain of dependent loads, so more waves in flight pay off; 8 spill too much.
// vb: the workgroup's index in the grid of shade_blocks workgroups; remap:
// blocks in XCD-interleaved order (own launch), else b = vb.
__device__ __forceinline__ void stage_shade(const BatchArgs& a, int vb, bool remap) {
  __shared__ int64_t dcache[kMaxLdsDraws];
  const int tid = threadIdx.x;
  const int b = remap ? (int)xcd_remap(vb, a.shade_blocks) : vb;
  const int j = __builtin_amdgcn_readfirstlane(
      job_lookup(a.shade_hint, a.job_shade_base, a.njobs, b));
  const JobDesc J = uniform_job(a, j);
  const int bxn = (J.W + kShadeW - 1) / kShadeW;
  const int local = b - J.shade_base;
  const int byi = local / bxn;
  const int r = tid / kShadeW;
  const int c = (local - byi * bxn) * kShadeW + (tid - r * kShadeW), row = byi * kShadeH + r;
  const bool inside = c < J.W && row < J.H;
  // The winner load is issued first, so that its latency overlaps the
  // draw-table fill below.  A tile without bin entries has no winners
  // written (the raster skips it): its pixels are background, and their
  // winners are not read.
  const uint32_t pix = (uint32_t)row * (uint32_t)J.W + (uint32_t)c;
  uint32_t id = kNoWinner;
  if (inside) {
    const int sw = __ffs(a.tile_w) - 1, sh = __ffs(a.tile_h) - 1;
    const uint32_t flag = a.tile_heavy[J.tile_base + (row >> sh) * J.tiles_x + (c >> sw)];
    if (flag != kEmptyTile) id = *px_at(J.win, 4u * pix);
  }
  // A job of one draw needs no draw table (ncached = -1).
  const bool one_draw = J.draw_end - J.draw_begin == 1;
  const int ncached = one_draw ? -1 : fill_draw_cache(dcache, a, J, tid, kShadeBlock);
  // The raster and scan are done with the tile counters, part counters and
  // look-back words: leave them zero for the next batch on this scratch.
  for (int64_t i = (int64_t)vb * kShadeBlock + tid; i < a.ntiles || i < a.scan_blocks;
       i += (int64_t)a.shade_blocks * kShadeBlock) {
    if (i < a.ntiles) {
      a.tile_count[(size_t)i * kCtrStride64] = 0;
      a.tile_done[i] = 0;
    }
    if (i < a.scan_blocks) a.scan_status[i] = 0;
  }
  if (!one_draw) __syncthreads();
  if (vb == 0) reduce_stats(a, tid);
  // After a bin-list or part-list overflow some tiles were not rastered (their
  // winner ids are stale): nothing is shaded, the batch is re-run.
  const bool overflow = a.counters[kBinOverflow] != 0ull || a.counters[kHeavyOverflow] != 0ull;
  if (inside && !overflow) shade_pixel(a, J, dcache, ncached, id, c, row);
}

__global__ __launch_bounds__(kShadeBlock, 7) void k_shade(BatchArgs a) {
  stage_shade(a, blockIdx.x, true);
}

// ---- the persistent kernel --------------------------------------------------
//
// One launch of a resident grid runs every stage of the batch, each stage as
// the same workgroups (vb, nvb) its own launch would run.  Workgroup w serves
// queue x = w % kPQueues (with round-robin dispatch, one XCD per queue).
// Stage s's workgroups are split into kPQueues contiguous ranges, one per
// queue, and dealt out in items of consecutive workgroups by an atomic head
// per (stage, queue).  A stage that waits for earlier ones starts once each
// of them is complete: a workgroup leaving stage s adds the items it ran to
// done[s][x] (after a release fence); the one completing queue x's count
// bumps fin[s], and the one completing fin[s] sets flag[s][y] for every
// queue y, which queue y's workgroups poll (one load per poll, bounded by a
// wall-clock limit).  Items are only ever waited on by items of later stages,
// and each queue's items are taken in order, so the launch drains while at
// least one workgroup of each queue runs.  A wait that hits the limit sets
// the error word: from then on workgroups skip the work of their items (still
// counting them), the launch drains, and the error reaches the host through
// the batch counters.  The last workgroup to finish clears the control words.

// Polls are relaxed loads: an acquire at agent scope invalidates the L2 of
// the XCD, and workgroups out of items poll while the others still work on
// the stage.  The one acquire fence after a wait orders what follows.
__device__ inline u64 ctl_load(const u64* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Stage workgroups of the batch; the raster's from the scan's counts (the
// grid of its own launch is a bound, whose workgroups past the count exit).
__device__ inline int64_t persist_stage_size(const BatchArgs& a, const PersistPlan& q, int s) {
  if (s != kPTile) return q.nvb[s];
  const int64_t nparts_all =
      a.counters[kHeavyOverflow] ? 0 : (int64_t)min(a.counters[kParts], (u64)a.parts_cap);
  const int64_t nlisted = (int64_t)min(a.counters[kListed], (u64)a.ntiles);
  return min((int64_t)q.nvb[s], nparts_all + nlisted);
}

// Queue x's share of a stage of n workgroups: first workgroup, count, and
// workgroups per item.
struct QueueShare {
  int64_t lo, len, per, items;
};
__device__ inline QueueShare queue_share(const PersistPlan& q, int s, int64_t n, int x, int G) {
  QueueShare r;
  r.lo = n * x / kPQueues;
  r.len = n * (x + 1) / kPQueues - r.lo;
  const int64_t wgs = G / kPQueues + (x < G % kPQueues ? 1 : 0);
  const int64_t want = max((int64_t)1, wgs * (int64_t)q.depth[s]);
  r.per = max((int64_t)1, (r.len + want - 1) / want);
  r.items = (r.len + r.per - 1) / r.per;
  return r;
}

// The persistent kernel's one argument.
struct PersistArgs {
  BatchArgs a;
  PersistPlan q;
};

// The kernel's argument in the kernel-argument segment, where it is the
// first and only explicit argument (offset 0).  The kernel takes the
// segment's address (only a kernel can: in a called function the intrinsic
// yields null) and hands it to the stage functions.  The argument itself is
// never named: taking its address makes the compiler copy it to private
// memory, which is not in the kernel-argument segment.
using KernargPersist = const __attribute__((address_space(4))) PersistArgs;

__device__ __forceinline__ KernargPersist* persist_kernarg() {
  return (KernargPersist*)__builtin_amdgcn_kernarg_segment_ptr();
}

// The argument through a pointer the compiler cannot see through: loads
// after it are not hoisted out of the loops around a stage call, where they
// would be live across the whole stage (and spill).  Loads through it stay
// scalar (constant address space).
__device__ __forceinline__ const PersistArgs& persist_args(KernargPersist* pk) {
  asm volatile("" : "+s"(pk));
  return *(const PersistArgs*)pk;
}

template <int TW>
__device__ __forceinline__ void persist_run(const BatchArgs& a, const PersistPlan& q, int s,
                                            int64_t vb) {
  switch (s) {
    case kPClear: stage_clear(a, (int)vb, q.nvb[s]); break;
    case kPVertex: stage_vertex(a, (int)vb, q.nvb[s]); break;
    case kPGeom: stage_geom(a, (int)vb, q.nvb[s]); break;
    case kPScan: stage_scan(a); break;
    case kPBin: stage_bin<true>(a, (int)vb, q.nvb[s]); break;
    case kPRows: stage_rows(a, (int)vb, q.nvb[s]); break;
    case kPTile: stage_tile<TW, TW, 256>(a, vb, false); break;
    default: stage_shade(a, (int)vb, false); break;
  }
}

// One stage of the persistent kernel for this workgroup: take its queue's
// items until none is left, run their workgroups, then report the items run
// (release, count, completion flags).  A function of its own per stage, not
// inlined: each stage's registers are allocated for it alone (inlined
// together, the stages spill), and the call is made once per stage.
template <int TW, int S>
__device__ __noinline__ void persist_stage(KernargPersist* pk_arg, bool err) {
  // The pointer arrives in vector registers (a function argument): it is
  // the same in every lane, which the compiler is told here.
  const uint64_t pv = (uint64_t)(uintptr_t)pk_arg;
  KernargPersist* const pk = (KernargPersist*)(uintptr_t)(
      ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(pv >> 32)) << 32) |
      (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)pv));
  __shared__ int64_t sh_item[2];
  __shared__ int64_t sh_share[4];  // the stage's lo, len, per, items for this queue
  const int tid = threadIdx.x;
  const int G = gridDim.x;
  const int x = blockIdx.x % kPQueues;
  int64_t n;
  {
    const PersistArgs& pa = persist_args(pk);
    n = persist_stage_size(pa.a, pa.q, S);
  }
  if (n == 0) return;
  const u64 t_begin = persist_args(pk).q.stamps ? wall_clock64() : 0ull;
  __syncthreads();  // every thread has read the previous stage's last item index
  if (tid == 0) {
    const PersistArgs& pa = persist_args(pk);
    const QueueShare sh = queue_share(pa.q, S, n, x, G);
    sh_share[0] = sh.lo;
    sh_share[1] = sh.len;
    sh_share[2] = sh.per;
    sh_share[3] = sh.items;
    sh_item[0] = (int64_t)atomicAdd(pa.q.ctl + (size_t)(kCtlHead + S * kPQueues + x) * kCtlStride,
                                    1ull);
  }
  __syncthreads();
  int64_t it = sh_item[0];
  int par = 0;
  u64 ran = 0;
  while (it < sh_share[3]) {
    u64 next = 0;
    if (tid == 0)  // the next item, in flight during this one
      next = atomicAdd(persist_args(pk).q.ctl + (size_t)(kCtlHead + S * kPQueues + x) * kCtlStride,
                       1ull);
    if (!err) {
      const int64_t lo = sh_share[0], per = sh_share[2];
      const int64_t b0 = lo + it * per;
      const int64_t b1 = min(b0 + per, lo + sh_share[1]);
      for (int64_t vb = b0; vb < b1; vb++) {
        const PersistArgs& pa = persist_args(pk);
        persist_run<TW>(pa.a, pa.q, S, vb);
        __syncthreads();  // LDS of this workgroup is reused by the next
      }
    }
    ran++;
    par ^= 1;
    if (tid == 0) sh_item[par] = (int64_t)next;
    __syncthreads();
    it = sh_item[par];
  }
  if (!ran) return;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (tid == 0) {
    const PersistPlan& q = persist_args(pk).q;
    if (q.stamps) {
      atomicMax(q.stamps + (size_t)(2 * S) * kStampStride, ~t_begin);
      atomicMax(q.stamps + (size_t)(2 * S + 1) * kStampStride, wall_clock64());
    }
    u64* const ctl = q.ctl;
    const u64 old = __hip_atomic_fetch_add(ctl + (size_t)(kCtlDone + S * kPQueues + x) * kCtlStride,
                                           ran, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (old + ran == (u64)sh_share[3]) {
      int nonempty = 0;
      for (int y = 0; y < kPQueues; y++) nonempty += queue_share(q, S, n, y, G).items > 0;
      const u64 f = __hip_atomic_fetch_add(ctl + (size_t)(kCtlFin + S) * kCtlStride, 1ull,
                                           __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      if (f + 1 == (u64)nonempty)
        for (int y = 0; y < kPQueues; y++)
          __hip_atomic_store(ctl + (size_t)(kCtlFlag + S * kPQueues + y) * kCtlStride, 1ull,
                             __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

template <int TW>
__global__ __launch_bounds__(256, 6) void k_persist(PersistArgs) {
  KernargPersist* const pk = persist_kernarg();
  __shared__ int sh_err;
  const int tid = threadIdx.x;
  bool err = false;
  int confirmed = -1;  // stages up to this one are known complete
  for (int s = 0; s < kPStages; s++) {
    const PersistPlan& q = persist_args(pk).q;
    const BatchArgs& a = persist_args(pk).a;
    const int x = blockIdx.x % kPQueues;
    u64* const ctl = q.ctl;
    auto word = [&](int i) -> u64* { return ctl + (size_t)i * kCtlStride; };
    if (q.nvb[s] == 0) continue;
    if (q.wait[s]) {
      // Wait for every earlier stage that has items (thread 0 polls).  A
      // limit of 0 gives up before polling (the host's test of this path).
      if (tid == 0) {
        if (q.time_limit == 0 && !err) {
          __hip_atomic_store(word(kCtlErr), 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          err = true;
        }
        for (int t = confirmed + 1; t < s && !err; t++) {
          if (q.nvb[t] == 0 || persist_stage_size(a, q, t) == 0) continue;
          const u64 t0 = wall_clock64();
          while (ctl_load(word(kCtlFlag + t * kPQueues + x)) == 0ull) {
            if (ctl_load(word(kCtlErr)) != 0ull) {
              err = true;
              break;
            }
            if (wall_clock64() - t0 > q.time_limit) {
              __hip_atomic_store(word(kCtlErr), 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              err = true;
              break;
            }
            __builtin_amdgcn_s_sleep(8);
          }
        }
        sh_err = err;
      }
      __syncthreads();
      err = sh_err != 0;
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      confirmed = s - 1;
    }
    switch (s) {
      case kPClear: persist_stage<TW, kPClear>(pk, err); break;
      case kPVertex: persist_stage<TW, kPVertex>(pk, err); break;
      case kPGeom: persist_stage<TW, kPGeom>(pk, err); break;
      case kPScan: persist_stage<TW, kPScan>(pk, err); break;
      case kPBin: persist_stage<TW, kPBin>(pk, err); break;
      case kPRows: persist_stage<TW, kPRows>(pk, err); break;
      case kPTile: persist_stage<TW, kPTile>(pk, err); break;
      default: persist_stage<TW, kPShade>(pk, err); break;
    }
  }
  if (tid == 0) {
    u64* const ctl = persist_args(pk).q.ctl;
    const u64 old = __hip_atomic_fetch_add(ctl + (size_t)kCtlExit * kCtlStride, 1ull,
                                           __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (old + 1 == (u64)gridDim.x) {  // every other workgroup is done with the words
      if (ctl_load(ctl + (size_t)kCtlErr * kCtlStride))
        atomicOr(&persist_args(pk).a.counters[kError], 1ull);
      for (int i = 0; i < kCtlWords; i++)
        __hip_atomic_store(ctl + (size_t)i * kCtlStride, 0ull, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// ---- scratch of the entry point that allocates nothing ----------------------
//
// mr_render_noalloc takes every scratch buffer of a frame (items, bins,
// checkpoints, tile counters, vertex records, winner ids, tables) from this
// one array instead of allocating.  Each device the library runs on holds its
// own copy for the life of the process.  Host code never passes `scratch`
// itself to a kernel or copy: it uses the device address that
// hipGetSymbolAddress returns for the current device, kept per device below.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return -1;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return 0;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != device) {
    (void)hipGetLastError();
    return -1;
  }
  void* ptr = nullptr;
  if (hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
    (void)hipGetLastError();
    return -1;
  }
  g_scratch_ptr[device] = ptr;
  return 0;
}

void* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

// ---- launchers -------------------------------------------------------------

static int clear_grid(const BatchArgs& a) {
  const int n = std::max(std::max(a.scan_blocks, a.ntiles), (int)kNumCounters);
  return std::min((n + 255) / 256, 4096);
}

static int vertex_grid(const BatchArgs& a) {
  if (a.total_verts == 0) return 0;
  const int64_t blocks = (a.total_verts + kVertexStep - 1) / kVertexStep;
  return (int)std::min<int64_t>(blocks, (int64_t)a.num_cus * 8);
}

hipError_t launch_clear(const BatchArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_clear, dim3(clear_grid(a)), dim3(256), 0, s, a);
  return hipGetLastError();
}

int geom_blocks_per_cu() {
  static const int n = [] {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_geom, 256, 0) != hipSuccess || b < 1) {
      (void)hipGetLastError();
      b = 8;
    }
    return b;
  }();
  return n;
}

hipError_t launch_vertex(const BatchArgs& a, hipStream_t s) {
  if (a.total_verts == 0) return hipSuccess;
  hipLaunchKernelGGL(k_vertex, dim3(vertex_grid(a)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_geometry(const BatchArgs& a, hipStream_t s) {
  if (a.total_tris == 0) return hipSuccess;
  hipLaunchKernelGGL(k_geom, dim3(geom_grid(a)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_scan(const BatchArgs& a, hipStream_t s) {
  const int32_t ckpt_blocks = a.ckpt_in_scan ? bin_grid(a) : 0;
  hipLaunchKernelGGL(k_scan, dim3(a.scan_blocks + ckpt_blocks), dim3(kScanBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_bin(const BatchArgs& a, hipStream_t s) {
  if (a.ckpt_in_scan)
    hipLaunchKernelGGL(k_bin_rows, dim3(bin_grid(a)), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_bin, dim3(bin_grid(a)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_rows(const BatchArgs& a, hipStream_t s) {
  if (a.ckpt_in_scan) return hipSuccess;
  hipLaunchKernelGGL(k_rows, dim3(bin_grid(a)), dim3(256), 0, s, a);
  return hipGetLastError();
}


bool tile_supported(int tw, int th) {
  return (tw == 8 && th == 8) || (tw == 16 && th == 16) || (tw == 32 && th == 32);
}

int default_raster_threads(int tw) { return tw == 8 ? 64 : 256; }

bool raster_threads_supported(int tw, int nt) {
  switch (tw) {
    case 8: return nt == 64 || nt == 128;
    case 16:
    case 32: return nt == 64 || nt == 128 || nt == 256 || nt == 512;
    default: return false;
  }
}

#define MR_TILE_CASE(TW, NT)                                                         \
  if (a.tile_w == TW && a.raster_threads == NT) {                                     \
    hipLaunchKernelGGL((k_tile<TW, TW, NT>), dim3(grid), dim3(NT), 0, s, a);          \
    return hipGetLastError();                                                         \
  }

hipError_t launch_tile(const BatchArgs& a, hipStream_t s) {
  const int grid = a.ntiles + (int)a.parts_cap;
  if (a.tile_w != a.tile_h) return hipErrorInvalidValue;
  MR_TILE_CASE(8, 64)
  MR_TILE_CASE(8, 128)
  MR_TILE_CASE(16, 64)
  MR_TILE_CASE(16, 128)
  MR_TILE_CASE(16, 256)
  MR_TILE_CASE(16, 512)
  MR_TILE_CASE(32, 64)
  MR_TILE_CASE(32, 128)
  MR_TILE_CASE(32, 256)
  MR_TILE_CASE(32, 512)
  return hipErrorInvalidValue;
}
#undef MR_TILE_CASE

hipError_t launch_shade(const BatchArgs& a, hipStream_t s) {
  const int grid = a.shade_blocks;
  hipLaunchKernelGGL(k_shade, dim3(grid), dim3(kShadeBlock), 0, s, a);
  return hipGetLastError();
}

// 16 x 16 tiles only: with 32 x 32 tiles the kernel's registers leave two
// workgroups per CU.
bool persistent_supported(int tw, int th) { return tw == 16 && th == 16; }

// Resident workgroups per CU of the persistent kernel, and the wall clock's
// ticks per second (device 0's figures stand for every device: one GPU model).
template <int TW>
static int persist_blocks_per_cu() {
  static const int n = [] {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_persist<TW>, 256, 0) != hipSuccess ||
        b < 1) {
      (void)hipGetLastError();
      b = 1;
    }
    return b;
  }();
  return n;
}

unsigned long long wall_clock_hz() {
  static const unsigned long long hz = [] {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
        khz <= 0) {
      (void)hipGetLastError();
      khz = 100000;  // 100 MHz
    }
    return (unsigned long long)khz * 1000ull;
  }();
  return hz;
}

hipError_t launch_persistent(const BatchArgs& a, bool clear, unsigned long long* ctl,
                             unsigned long long* stamps, uint64_t wait_us, hipStream_t s) {
  if (!persistent_supported(a.tile_w, a.tile_h) || !ctl) return hipErrorInvalidValue;
  PersistPlan q;
  memset(&q, 0, sizeof(q));
  q.nvb[kPClear] = clear ? clear_grid(a) : 0;
  q.nvb[kPVertex] = vertex_grid(a);
  q.nvb[kPGeom] = a.total_tris ? geom_grid(a) : 0;
  q.nvb[kPScan] = a.scan_blocks;
  q.nvb[kPBin] = bin_grid(a);
  q.nvb[kPRows] = bin_grid(a);
  q.nvb[kPTile] = (int32_t)std::min<int64_t>((int64_t)a.ntiles + (int64_t)a.parts_cap, INT32_MAX);
  q.nvb[kPShade] = a.shade_blocks;
  // Items per queue workgroup: one for the even stages, several for the
  // raster and shading, whose workgroups differ in cost; the scan's
  // workgroups one at a time (each waits on its predecessors).
  const int32_t depth[kPStages] = {1, 1, 1, 1 << 30, 1, 1, 8, 4};
  const int32_t wait[kPStages] = {0, 0, 1, 1, 1, 1, 1, 1};
  for (int i = 0; i < kPStages; i++) {
    q.depth[i] = depth[i];
    q.wait[i] = wait[i];
  }
  q.ctl = ctl;
  q.stamps = stamps;
  q.time_limit = wait_us * (wall_clock_hz() / 1000000ull);
  const int grid = std::max(persist_blocks_per_cu<16>() * std::max(a.num_cus, 1), kPQueues);
  PersistArgs pa;
  pa.a = a;
  pa.q = q;
  hipLaunchKernelGGL(k_persist<16>, dim3(grid), dim3(256), 0, s, pa);
  return hipGetLastError();
}

}  // namespace mr
