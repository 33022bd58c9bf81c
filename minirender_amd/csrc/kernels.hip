// kernels.hip -- the rasterization pipeline of minirender's Renderer::render()
// (src/Renderer.cpp:311-381) as HIP kernels for gfx950 (wave64).
//
//   k_clear     counters, look-back words and tile counts of the batch
//   k_geom      (a)+(b) vertex transform, near clip, projection, setup and
//               cull of every input triangle; order-preserving compaction of
//               the surviving (sub-)triangles (ballot + wave prefix + block
//               scan + decoupled look-back), then per-tile counts
//   k_scan      (c) exclusive scan of the tile counts (decoupled look-back)
//   k_bin       (c) scatter of record ids into the tiles' bin lists
//   k_tile      (d)+(e) per-tile raster with the tile's 64-bit
//               (depth key, record id) buffer in LDS and ds_min_u64, then
//               deferred Blinn-Phong + texture of each pixel's winner
//
// Record ids increase with draw order, so the minimum of (depth key, record
// id) is the first-drawn triangle of minimum depth -- the winner of the
// reference's sequential strict `z < depth` test (Renderer.cpp:267).
#include <hip/hip_runtime.h>

#include "common.h"
#include "exact.h"
#include "pipeline.h"

namespace mr {

using u64 = unsigned long long;

constexpr u64 kFlagAgg = 1ull << 62;
constexpr u64 kFlagPre = 2ull << 62;
constexpr u64 kValMask = (1ull << 62) - 1;
constexpr int kSmallArea = 24;  // pixels of a tile walked by one lane

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

__global__ void k_clear(BatchArgs a) {
  const int n = max(max(a.geom_chunks, a.scan_blocks),
                    max(a.ntiles, (int)kNumCounters));
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    if (i < kNumCounters) a.counters[i] = 0;
    if (i < a.geom_chunks) a.geom_status[i] = 0;
    if (i < a.scan_blocks) a.scan_status[i] = 0;
    if (i < a.ntiles) a.tile_count[i] = 0;
  }
}

__device__ inline int find_draw(const DrawDesc* draws, int n, int64_t g) {
  int lo = 0, hi = n - 1;  // last draw with tri_begin <= g
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (draws[mid].tri_begin <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ inline int find_job(const JobDesc* jobs, int n, int tile) {
  int lo = 0, hi = n - 1;  // last job with tile_base <= tile
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].tile_base <= tile) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ inline f3 load3(const float* p, int i) {
  return f3{p[3 * i], p[3 * i + 1], p[3 * i + 2]};
}

// (a)+(b): one lane per input triangle (Renderer.cpp:340-380, 163-236).
__global__ __launch_bounds__(256) void k_geom(BatchArgs a) {
  __shared__ int sh_chunk;
  __shared__ uint32_t sh_wave[4];
  __shared__ u64 sh_excl;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int sw = __ffs(a.tile_w) - 1, sh = __ffs(a.tile_h) - 1;
  for (;;) {
    if (tid == 0) sh_chunk = (int)atomicAdd(&a.counters[kGeomTicket], 1ull);
    __syncthreads();
    const int chunk = sh_chunk;
    if (chunk >= a.geom_chunks) break;
    const int64_t g = (int64_t)chunk * kGeomChunk + tid;
    TriRec r0, r1;
    int nrec = 0;
    bool clipped = false;
    const JobDesc* J = nullptr;
    if (g < a.total_tris) {
      const int d = find_draw(a.draws, a.ndraws, g);
      const DrawDesc& D = a.draws[d];
      J = &a.jobs[D.job];
      const int t = (int)(g - D.tri_begin);
      const f3 e0 = xform(D.mv, load3(D.pos, D.idx[3 * t]));
      const f3 e1 = xform(D.mv, load3(D.pos, D.idx[3 * t + 1]));
      const f3 e2 = xform(D.mv, load3(D.pos, D.idx[3 * t + 2]));
      const float zn = J->znear;
      const bool persp = J->persp != 0;
      const float w = (float)J->W, h = (float)J->H;
      const bool b0 = e0.z > zn, b1 = e1.z > zn, b2 = e2.z > zn;
      if (b0 || b1 || b2) {  // near test, Renderer.cpp:169-177
        if (!(b0 && b1 && b2)) {
          clipped = true;
          Vtx v0, v1, v2, out[6];
          v0.p = e0; v1.p = e1; v2.p = e2;
          v0.n = v1.n = v2.n = f3{0.0f, 0.0f, 0.0f};
          v0.uv = v1.uv = v2.uv = f2{0.0f, 0.0f};
          const int ns = clip_triangle(zn, v0, v1, v2, out);
#pragma unroll
          for (int s = 0; s < 2; s++) {
            TriRec tr;
            if (s < ns && setup_triangle(J->P, persp, w, h, out[3 * s].p,
                                         out[3 * s + 1].p, out[3 * s + 2].p, tr)) {
              tr.tri_sub = ((uint32_t)t << 2) | 2u | (uint32_t)s;
              tr.draw = (uint32_t)d;
              tr.job = (uint32_t)D.job;
              if (nrec == 0) r0 = tr; else r1 = tr;
              ++nrec;
            }
          }
        }
      } else if (setup_triangle(J->P, persp, w, h, e0, e1, e2, r0)) {
        r0.tri_sub = (uint32_t)t << 2;
        r0.draw = (uint32_t)d;
        r0.job = (uint32_t)D.job;
        nrec = 1;
      }
    }
    // Order-preserving compaction: wave ballots, block scan, look-back.
    const u64 bal1 = __ballot(nrec >= 1), bal2 = __ballot(nrec == 2);
    const u64 lt = (1ull << lane) - 1ull;
    const uint32_t wpre = __popcll(bal1 & lt) + __popcll(bal2 & lt);
    const u64 balc = __ballot(clipped);
    if (lane == 0) {
      sh_wave[wave] = __popcll(bal1) + __popcll(bal2);
      if (balc) atomicAdd(&a.counters[kClipped], (u64)__popcll(balc));
    }
    __syncthreads();
    uint32_t bpre = 0, btot = 0;
#pragma unroll
    for (int w = 0; w < 4; w++) {
      const uint32_t v = sh_wave[w];
      if (w < wave) bpre += v;
      btot += v;
    }
    if (tid == 0) {
      const u64 excl = lookback(a.geom_status, chunk, btot, a.counters);
      sh_excl = excl;
      if (chunk == a.geom_chunks - 1) a.counters[kRecords] = excl + btot;
    }
    __syncthreads();
    const u64 base = sh_excl + bpre + wpre;
    if (nrec >= 1) a.recs[base] = r0;
    if (nrec == 2) a.recs[base + 1] = r1;
    // Per-tile counts of the records just written.
#pragma unroll
    for (int s = 0; s < 2; s++) {
      if (s < nrec) {
        const TriRec& r = s == 0 ? r0 : r1;
        uint32_t* tc = a.tile_count + J->tile_base;
        const int tx0 = r.bx0 >> sw, tx1 = r.bx1 >> sw;
        const int ty0 = r.by0 >> sh, ty1 = r.by1 >> sh;
        for (int ty = ty0; ty <= ty1; ty++)
          for (int tx = tx0; tx <= tx1; tx++)
            atomicAdd(&tc[ty * J->tiles_x + tx], 1u);
      }
    }
    __syncthreads();  // sh_chunk / sh_wave are rewritten next iteration
  }
}

// (c) exclusive scan of tile counts -> tile offsets and cursors.
__global__ __launch_bounds__(kScanBlock) void k_scan(BatchArgs a) {
  __shared__ int sh_blk;
  __shared__ u64 sh_w[kScanBlock / 64];
  __shared__ u64 sh_excl;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) sh_blk = (int)atomicAdd(&a.counters[kScanTicket], 1ull);
  __syncthreads();
  const int blk = sh_blk;
  const int base = blk * (kScanBlock * kScanItems) + tid * kScanItems;
  uint32_t v[kScanItems];
  u64 sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    v[i] = (base + i < a.ntiles) ? a.tile_count[base + i] : 0u;
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
    if (blk == a.scan_blocks - 1) a.counters[kBinEntries] = excl + btot;
  }
  __syncthreads();
  u64 run = sh_excl + wpre + (x - sum);
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    if (base + i < a.ntiles) {
      const uint32_t o = run > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)run;
      a.tile_off[base + i] = o;
      a.tile_cursor[base + i] = o;
    }
    run += v[i];
  }
}

// (c) scatter record ids into the bin lists of the tiles they overlap.
__global__ __launch_bounds__(256) void k_bin(BatchArgs a) {
  const u64 nrec = a.counters[kRecords];
  const int sw = __ffs(a.tile_w) - 1, sh = __ffs(a.tile_h) - 1;
  for (u64 i = (u64)blockIdx.x * 256 + threadIdx.x; i < nrec;
       i += (u64)gridDim.x * 256) {
    const uint4 q = reinterpret_cast<const uint4*>(a.recs + i)[3];
    const int bx0 = q.x & 0xFFFFu, bx1 = q.x >> 16;
    const int by0 = q.y & 0xFFFFu, by1 = q.y >> 16;
    const JobDesc& J = a.jobs[q.w];
    uint32_t* cur = a.tile_cursor + J.tile_base;
    for (int ty = by0 >> sh; ty <= (by1 >> sh); ty++)
      for (int tx = bx0 >> sw; tx <= (bx1 >> sw); tx++) {
        const uint32_t pos = atomicAdd(&cur[ty * J.tiles_x + tx], 1u);
        if ((u64)pos < a.bins_cap) a.bins[pos] = (uint32_t)i;
        else atomicOr(&a.counters[kBinOverflow], 1ull);
      }
  }
}

// Depth test of one covered-or-not pixel into the tile's LDS key buffer.
__device__ inline void raster_pixel(u64* keys, int p, float e1, float e2,
                                    float z0, float z1, float z2, bool persp,
                                    uint32_t rid) {
  if (!covered(e1, e2)) return;
  float k[3];
  const float z = depth_of(e1, e2, z0, z1, z2, persp, k);
  if (!(z < kClearDepth)) return;  // strict test against the clear value
  const u64 key = ((u64)depth_key(z) << 32) | rid;
  if (key < keys[p]) atomicMin(&keys[p], key);
}

// Deferred shading of one pixel from its winning record (Renderer.cpp:271-305).
__device__ void shade_pixel(const BatchArgs& a, const JobDesc& J, u64 key,
                            int c, int row) {
  const size_t pix = (size_t)row * (size_t)J.W + (size_t)c;
  if (key == ~0ull) {  // clear(), Renderer.cpp:113-119
    if (J.rgb8) {
      J.rgb8[pix * 3 + 0] = quantize(J.bg[0]);
      J.rgb8[pix * 3 + 1] = quantize(J.bg[1]);
      J.rgb8[pix * 3 + 2] = quantize(J.bg[2]);
    }
    if (J.depth) J.depth[pix] = kClearDepth;
    if (J.winner) J.winner[pix] = kNoWinner;
    if (J.rgb) {
      J.rgb[pix * 3 + 0] = J.bg[0];
      J.rgb[pix * 3 + 1] = J.bg[1];
      J.rgb[pix * 3 + 2] = J.bg[2];
    }
    if (J.normals && J.save_normals) {
      J.normals[pix * 3 + 0] = 0.0f;
      J.normals[pix * 3 + 1] = 0.0f;
      J.normals[pix * 3 + 2] = 1.0f;
    }
    return;
  }
  const uint32_t rid = (uint32_t)key;
  const TriRec r = a.recs[rid];
  const DrawDesc& D = a.draws[r.draw];
  float e1, e2;
  edge_start(r.n1x, r.n1y, r.n2x, r.n2y, r.p0x, r.p0y, r.p2x, r.p2y,
             (float)r.bx0 + 0.5f, (float)row + 0.5f, e1, e2);
  const int n = c - (int)r.bx0;
  e1 = advance(e1, r.n1x, n);
  e2 = advance(e2, r.n2x, n);
  float k[3];
  const float z = depth_of(e1, e2, r.z0, r.z1, r.z2, J.persp != 0, k);

  // The triangle's eye-space vertices, recomputed as paintMesh does
  // (Renderer.cpp:344-379), and re-clipped for a clipped sub-triangle.
  const int tri = (int)(r.tri_sub >> 2);
  Vtx v0, v1, v2;
  {
    const int* ip = D.idx + 3 * tri;
    const int* in = D.nidx + 3 * tri;
    v0.p = xform(D.mv, load3(D.pos, ip[0]));
    v1.p = xform(D.mv, load3(D.pos, ip[1]));
    v2.p = xform(D.mv, load3(D.pos, ip[2]));
    v0.n = xform(D.nm, load3(D.nrm, in[0]));
    v1.n = xform(D.nm, load3(D.nrm, in[1]));
    v2.n = xform(D.nm, load3(D.nrm, in[2]));
    if (D.tidx) {
      const int* it = D.tidx + 3 * tri;
      v0.uv = f2{D.uv[2 * it[0]], D.uv[2 * it[0] + 1]};
      v1.uv = f2{D.uv[2 * it[1]], D.uv[2 * it[1] + 1]};
      v2.uv = f2{D.uv[2 * it[2]], D.uv[2 * it[2] + 1]};
    } else {
      v0.uv = v1.uv = v2.uv = f2{0.0f, 0.0f};
    }
  }
  if (r.tri_sub & 2u) {
    Vtx out[6];
    clip_triangle(J.znear, v0, v1, v2, out);
    if (r.tri_sub & 1u) { v0 = out[3]; v1 = out[4]; v2 = out[5]; }
    else { v0 = out[0]; v1 = out[1]; v2 = out[2]; }
  }

  f3 color = f3{D.diffuse[0], D.diffuse[1], D.diffuse[2]};
  if (D.tex) {  // :271-275
    const float u = (k[0] * v0.uv.x + k[1] * v1.uv.x) + k[2] * v2.uv.x;
    const float v = (k[0] * v0.uv.y + k[1] * v1.uv.y) + k[2] * v2.uv.y;
    color = texel(D.tex, D.tex_format, D.tex_cols, tex_index(v, D.tex_rows),
                  tex_index(u, D.tex_cols));
  }
  f3 value = f3{D.emissive[0], D.emissive[1], D.emissive[2]};
  f3 normal = f3{0.0f, 0.0f, 1.0f};
  if (J.lighting) {  // :279-304, FAST_LIGHT
    const f3 position = add3(add3(scale3(k[0], v0.p), scale3(k[1], v1.p)),
                             scale3(k[2], v2.p));
    const f3 L = f3{J.light[0], J.light[1], J.light[2]};
    const f3 lightdir = J.light_is_point ? norm3(sub3(L, position)) : L;
    normal = add3(add3(scale3(k[0], v0.n), scale3(k[1], v1.n)), scale3(k[2], v2.n));
    const float diff = amax(0.0f, dot3(normal, lightdir)) / len3(normal) + J.ambient;
    value = add3(value, scale3(diff, color));
    if (D.shininess != 0.0f) {
      const f3 view = norm3(position);
      const f3 hv = sub3(lightdir, view);
      const float sp = powf(amax(dot3(hv, normal), 0.0f) / (len3(hv) * len3(normal)),
                            D.shininess);
      value = add3(value, scale3(sp, f3{D.specular[0], D.specular[1], D.specular[2]}));
    }
    if (J.normals && J.save_normals) {
      J.normals[pix * 3 + 0] = normal.x;
      J.normals[pix * 3 + 1] = normal.y;
      J.normals[pix * 3 + 2] = normal.z;
    }
  } else if (J.normals && J.save_normals) {
    J.normals[pix * 3 + 0] = 0.0f;
    J.normals[pix * 3 + 1] = 0.0f;
    J.normals[pix * 3 + 2] = 1.0f;
  }
  if (J.rgb8) {
    J.rgb8[pix * 3 + 0] = quantize(value.x);
    J.rgb8[pix * 3 + 1] = quantize(value.y);
    J.rgb8[pix * 3 + 2] = quantize(value.z);
  }
  if (J.depth) J.depth[pix] = z;
  if (J.winner) J.winner[pix] = D.order_base + 2u * (uint32_t)tri + (r.tri_sub & 1u);
  if (J.rgb) {
    J.rgb[pix * 3 + 0] = value.x;
    J.rgb[pix * 3 + 1] = value.y;
    J.rgb[pix * 3 + 2] = value.z;
  }
}

// (d)+(e): one workgroup per tile, persistent over a tile ticket.
template <int TW, int TH, int NT>
__global__ __launch_bounds__(NT) void k_tile(BatchArgs a) {
  constexpr int NP = TW * TH;
  __shared__ u64 keys[NP];
  __shared__ TriRec big[NT];
  __shared__ uint32_t bigid[NT];
  __shared__ int nbig;
  __shared__ int sh_tile;
  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0) sh_tile = (int)atomicAdd(&a.counters[kTileTicket], 1ull);
    __syncthreads();
    const int tile = sh_tile;
    if (tile >= a.ntiles) break;
    const JobDesc& J = a.jobs[find_job(a.jobs, a.njobs, tile)];
    const bool persp = J.persp != 0;
    const int local = tile - J.tile_base;
    const int tyi = local / J.tiles_x, txi = local - tyi * J.tiles_x;
    const int x0 = txi * TW, y0 = tyi * TH;
    for (int p = tid; p < NP; p += NT) keys[p] = ~0ull;
    const uint32_t off = a.tile_off[tile];
    uint32_t cnt = a.tile_count[tile];
    if ((u64)off >= a.bins_cap) cnt = 0;
    else if ((u64)off + cnt > a.bins_cap) cnt = (uint32_t)(a.bins_cap - off);
    __syncthreads();
    for (uint32_t base = 0; base < cnt; base += NT) {
      if (tid == 0) nbig = 0;
      __syncthreads();
      if (base + tid < cnt) {
        const uint32_t rid = a.bins[off + base + tid];
        const TriRec r = a.recs[rid];
        const int cx0 = max((int)r.bx0, x0), cx1 = min((int)r.bx1, x0 + TW - 1);
        const int cy0 = max((int)r.by0, y0), cy1 = min((int)r.by1, y0 + TH - 1);
        if ((cx1 - cx0 + 1) * (cy1 - cy0 + 1) <= kSmallArea) {
          // One lane walks the triangle's rows inside the tile, as the
          // reference's loops do (Renderer.cpp:238-269).
          const float x0f = (float)r.bx0 + 0.5f;
          const int lead = cx0 - (int)r.bx0;
          for (int row = cy0; row <= cy1; row++) {
            float e1, e2;
            edge_start(r.n1x, r.n1y, r.n2x, r.n2y, r.p0x, r.p0y, r.p2x, r.p2y,
                       x0f, (float)row + 0.5f, e1, e2);
            e1 = advance(e1, r.n1x, lead);
            e2 = advance(e2, r.n2x, lead);
            const int prow = (row - y0) * TW - x0;
            for (int c = cx0; c <= cx1; c++) {
              raster_pixel(keys, prow + c, e1, e2, r.z0, r.z1, r.z2, persp, rid);
              e1 = e1 + r.n1x;
              e2 = e2 + r.n2x;
            }
          }
        } else {
          const int s = atomicAdd(&nbig, 1);
          big[s] = r;
          bigid[s] = rid;
        }
      }
      __syncthreads();
      const int nb = nbig;
      for (int b = 0; b < nb; b++) {  // large triangles: one lane per pixel
        const TriRec& r = big[b];
        for (int p = tid; p < NP; p += NT) {
          const int c = x0 + (p % TW), row = y0 + (p / TW);
          if (c < (int)r.bx0 || c > (int)r.bx1 || row < (int)r.by0 || row > (int)r.by1)
            continue;
          float e1, e2;
          edge_start(r.n1x, r.n1y, r.n2x, r.n2y, r.p0x, r.p0y, r.p2x, r.p2y,
                     (float)r.bx0 + 0.5f, (float)row + 0.5f, e1, e2);
          const int n = c - (int)r.bx0;
          e1 = advance(e1, r.n1x, n);
          e2 = advance(e2, r.n2x, n);
          raster_pixel(keys, p, e1, e2, r.z0, r.z1, r.z2, persp, bigid[b]);
        }
      }
      __syncthreads();
    }
    for (int p = tid; p < NP; p += NT) {
      const int c = x0 + (p % TW), row = y0 + (p / TW);
      if (c < J.W && row < J.H) shade_pixel(a, J, keys[p], c, row);
    }
    __syncthreads();
  }
}

// ---- launchers -------------------------------------------------------------

hipError_t launch_clear(const BatchArgs& a, hipStream_t s) {
  const int n = max(max(a.geom_chunks, a.scan_blocks), max(a.ntiles, (int)kNumCounters));
  const int grid = min((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_clear, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_geometry(const BatchArgs& a, hipStream_t s) {
  if (a.geom_chunks == 0) return hipSuccess;
  const int grid = min(a.geom_chunks, a.num_cus * 8);
  hipLaunchKernelGGL(k_geom, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_scan(const BatchArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_scan, dim3(a.scan_blocks), dim3(kScanBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_bin(const BatchArgs& a, hipStream_t s) {
  const int64_t maxrec = 2 * a.total_tris;
  if (maxrec == 0) return hipSuccess;
  const int64_t blocks = (maxrec + 255) / 256;
  const int grid = (int)(blocks < (int64_t)a.num_cus * 16 ? blocks : (int64_t)a.num_cus * 16);
  hipLaunchKernelGGL(k_bin, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

bool tile_supported(int tw, int th) {
  return (tw == 8 && th == 8) || (tw == 16 && th == 16) || (tw == 32 && th == 32);
}

hipError_t launch_tile(const BatchArgs& a, hipStream_t s) {
  const int grid = min(a.ntiles, a.num_cus * 8);
  if (a.tile_w == 8 && a.tile_h == 8)
    hipLaunchKernelGGL((k_tile<8, 8, 64>), dim3(grid), dim3(64), 0, s, a);
  else if (a.tile_w == 16 && a.tile_h == 16)
    hipLaunchKernelGGL((k_tile<16, 16, 256>), dim3(grid), dim3(256), 0, s, a);
  else if (a.tile_w == 32 && a.tile_h == 32)
    hipLaunchKernelGGL((k_tile<32, 32, 256>), dim3(grid), dim3(256), 0, s, a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace mr
