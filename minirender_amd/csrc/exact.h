// exact.h -- fp32 arithmetic of the rasterizer, shared by every kernel.
//
// Each function states the reference lines (aslze/minirender
// src/Renderer.cpp) whose operation order it keeps.  Device code is built
// with -ffp-contract=off: every a*b+c below is two rounded operations, as on
// the reference's x86-64 SSE build.  Division and sqrt are IEEE correctly
// rounded on gfx950 (hipcc default), fp32 denormals are preserved.
//
// The functions are plain inline code so a CPU build (g++) of this header can
// check them against straightforward serial loops.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "common.h"

namespace mr {

struct f2 { float x, y; };
struct f3 { float x, y, z; };
struct Vtx { f3 p, n; f2 uv; };  // Vertex, Scene.h:14-24

MR_HD inline uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
MR_HD inline float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }
MR_HD inline double pow2d(int e) {  // exact 2^e for -1022 <= e <= 1023
  return __builtin_bit_cast(double, (uint64_t)(1023 + e) << 52);
}

// ASL min/max/clamp: a<b?a:b, a>b?a:b, min(max(x,a),b).
MR_HD inline float amin(float a, float b) { return a < b ? a : b; }
MR_HD inline float amax(float a, float b) { return a > b ? a : b; }
MR_HD inline float aclamp(float x, float a, float b) { return amin(amax(x, a), b); }
// The bounding-box minimum / maximum of triangle setup, as the hardware
// min / max (one instruction; amin and amax compile to a compare and a
// select).  Equal to amin / amax but for NaN and the sign of a zero: a zero
// of either sign clamps and floors to the same pixel, and a triangle with a
// NaN coordinate has a NaN area and is rejected by setup either way.
MR_HD inline float bmin(float a, float b) { return fminf(a, b); }
MR_HD inline float bmax(float a, float b) { return fmaxf(a, b); }
MR_HD inline float bclamp(float x, float a, float b) { return bmin(bmax(x, a), b); }

MR_HD inline f3 add3(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
MR_HD inline f3 sub3(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
MR_HD inline f3 scale3(float k, f3 a) { return f3{k * a.x, k * a.y, k * a.z}; }
MR_HD inline float dot3(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
MR_HD inline float len3(f3 a) { return sqrtf(dot3(a, a)); }
MR_HD inline f3 norm3(f3 a) {
  float l = len3(a);
  return f3{a.x / l, a.y / l, a.z / l};
}

// ASL Matrix4 * Vec3, affine, rows 0..2 of M (Renderer.cpp:345,348).
MR_HD inline f3 xform(const float* M, f3 p) {
  return f3{((M[0] * p.x + M[1] * p.y) + M[2] * p.z) + M[3],
            ((M[4] * p.x + M[5] * p.y) + M[6] * p.z) + M[7],
            ((M[8] * p.x + M[9] * p.y) + M[10] * p.z) + M[11]};
}

// htransform, Renderer.cpp:13-20.
MR_HD inline f3 htransform(const float* M, f3 p) {
  float iw = 1.0f / (((M[12] * p.x + M[13] * p.y) + M[14] * p.z) + M[15]);
  return f3{(((M[0] * p.x + M[1] * p.y) + M[2] * p.z) + M[3]) * iw,
            (((M[4] * p.x + M[5] * p.y) + M[6] * p.z) + M[7]) * iw,
            (((M[8] * p.x + M[9] * p.y) + M[10] * p.z) + M[11]) * iw};
}

// Renderer.cpp:324-326.
MR_HD inline float znear_of(const float* P) {
  bool persp = P[15] == 0.0f;
  float zn = persp ? P[11] / (P[10] - 1.0f) : (P[11] + 1.0f) / P[10];
  return -zn;
}

// ---------------------------------------------------------------------------
// Exact repetition of the edge-walk recurrence e += n.x (Renderer.cpp:243).
//
// ff_add(s, d, n) returns the value of `for (i < n) s = s + d;` in IEEE fp32
// round-to-nearest-even, without doing n dependent adds.  While s and the
// exact sum s + d stay inside one binade [2^E, 2^(E+1)) (scaled to units of
// u = 2^(E-23), s is an integer S in [2^23, 2^24)), every step adds the same
// integer D = round(d/u) -- unless d/u is exactly half way between integers
// (then the rounding depends on S's parity).  So the loop jumps over every
// run of such steps in one multiply, and takes single IEEE steps where the
// premise fails: zero/subnormal/inf/NaN operands, |d| >= 2^E, ties.  All
// double arithmetic below is exact (operands need < 53 bits) except the one
// division, whose result is corrected by exact comparisons.
// ---------------------------------------------------------------------------
MR_HD inline float ff_add(float s, float d, int n) {
  while (n > 0) {
    uint32_t sb = f2u(s);
    uint32_t se = (sb >> 23) & 0xFFu;
    uint32_t db = f2u(d);
    if (se == 0u || se == 0xFFu || ((db >> 23) & 0xFFu) == 0xFFu) {
      s = s + d;  // zero / subnormal / inf / NaN: one IEEE step
      --n;
      continue;
    }
    if ((db & 0x7FFFFFFFu) == 0u) return s;  // s + (+-0) == s for s != 0
    const int E = (int)se - 127;
    const int64_t S = (int64_t)((sb & 0x7FFFFFu) | 0x800000u);
    double q = (double)d * pow2d(23 - E);  // d in units of u (exact)
    if (sb >> 31) q = -q;                   // mirror: work on |s|
    if (fabs(q) >= 8388608.0) {             // |d| >= 2^E: may leave the binade
      s = s + d;
      --n;
      continue;
    }
    const double fq = floor(q);
    if (q - fq == 0.5) {  // tie: rounding depends on the parity of S
      s = s + d;
      --n;
      continue;
    }
    const int64_t D = (int64_t)rint(q);
    if (D == 0) {
      // |d| < u/2: s is a fixed point of s = s + d unless s + d drops below
      // 2^E, where the spacing halves (s a power of two stepping down).
      if ((double)S + q >= 8388608.0) return s;
      s = s + d;
      --n;
      continue;
    }
    int64_t m;
    if (D > 0) {  // step k valid while S + k*D + q < 2^24
      const double R = (16777216.0 - (double)S) - q;
      if (R <= 0.0) {
        m = 0;
      } else {
        m = (int64_t)ceil(R / (double)D);
        while (m > 0 && !((double)(S + (m - 1) * D) + q < 16777216.0)) --m;
        while ((double)(S + m * D) + q < 16777216.0) ++m;
      }
    } else {  // step k valid while S + k*D + q >= 2^23
      const double R = ((double)S + q) - 8388608.0;
      if (R < 0.0) {
        m = 0;
      } else {
        m = (int64_t)floor(R / (double)(-D)) + 1;
        while (m > 0 && !((double)(S + (m - 1) * D) + q >= 8388608.0)) --m;
        while ((double)(S + m * D) + q >= 8388608.0) ++m;
      }
    }
    if (m == 0) {
      s = s + d;
      --n;
      continue;
    }
    if (m > n) m = n;
    const float r = (float)((double)(S + m * D) * pow2d(E - 23));
    s = (sb >> 31) ? -r : r;
    n -= (int)m;
  }
  return s;
}

// s after n steps of s = s + d: serial for short runs (one add each), ff_add
// (a few hundred instructions of fp64 bookkeeping) for long ones.
MR_HD inline float advance(float s, float d, int n) {
  if (n <= 256) {
    for (int i = 0; i < n; i++) s = s + d;
    return s;
  }
  return ff_add(s, d, n);
}

// Both edge values of a row after n columns: the two recurrences interleaved
// and unrolled by 8, so a wavefront spends about one instruction per add on
// loop control instead of three or four (lanes with different n run the same
// instruction stream under their execution mask).
MR_HD inline void advance2(float& e1, float d1, float& e2, float d2, int n) {
  if (n > 256) {
    e1 = ff_add(e1, d1, n);
    e2 = ff_add(e2, d2, n);
    return;
  }
  float a = e1, b = e2;
  for (int i = n >> 3; i > 0; i--) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      a = a + d1;
      b = b + d2;
    }
  }
  if (n & 4) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      a = a + d1;
      b = b + d2;
    }
  }
  if (n & 2) {
    a = a + d1;
    b = b + d2;
    a = a + d1;
    b = b + d2;
  }
  if (n & 1) {
    a = a + d1;
    b = b + d2;
  }
  e1 = a;
  e2 = b;
}

// Row start of the edge walk at the bbox's left column (Renderer.cpp:240-242).
// x0f = floor(pmin.x) + 0.5f, y = row + 0.5f.
MR_HD inline void edge_start(float n1x, float n1y, float n2x, float n2y,
                             float p0x, float p0y, float p2x, float p2y,
                             float x0f, float y, float& e1, float& e2) {
  e1 = n1x * (x0f - p2x) + n1y * (y - p2y);
  e2 = n2x * (x0f - p0x) + n2y * (y - p0y);
}

// Coverage (Renderer.cpp:245): inclusive edges, no tie rule.
MR_HD inline bool covered(float e1, float e2) {
  return !(e1 < 0.0f || e2 < 0.0f || (1.0f - e1) - e2 < 0.0f);
}

// Depth and perspective-corrected weights (Renderer.cpp:247-261).
MR_HD inline float depth_of(float e1, float e2, float z0, float z1, float z2,
                            bool persp, float k[3]) {
  k[0] = (1.0f - e1) - e2;
  k[1] = e1;
  k[2] = e2;
  if (persp) {
    float z = 1.0f / ((k[0] * z0 + k[1] * z1) + k[2] * z2);
    k[0] = k[0] * (z0 * z);
    k[1] = k[1] * (z1 * z);
    k[2] = k[2] * (z2 * z);
    return z;
  }
  return ((k[0] * z0 + k[1] * z1) + k[2] * z2) + 0.0f * 1.0f;  // + k[3]*zz[3]
}

// Order-preserving 32-bit key of a depth (NaN excluded by the caller):
// -0 and +0 share a key so that equal depths tie-break on draw order, as the
// reference's strict `z < depth` does (Renderer.cpp:267).
MR_HD inline uint32_t depth_key(float z) {
  uint32_t b = f2u(z);
  if (b == 0x80000000u) b = 0u;
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// clip, Renderer.cpp:121-129.
MR_HD inline Vtx clip_edge(float z, const Vtx& a, const Vtx& b) {
  float d = b.p.z - a.p.z;
  float k = (fabsf(d) < 1e-6f) ? 0.5f : (z - a.p.z) / d;
  float k1 = 1.0f - k;
  Vtx v;
  v.p = add3(scale3(k, b.p), scale3(k1, a.p));
  v.n = add3(scale3(k, b.n), scale3(k1, a.n));
  v.uv.x = k * b.uv.x + k1 * a.uv.x;
  v.uv.y = k * b.uv.y + k1 * a.uv.y;
  return v;
}

// clipTriangle, Renderer.cpp:131-161, for a triangle with at least one vertex
// beyond the near plane and not all three.  Returns the sub-triangles in draw
// order, (a0,a1,a2) and (b0,b1,b2), and their number (1 or 2).
// The reference rotates (a,b,c) -> (c,a,b) while v0 is not the farthest
// (Renderer.cpp:136-140).  That stops after at most two rotations (also with
// NaNs: some state always has a v0 that no other z exceeds), so the rotation
// count is picked directly: 0 if a is farthest, else 1 if c is, else 2.
MR_HD inline int rotations(float za, float zb, float zc) {
  if (!(za < zb || za < zc)) return 0;
  if (!(zc < za || zc < zb)) return 1;
  return 2;
}

MR_HD inline int clip_triangle(float z, Vtx v0, Vtx v1, Vtx v2, Vtx& a0, Vtx& a1,
                               Vtx& a2, Vtx& b0, Vtx& b1, Vtx& b2) {
  {
    const int r = rotations(v0.p.z, v1.p.z, v2.p.z);
    const Vtx A = v0, B = v1, C = v2;
    v0 = r == 0 ? A : (r == 1 ? C : B);
    v1 = r == 0 ? B : (r == 1 ? A : C);
    v2 = r == 0 ? C : (r == 1 ? B : A);
  }
  // Every output is assigned on every path (keeps them in registers).
  const bool beyond1 = v1.p.z > z, beyond2 = v2.p.z > z;
  const Vtx p = clip_edge(z, v0, beyond1 ? v2 : v1);                 // v02 | v01
  const Vtx q = clip_edge(z, beyond1 || beyond2 ? v1 : v0, v2);      // v12 | v02
  if (beyond1) {         // v0, v1 beyond: (v02, v12, v2)
    a0 = p; a1 = q; a2 = v2;
  } else if (beyond2) {  // v0, v2 beyond: (v01, v1, v12)
    a0 = p; a1 = v1; a2 = q;
  } else {               // only v0 beyond: (v01, v1, v2), (v01, v2, v02)
    a0 = p; a1 = v1; a2 = v2;
  }
  b0 = p; b1 = v2; b2 = q;
  return (beyond1 || beyond2) ? 1 : 2;
}

// Position part of clip (Renderer.cpp:123,125): k depends on positions only,
// so this equals clip_edge's position bit for bit.
MR_HD inline f3 clip_pos(float z, f3 a, f3 b) {
  float d = b.z - a.z;
  float k = (fabsf(d) < 1e-6f) ? 0.5f : (z - a.z) / d;
  return add3(scale3(k, b), scale3(1.0f - k, a));
}

// clipTriangle on positions only, for setup and binning.
MR_HD inline int clip_positions(float z, f3 v0, f3 v1, f3 v2, f3& a0, f3& a1,
                                f3& a2, f3& b0, f3& b1, f3& b2) {
  {
    const int r = rotations(v0.z, v1.z, v2.z);
    const f3 A = v0, B = v1, C = v2;
    v0 = r == 0 ? A : (r == 1 ? C : B);
    v1 = r == 0 ? B : (r == 1 ? A : C);
    v2 = r == 0 ? C : (r == 1 ? B : A);
  }
  // Every output is assigned on every path (keeps them in registers).
  const bool beyond1 = v1.z > z, beyond2 = v2.z > z;
  const f3 p = clip_pos(z, v0, beyond1 ? v2 : v1);    // v02 | v01
  const f3 q = clip_pos(z, beyond1 || beyond2 ? v1 : v0, v2);  // v12 | v02
  if (beyond1) {         // v0, v1 beyond: (v02, v12, v2)
    a0 = p; a1 = q; a2 = v2;
  } else if (beyond2) {  // v0, v2 beyond: (v01, v1, v12)
    a0 = p; a1 = v1; a2 = q;
  } else {               // only v0 beyond: (v01, v1, v2), (v01, v2, v02)
    a0 = p; a1 = v1; a2 = v2;
  }
  b0 = p; b1 = v2; b2 = q;
  return (beyond1 || beyond2) ? 1 : 2;
}

// Raster state of one (sub-)triangle after setup.
struct TriSetup {
  float n1x, n1y, n2x, n2y;  // edge normals, Renderer.cpp:215-216
  float p0x, p0y, p2x, p2y;  // pixel positions of vertices 0 and 2
  float z0, z1, z2;          // -1/z_eye (perspective) or NDC z (ortho)
  int bx0, bx1, by0, by1;    // clamped pixel bbox, inclusive
};

// A vertex projected to the screen (Renderer.cpp:191-203, 225-227): pixel
// position and the depth attribute the raster interpolates, -1/z_eye under a
// perspective projection or NDC z under an orthographic one.  It depends on
// the vertex alone, so it can be computed once per vertex.
struct ScreenVtx {
  float px, py, zv;
};

MR_HD inline ScreenVtx project_vertex(const float* P, bool persp, float w, float h, f3 e) {
  const f3 ndc = htransform(P, e);
  ScreenVtx v;
  v.px = (1.0f + ndc.x) * (w / 2.0f);
  v.py = (1.0f - ndc.y) * (h / 2.0f);
  v.zv = persp ? -1.0f / e.z : ndc.z;
  return v;
}

// Last pixel index the reference's loop `for (v = floor(lo) + 0.5f;
// v <= hi + 0.5f; v++)` visits (Renderer.cpp:238,243), for 0 <= hi < 2^23.
// Pixel centres k + 0.5f are exact, so the last one is the largest k + 0.5
// not above fl(hi + 0.5f).  That sum rounds up when hi lies within half an
// ulp of the result below a power of two (1023.99994f + 0.5f == 1024.5f):
// the loop then visits floor(hi) + 1.  fl(hi + 0.5f) - 0.5f is exact here.
MR_HD inline int last_center(float hi) {
  const float end = hi + 0.5f;
  return (int)floorf(end - 0.5f);
}

// Triangle setup of paintTriangle (Renderer.cpp:193-236) from projected
// vertices.  Returns false when the triangle is rejected off-screen or
// culled (front faces are counter-clockwise in NDC, :207).
MR_HD inline bool setup_screen(float w, float h, const ScreenVtx& v0, const ScreenVtx& v1,
                               const ScreenVtx& v2, TriSetup& r) {
  const float p0x = v0.px, p0y = v0.py, p1x = v1.px, p1y = v1.py, p2x = v2.px, p2y = v2.py;
  float minx = bmin(bmin(bmin(1e30f, p0x), p1x), p2x);
  float miny = bmin(bmin(bmin(1e30f, p0y), p1y), p2y);
  float maxx = bmax(bmax(bmax(-1e30f, p0x), p1x), p2x);
  float maxy = bmax(bmax(bmax(-1e30f, p0y), p1y), p2y);
  if (maxx < 0 || maxy < 0 || minx > w || miny > h) return false;
  float ar = (p0x - p1x) * (p2y - p1y) - (p0y - p1y) * (p2x - p1x);
  if (!(ar > 0.0f)) {
    // a <= 0 is culled (:207); a NaN passes the cull in the reference but
    // every edge value is then NaN and so is z: nothing is written.
    return false;
  }
  float i2a = -1.0f / ar;
  r.n1x = -(p0y - p2y) * i2a;
  r.n1y = (p0x - p2x) * i2a;
  r.n2x = -(p1y - p0y) * i2a;
  r.n2y = (p1x - p0x) * i2a;
  r.p0x = p0x; r.p0y = p0y; r.p2x = p2x; r.p2y = p2y;
  minx = bclamp(minx, 0.0f, w - 1);
  maxx = bclamp(maxx, 0.0f, w - 1);
  miny = bclamp(miny, 0.0f, h - 1);
  maxy = bclamp(maxy, 0.0f, h - 1);
  r.bx0 = (int)floorf(minx);
  r.bx1 = last_center(maxx);
  r.by0 = (int)floorf(miny);
  r.by1 = last_center(maxy);
  r.z0 = v0.zv;
  r.z1 = v1.zv;
  r.z2 = v2.zv;
  return true;
}

// What shading needs of setup_screen for a triangle that passed it (a
// winner): edge normals, vertices 0 and 2, depths and the bbox's columns,
// by the same arithmetic.
MR_HD inline void setup_edges(float w, const ScreenVtx& v0, const ScreenVtx& v1,
                              const ScreenVtx& v2, TriSetup& r) {
  const float p0x = v0.px, p0y = v0.py, p1x = v1.px, p1y = v1.py, p2x = v2.px, p2y = v2.py;
  const float minx = bmin(bmin(bmin(1e30f, p0x), p1x), p2x);
  const float maxx = bmax(bmax(bmax(-1e30f, p0x), p1x), p2x);
  const float ar = (p0x - p1x) * (p2y - p1y) - (p0y - p1y) * (p2x - p1x);
  const float i2a = -1.0f / ar;
  r.n1x = -(p0y - p2y) * i2a;
  r.n1y = (p0x - p2x) * i2a;
  r.n2x = -(p1y - p0y) * i2a;
  r.n2y = (p1x - p0x) * i2a;
  r.p0x = p0x; r.p0y = p0y; r.p2x = p2x; r.p2y = p2y;
  r.bx0 = (int)floorf(bclamp(minx, 0.0f, w - 1));
  r.bx1 = last_center(bclamp(maxx, 0.0f, w - 1));
  r.z0 = v0.zv;
  r.z1 = v1.zv;
  r.z2 = v2.zv;
}

// Triangle setup (Renderer.cpp:179-236) for eye-space positions.
MR_HD inline bool setup_triangle(const float* P, bool persp, float w, float h,
                                 f3 a, f3 b, f3 c, TriSetup& r) {
  return setup_screen(w, h, project_vertex(P, persp, w, h, a),
                      project_vertex(P, persp, w, h, b), project_vertex(P, persp, w, h, c), r);
}

// Range image point of pixel (i = row, j = column) from its depth
// (Renderer.cpp:388-415): back-projection through the projection's diagonal
// terms, (0,0,0) past the far plane.
MR_HD inline float range_far_depth(const float* P) {
  const bool persp = P[15] == 0.0f;
  const float zfar = persp ? P[11] / (P[10] + 1) : (P[11] - 1) / P[10];
  return persp ? zfar : 1.0f;
}

MR_HD inline f3 range_point(const float* P, float fardepth, float w, float h, int i, int j,
                            float depth) {
  if (depth > fardepth) return f3{0.0f, 0.0f, 0.0f};
  const float u = (j + 0.5f) / (w / 2) - 1;
  const float v = -(i + 0.5f) / (h / 2) + 1;
  const float z = -depth;
  const float x = -(u + P[2]) * z / P[0];
  const float y = -(v + P[6]) * z / P[5];
  return f3{x, y, z};
}

// fract + nearest texel index of Renderer.cpp:274, clamped to n-1 where the
// reference would read one past the end (fract rounding to 1.0f).
MR_HD inline int tex_index(float u, int n) {
  float fr = u - floorf(u);
  float v = fr * (float)n;
  int i = (v >= 0.0f && v < 2147483648.0f) ? (int)v : 0;
  return i > n - 1 ? n - 1 : i;
}

MR_HD inline f3 texel(const void* tex, int fmt, int cols, int i, int j) {
  size_t o = ((size_t)i * (size_t)cols + (size_t)j) * 3;
  if (fmt == 0) {  // RGB8: byte / 255 (io.cpp:405-411)
    const uint8_t* p = (const uint8_t*)tex + o;
    return f3{(float)p[0] / 255.0f, (float)p[1] / 255.0f, (float)p[2] / 255.0f};
  }
  const float* p = (const float*)tex + o;
  return f3{p[0], p[1], p[2]};
}

// savePPM quantization, io.cpp:356-361.
MR_HD inline uint8_t quantize(float v) {
  return (uint8_t)aclamp(v * 255.0f, 0.0f, 255.0f);
}

}  // namespace mr
