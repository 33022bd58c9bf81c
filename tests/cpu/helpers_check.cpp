// CPU check of device helpers whose form differs from the reference's:
//   remap  xcd_remap (common.h) is a bijection on [0, n) for every n tested;
//   clip   clip_positions / clip_triangle (exact.h) pick the same rotation
//          and emit the same sub-triangles, bit for bit, as the reference's
//          rotate-while-loop clipTriangle (Renderer.cpp:131-161), NaNs and
//          ties included.
//   bound  last_center (exact.h) is the last pixel centre the reference's
//          fp32 loop `for (x = floor(lo) + 0.5f; x <= hi + 0.5f; x++)`
//          visits (Renderer.cpp:238,243), for every float within 64 ulps of
//          each power of two from 2 to 16384 and for random floats below
//          16384 (where hi + 0.5f rounds up, the loop visits one more).
// Built with g++ -O2 -ffp-contract=off.  Prints "ok" or the first failure.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "exact.h"

using namespace mr;

static uint64_t st = 7;
static uint64_t rnd() {
  uint64_t z = (st += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static float rz() {
  switch (rnd() % 6) {
    case 0: return (float)((int)(rnd() % 7) - 3);  // ties
    case 1: return NAN;
    default: return (float)((double)(rnd() >> 11) / 9007199254740992.0 * 8.0 - 6.0);
  }
}
static bool same(float a, float b) {
  uint32_t x, y;
  memcpy(&x, &a, 4);
  memcpy(&y, &b, 4);
  return x == y || (std::isnan(a) && std::isnan(b));
}
static bool same3(f3 a, f3 b) { return same(a.x, b.x) && same(a.y, b.y) && same(a.z, b.z); }
static bool samev(const Vtx& a, const Vtx& b) {
  return same3(a.p, b.p) && same3(a.n, b.n) && same(a.uv.x, b.uv.x) && same(a.uv.y, b.uv.y);
}

// The reference's loop form (Renderer.cpp:131-161).
static int clip_loop(float z, Vtx v[3], Vtx out[6]) {
  int guard = 0;
  while (v[0].p.z < v[1].p.z || v[0].p.z < v[2].p.z) {
    Vtx t = v[0]; v[0] = v[1]; v[1] = t;
    t = v[0]; v[0] = v[2]; v[2] = t;
    if (++guard > 3) return -1;
  }
  if (v[1].p.z > z) {
    out[0] = clip_edge(z, v[0], v[2]); out[1] = clip_edge(z, v[1], v[2]); out[2] = v[2];
    return 1;
  } else if (v[2].p.z > z) {
    out[0] = clip_edge(z, v[0], v[1]); out[1] = v[1]; out[2] = clip_edge(z, v[1], v[2]);
    return 1;
  }
  Vtx v01 = clip_edge(z, v[0], v[1]), v02 = clip_edge(z, v[0], v[2]);
  out[0] = v01; out[1] = v[1]; out[2] = v[2];
  out[3] = v01; out[4] = v[2]; out[5] = v02;
  return 2;
}

// The reference's loop, run from a few columns below hi: index of the last
// centre visited.
static int loop_last(float hi) {
  float start = floorf(hi) - 3.0f;
  if (start < 0.0f) start = 0.0f;
  int last = -1;
  for (float x = start + 0.5f; x <= hi + 0.5f; x++) last = (int)x;
  return last;
}

static bool check_bound(float hi, long* rounded_up) {
  const int want = loop_last(hi), got = last_center(hi);
  if (want != got) {
    printf("bound mismatch: hi=%a (%.9g) loop %d last_center %d\n", hi, hi, want, got);
    return false;
  }
  if (want != (int)floorf(hi)) ++*rounded_up;
  return true;
}

int main(int argc, char** argv) {
  const long cases = argc > 1 ? atol(argv[1]) : 200000;
  // bound
  long rounded_up = 0;
  for (int k = 1; k <= 14; k++) {
    const float p = ldexpf(1.0f, k);
    float lo = p, hi = p;
    for (int u = 0; u < 64; u++) {
      lo = nextafterf(lo, 0.0f);
      hi = nextafterf(hi, 1e30f);
    }
    for (float v = lo; v <= hi; v = nextafterf(v, 1e30f))
      if (!check_bound(v, &rounded_up)) return 1;
  }
  for (long c = 0; c < cases; c++) {
    const float v = (float)((double)(rnd() >> 11) / 9007199254740992.0 * 16384.0);
    if (!check_bound(v, &rounded_up)) return 1;
  }
  if (rounded_up == 0) {
    printf("bound: no case where the loop visits floor(hi) + 1\n");
    return 1;
  }
  // remap
  for (int64_t n = 1; n <= 4096; n++) {
    std::vector<char> seen(n, 0);
    for (int64_t b = 0; b < n; b++) {
      const int64_t r = xcd_remap(b, n);
      if (r < 0 || r >= n || seen[r]) {
        printf("remap not a bijection: n=%lld b=%lld r=%lld\n", (long long)n, (long long)b, (long long)r);
        return 1;
      }
      seen[r] = 1;
    }
  }
  // clip
  for (long c = 0; c < cases; c++) {
    Vtx v[3];
    for (int i = 0; i < 3; i++) {
      v[i].p = f3{rz(), rz(), rz()};
      v[i].n = f3{rz(), rz(), rz()};
      v[i].uv = f2{rz(), rz()};
    }
    const float z = (rnd() % 4 == 0) ? (float)((int)(rnd() % 5) - 2) : rz();
    // only the inputs the reference clips: some vertex beyond, not all
    const bool b0 = v[0].p.z > z, b1 = v[1].p.z > z, b2 = v[2].p.z > z;
    if (!(b0 || b1 || b2) || (b0 && b1 && b2)) continue;
    Vtx ref[6], w[3] = {v[0], v[1], v[2]};
    const int nr = clip_loop(z, w, ref);
    if (nr < 0) { printf("reference loop did not stop\n"); return 1; }
    Vtx a0, a1, a2, c0, c1, c2;
    const int ng = clip_triangle(z, v[0], v[1], v[2], a0, a1, a2, c0, c1, c2);
    f3 p0, p1, p2, q0, q1, q2;
    const int np = clip_positions(z, v[0].p, v[1].p, v[2].p, p0, p1, p2, q0, q1, q2);
    bool ok = ng == nr && np == nr && samev(a0, ref[0]) && samev(a1, ref[1]) && samev(a2, ref[2]) &&
              same3(p0, ref[0].p) && same3(p1, ref[1].p) && same3(p2, ref[2].p);
    if (ok && nr == 2)
      ok = samev(c0, ref[3]) && samev(c1, ref[4]) && samev(c2, ref[5]) &&
           same3(q0, ref[3].p) && same3(q1, ref[4].p) && same3(q2, ref[5].p);
    if (!ok) {
      printf("clip mismatch case %ld: z=%a zs=%a %a %a n=%d/%d/%d\n", c, z, v[0].p.z, v[1].p.z,
             v[2].p.z, nr, ng, np);
      return 1;
    }
  }
  printf("ok\n");
  return 0;
}
