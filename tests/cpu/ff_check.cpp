// CPU check of mr::ff_add / mr::advance / mr::advance2 (minirender_amd/csrc/exact.h):
// n repetitions of the fp32 recurrence s = s + d (Renderer.cpp:243) must be
// bit-identical to the serial loop.  Built with g++ -O2 -ffp-contract=off.
//
// usage: ff_check <cases> <seed>; prints "ok <cases>" or the first mismatch.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "exact.h"

static uint64_t rng_state;
static uint64_t next_u64() {  // splitmix64
  uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double uni() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }
static float bits(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static uint32_t fbits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

static float rand_start() {
  switch (next_u64() % 8) {
    case 0: return (float)(uni() * 20.0 - 10.0);
    case 1: return (float)(uni() * 2.0 - 1.0);
    case 2: {  // log-uniform magnitude, random sign
      float m = (float)std::ldexp(1.0 + uni(), (int)(next_u64() % 160) - 140);
      return (next_u64() & 1) ? -m : m;
    }
    case 3: {  // a power of two, +-few ulps
      float p = std::ldexp(1.0f, (int)(next_u64() % 40) - 30);
      uint32_t b = fbits(p) + (uint32_t)(next_u64() % 7) - 3u;
      float f = bits(b);
      return (next_u64() & 1) ? -f : f;
    }
    case 4: return bits((uint32_t)(next_u64() % 0x00800000u)) * ((next_u64() & 1) ? -1.0f : 1.0f);
    case 5: return (next_u64() & 1) ? 0.0f : -0.0f;
    case 6: return bits((uint32_t)next_u64());  // any pattern, incl. inf/NaN
    default: {  // an edge value like the rasterizer's: n . (pt - p)
      float nx = (float)(uni() * 0.2 - 0.1), px = (float)(uni() * 1920.0);
      float ny = (float)(uni() * 0.2 - 0.1), py = (float)(uni() * 1080.0);
      float x = std::floor((float)(uni() * 1920.0)) + 0.5f, y = std::floor((float)(uni() * 1080.0)) + 0.5f;
      return nx * (x - px) + ny * (y - py);
    }
  }
}

static float rand_step() {
  switch (next_u64() % 7) {
    case 0: return (float)(uni() * 0.2 - 0.1);
    case 1: {
      float m = (float)std::ldexp(1.0 + uni(), (int)(next_u64() % 50) - 40);
      return (next_u64() & 1) ? -m : m;
    }
    case 2: {  // few significant bits: many rounding ties
      int odd = (int)(next_u64() % 64) * 2 + 1;
      float m = std::ldexp((float)odd, (int)(next_u64() % 30) - 36);
      return (next_u64() & 1) ? -m : m;
    }
    case 3: {  // exact power of two
      float m = std::ldexp(1.0f, (int)(next_u64() % 40) - 30);
      return (next_u64() & 1) ? -m : m;
    }
    case 4: return bits((uint32_t)(next_u64() % 0x00800000u)) * ((next_u64() & 1) ? -1.0f : 1.0f);
    case 5: return (next_u64() & 1) ? 0.0f : -0.0f;
    default: return 1.0f / (float)(1 + next_u64() % 3000) * ((next_u64() & 1) ? -1.0f : 1.0f);
  }
}

// Equal bit patterns, or both NaN (payloads may differ).
static bool same_value(float x, float y) {
  return fbits(x) == fbits(y) || (std::isnan(x) && std::isnan(y));
}

int main(int argc, char** argv) {
  const long cases = argc > 1 ? atol(argv[1]) : 200000;
  rng_state = argc > 2 ? strtoull(argv[2], nullptr, 10) : 1;
  for (long c = 0; c < cases; c++) {
    const float s0 = rand_start(), d = rand_step();
    const int n = (next_u64() % 4 == 0) ? (int)(next_u64() % 20000) : (int)(next_u64() % 2000);
    float ref = s0;
    for (int i = 0; i < n; i++) ref = ref + d;
    const float got = mr::ff_add(s0, d, n);
    const float got2 = mr::advance(s0, d, n);
    // advance2 on the pair (s0 by d) and (d by s0): two recurrences at once
    float ref2 = d;
    for (int i = 0; i < n; i++) ref2 = ref2 + s0;
    float a2 = s0, b2 = d;
    mr::advance2(a2, d, b2, s0, n);
    const bool same = same_value(ref, got) && same_value(ref, got2) && same_value(ref, a2) &&
                      same_value(ref2, b2);
    if (!same) {
      printf("mismatch s=%a d=%a n=%d ref=%a ff=%a adv=%a adv2=%a/%a\n", s0, d, n, ref, got, got2, a2, b2);
      return 1;
    }
  }
  printf("ok %ld\n", cases);
  return 0;
}
