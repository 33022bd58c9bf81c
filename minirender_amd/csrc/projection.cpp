// projection.cpp -- projection matrices of the reference API
// (src/Renderer.cpp:40-83, include/minirender/Renderer.h:9-14), row-major,
// evaluated in fp32 in the reference's expression order.
#include <cmath>
#include <cstring>

#include "minirender_amd.h"
#include "exact.h"

namespace {

void set16(float* o, float a0, float a1, float a2, float a3, float b0, float b1,
           float b2, float b3, float c0, float c1, float c2, float c3, float d0,
           float d1, float d2, float d3) {
  const float v[16] = {a0, a1, a2, a3, b0, b1, b2, b3, c0, c1, c2, c3, d0, d1, d2, d3};
  memcpy(o, v, sizeof(v));
}

}  // namespace

extern "C" {

// Renderer.cpp:40-47 (glOrtho).
int mr_projection_ortho(float l, float r, float b, float t, float n, float f, float* o) {
  if (!o) return MR_ERR_ARG;
  set16(o, 2.0f / (r - l), 0, 0, -(r + l) / (r - l),
        0, 2.0f / (t - b), 0, -(t + b) / (t - b),
        0, 0, -2.0f / (f - n), -(f + n) / (f - n),
        0, 0, 0, 1);
  return MR_OK;
}

// Renderer.cpp:49-56 (glFrustum).
int mr_projection_perspective(float l, float r, float b, float t, float n, float f, float* o) {
  if (!o) return MR_ERR_ARG;
  set16(o, 2.0f * n / (r - l), 0, (r + l) / (r - l), 0,
        0, 2.0f * n / (t - b), (t + b) / (t - b), 0,
        0, 0, -(f + n) / (f - n), -2.0f * f * n / (f - n),
        0, 0, -1, 0);
  return MR_OK;
}

// Renderer.cpp:68-72: vertical field of view.
int mr_projection_frustum(float fov, float aspect, float n, float f, float* o) {
  const float t = tanf(fov / 2.0f);
  return mr_projection_perspective(-n * t * aspect, n * t * aspect, -n * t, n * t, n, f, o);
}

// Renderer.cpp:74-78: horizontal field of view.
int mr_projection_frustum_h(float fov, float aspect, float n, float f, float* o) {
  const float t = tanf(fov / 2.0f);
  return mr_projection_perspective(-n * t, n * t, -n * t / aspect, n * t / aspect, n, f, o);
}

// Renderer.cpp:80-83.
int mr_projection_ortho_fov(float fov, float aspect, float n, float f, float* o) {
  return mr_projection_ortho(-fov * aspect / 2.0f, fov * aspect / 2.0f, -fov / 2.0f,
                             fov / 2.0f, n, f, o);
}

// Renderer.cpp:58-66: OpenCV-style intrinsics K (row-major 4x4, 3x3 used).
int mr_projection_cv(const float* K, float w, float h, float n, float f, float* o) {
  if (!K || !o) return MR_ERR_ARG;
  set16(o, K[0] * 2.0f / w, K[1] * 2.0f / w, -2.0f * K[2] / w + 1.0f, 0,
        K[4] * 2.0f / h, K[5] * 2.0f / h, 2.0f * K[6] / h - 1.0f, 0,
        0, 0, -(f + n) / (f - n), -2.0f * f * n / (f - n),
        0, 0, -1, 0);
  return MR_OK;
}

int mr_znear(const float* P, float* zn) {
  if (!P || !zn) return MR_ERR_ARG;
  *zn = mr::znear_of(P);
  return MR_OK;
}

}  // extern "C"
