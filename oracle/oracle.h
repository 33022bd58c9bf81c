/*
 * oracle.h -- CPU restatement of minirender's Renderer::render().
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product path (minirender_amd) never
 * calls it and fails loudly when its HIP library is missing.
 *
 * The restatement follows src/Renderer.cpp of aslze/minirender operation by
 * operation, single-threaded, in fp32 with no FMA contraction (built with
 * -O3 -ffp-contract=off, no -march).  It takes the same mr_frame/mr_draw
 * description as the HIP path (include/minirender_amd.h) so both see
 * identical host-side matrices.
 *
 * Pinning (tests/test_oracle_kat.py): projection closed forms, znear,
 * exact-arithmetic coverage/depth on dyadic triangles, inclusive fill rule,
 * first-drawn-wins ties, back-face cull sign, clip case analysis and areas,
 * Lambert/Blinn-Phong closed forms, nearest texture lookup, perspective
 * depth of a tilted plane, RGB8 quantization.  The ASL conventions that
 * only the real library could pin (Matrix4*Vec3 summation order,
 * Vec3::normalized) are fixed as documented in oracle.c.
 */
#ifndef MINIRENDER_ORACLE_H
#define MINIRENDER_ORACLE_H

#include "../include/minirender_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host copy of a TriMesh (Scene.h:73-87). */
typedef struct mro_mesh {
  const float* positions;  /* npositions x 3 */
  int32_t npositions;
  const float* normals;    /* nnormals x 3   */
  int32_t nnormals;
  const float* uvs;        /* nuvs x 2, may be NULL */
  int32_t nuvs;
  const int32_t* indices;         /* ntriangles x 3 */
  const int32_t* normal_indices;  /* ntriangles x 3 */
  const int32_t* uv_indices;      /* ntriangles x 3, may be NULL */
  int32_t ntriangles;
} mro_mesh;

typedef struct mro_texture {
  const void* texels;  /* rows x cols x 3, MR_TEX_RGB8 or MR_TEX_RGB32F */
  int32_t format;
  int32_t rows, cols;
} mro_texture;

/* Renderer::render() (Renderer.cpp:311-332) on the CPU.  draws[i].mesh and
 * draws[i].texture index `meshes` / `textures`.  Outputs are host pointers
 * (out->on_device must be 0); any may be NULL.  Returns 0 or MR_ERR_*. */
int mro_render(const mr_frame* frame, const mro_mesh* meshes, int32_t nmeshes,
               const mro_texture* textures, int32_t ntextures,
               const mr_outputs* out);

/* Number of input triangles that crossed the near plane in the last call of
 * mro_render on this thread (statistics for tests). */
int64_t mro_last_clipped(void);

/* The near-plane depth the oracle uses (Renderer.cpp:324-326). */
float mro_znear(const float projection[16]);

/* Exposed pieces of the pipeline, for known-answer tests. */
/* paintTriangle (Renderer.cpp:163-309) of one world-space triangle with the
 * given frame state, into caller-provided buffers (depth pre-cleared by the
 * caller).  vtx = 3 x (px,py,pz, nx,ny,nz, u,v).  order = draw index of
 * clip sub-triangle 0.  Returns the number of sub-triangles painted. */
int mro_paint_triangle(const mr_frame* frame, const mr_draw* material,
                       const mro_texture* texture, const float vtx[24],
                       uint32_t order, float* image, float* depth,
                       uint32_t* winner, float* normals);

/* clipTriangle (Renderer.cpp:131-161) alone: writes up to 2 output triangles
 * (2 x 3 x 8 floats) and returns how many. */
int mro_clip_triangle(float znear, const float vtx[24], float out[48]);

#ifdef __cplusplus
}
#endif
#endif
