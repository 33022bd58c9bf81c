/*
 * oracle.c -- CPU restatement of aslze/minirender Renderer::render().
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the HIP path and
 * the timed CPU baseline.  Never linked into the product library.
 *
 * Every function cites the reference lines it restates.  Arithmetic is fp32,
 * evaluated left to right one IEEE operation at a time; the file must be
 * compiled with -ffp-contract=off so that no a*b+c is fused.
 *
 * Conventions of the ASL library (aslze/asl 1.11.14, not in the reference
 * tree), fixed here because the reference tests nothing that pins them:
 *   Matrix4 * Vec3 (affine, no divide)  ((m0*x + m1*y) + m2*z) + m3 per row,
 *                                        the same order as htransform
 *                                        (Renderer.cpp:15-19);
 *   Vec2 a ^ b                           a.x*b.y - a.y*b.x;
 *   Vec2 perpend()                       (-y, x);
 *   dot products                         left to right;
 *   min(a,b) / max(a,b)                  a<b?a:b / a>b?a:b, clamp(x,a,b) =
 *                                        min(max(x,a),b);
 *   Vec3 length / normalized             sqrtf((x*x+y*y)+z*z), v / length
 *                                        per component;
 *   fract(x)                             x - floorf(x).
 * Perp and cross are pinned by the invariant e1(p1) = e2(p2) = 1 (tested),
 * the others only matter for colour (<= 1 LSB) or are exact.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float x, y; } vec2f;
typedef struct { float x, y, z; } vec3f;
typedef struct { vec3f p, n; vec2f uv; } vert; /* Vertex, Scene.h:14-24 */

static float amin(float a, float b) { return a < b ? a : b; }
static float amax(float a, float b) { return a > b ? a : b; }
static float aclamp(float x, float a, float b) { return amin(amax(x, a), b); }

static vec3f v3_add(vec3f a, vec3f b) { vec3f r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static vec3f v3_sub(vec3f a, vec3f b) { vec3f r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static vec3f v3_scale(float k, vec3f a) { vec3f r = {k * a.x, k * a.y, k * a.z}; return r; }
static float v3_dot(vec3f a, vec3f b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static float v3_len(vec3f a) { return sqrtf(v3_dot(a, a)); }
static vec3f v3_norm(vec3f a) {
  float l = v3_len(a);
  vec3f r = {a.x / l, a.y / l, a.z / l};
  return r;
}

/* ASL Matrix4 * Vec3, affine (Renderer.cpp:345,348).  M row-major 4x4. */
static vec3f xform(const float* M, vec3f p) {
  vec3f r;
  r.x = ((M[0] * p.x + M[1] * p.y) + M[2] * p.z) + M[3];
  r.y = ((M[4] * p.x + M[5] * p.y) + M[6] * p.z) + M[7];
  r.z = ((M[8] * p.x + M[9] * p.y) + M[10] * p.z) + M[11];
  return r;
}

/* htransform, Renderer.cpp:13-20. */
static vec3f htransform(const float* M, vec3f p) {
  float iw = 1.0f / (((M[12] * p.x + M[13] * p.y) + M[14] * p.z) + M[15]);
  vec3f r;
  r.x = (((M[0] * p.x + M[1] * p.y) + M[2] * p.z) + M[3]) * iw;
  r.y = (((M[4] * p.x + M[5] * p.y) + M[6] * p.z) + M[7]) * iw;
  r.z = (((M[8] * p.x + M[9] * p.y) + M[10] * p.z) + M[11]) * iw;
  return r;
}

/* Renderer.cpp:324-326. */
float mro_znear(const float P[16]) {
  int persp = P[15] == 0.0f;
  float zn = persp ? P[11] / (P[10] - 1.0f) : (P[11] + 1.0f) / P[10];
  return -zn;
}

typedef struct ostate {
  int W, H;
  float w, h;
  const float* P;
  int persp;
  float znear;
  vec3f light;
  int light_is_point;
  float ambient;
  int lighting, save_normals;
  /* current material (Renderer.cpp:227-234) */
  vec3f diffuse, specular, emissive;
  float shininess;
  const mro_texture* tex; /* non-NULL iff hastexture */
  /* outputs */
  float* image;
  float* depth;
  uint32_t* winner;
  float* normals;
  int64_t clipped;
} ostate;

/* Texel fetch: loadPPM's byte/255 (io.cpp:405-411) or a float texture. */
static vec3f texel(const mro_texture* t, int i, int j) {
  size_t o = ((size_t)i * (size_t)t->cols + (size_t)j) * 3;
  vec3f c;
  if (t->format == MR_TEX_RGB8) {
    const uint8_t* p = (const uint8_t*)t->texels + o;
    c.x = (float)p[0] / 255.0f;
    c.y = (float)p[1] / 255.0f;
    c.z = (float)p[2] / 255.0f;
  } else {
    const float* p = (const float*)t->texels + o;
    c.x = p[0]; c.y = p[1]; c.z = p[2];
  }
  return c;
}

/* int(fract(u) * n) of Renderer.cpp:274.  fract can round to exactly 1.0f
 * (u a tiny negative); the reference then reads row n (out of bounds), here
 * the index is clamped to n-1.  A NaN product maps to 0. */
static int tex_index(float u, int n) {
  float fr = u - floorf(u);
  float v = fr * (float)n;
  int i = (v >= 0.0f && v < 2147483648.0f) ? (int)v : 0;
  return i > n - 1 ? n - 1 : i;
}

static void paint(ostate* s, const vert* v0, const vert* v1, const vert* v2,
                  int world, uint32_t order);

/* clip, Renderer.cpp:121-129. */
static vert clip_edge(float z, const vert* a, const vert* b) {
  float d = b->p.z - a->p.z;
  float k = (fabsf(d) < 1e-6f) ? 0.5f : (z - a->p.z) / d;
  float k1 = 1.0f - k;
  vert v;
  v.p = v3_add(v3_scale(k, b->p), v3_scale(k1, a->p));
  v.n = v3_add(v3_scale(k, b->n), v3_scale(k1, a->n));
  v.uv.x = k * b->uv.x + k1 * a->uv.x;
  v.uv.y = k * b->uv.y + k1 * a->uv.y;
  return v;
}

/* clipTriangle, Renderer.cpp:131-161.  Emits sub-triangles in order; the
 * emitter receives sub-index 0 or 1. */
typedef void (*emit_fn)(void* user, const vert* a, const vert* b,
                        const vert* c, int sub);

static int clip_triangle(float z, vert v[3], emit_fn emit, void* user) {
  if (v[0].p.z > z && v[1].p.z > z && v[2].p.z > z) return 0;
  while (v[0].p.z < v[1].p.z || v[0].p.z < v[2].p.z) {
    vert t = v[0]; v[0] = v[1]; v[1] = t; /* swap(v0, v1) */
    t = v[0]; v[0] = v[2]; v[2] = t;      /* swap(v0, v2) */
  }
  if (v[1].p.z > z) {
    vert v02 = clip_edge(z, &v[0], &v[2]);
    vert v12 = clip_edge(z, &v[1], &v[2]);
    emit(user, &v02, &v12, &v[2], 0);
    return 1;
  } else if (v[2].p.z > z) {
    vert v01 = clip_edge(z, &v[0], &v[1]);
    vert v12 = clip_edge(z, &v[1], &v[2]);
    emit(user, &v01, &v[1], &v12, 0);
    return 1;
  } else {
    vert v01 = clip_edge(z, &v[0], &v[1]);
    vert v02 = clip_edge(z, &v[0], &v[2]);
    emit(user, &v01, &v[1], &v[2], 0);
    emit(user, &v01, &v[2], &v02, 1);
    return 2;
  }
}

typedef struct { ostate* s; uint32_t order; } paint_ctx;

static void emit_paint(void* user, const vert* a, const vert* b, const vert* c,
                       int sub) {
  paint_ctx* pc = (paint_ctx*)user;
  paint(pc->s, a, b, c, 0, pc->order + (uint32_t)sub);
}

/* paintTriangle, Renderer.cpp:163-309 (FAST_LIGHT variant, :288-299). */
static void paint(ostate* s, const vert* v0, const vert* v1, const vert* v2,
                  int world, uint32_t order) {
  const vert* vv[3] = {v0, v1, v2};
  if (world && (v0->p.z > s->znear || v1->p.z > s->znear || v2->p.z > s->znear)) {
    if (v0->p.z > s->znear && v1->p.z > s->znear && v2->p.z > s->znear) return;
    vert verts[3] = {*v0, *v1, *v2};
    paint_ctx pc = {s, order};
    s->clipped++;
    clip_triangle(s->znear, verts, emit_paint, &pc);
    return;
  }

  const float w = s->w, h = s->h;
  vec3f ndc[3];
  vec2f p[3];
  vec2f pmin = {1e30f, 1e30f}, pmax = {-1e30f, -1e30f};
  for (int i = 0; i < 3; i++) ndc[i] = htransform(s->P, vv[i]->p);
  for (int i = 0; i < 3; i++) {
    p[i].x = (1.0f + ndc[i].x) * (w / 2.0f);
    p[i].y = (1.0f - ndc[i].y) * (h / 2.0f);
    pmin.x = amin(pmin.x, p[i].x);
    pmin.y = amin(pmin.y, p[i].y);
    pmax.x = amax(pmax.x, p[i].x);
    pmax.y = amax(pmax.y, p[i].y);
  }
  if (pmax.x < 0 || pmax.y < 0 || pmin.x > w || pmin.y > h) return;

  /* a = (p0 - p1) ^ (p2 - p1), :205 */
  float ax = p[0].x - p[1].x, ay = p[0].y - p[1].y;
  float bx = p[2].x - p[1].x, by = p[2].y - p[1].y;
  float a = ax * by - ay * bx;
  if (a <= 0) return; /* back-face cull, :207-210 */
  float i2a = (a == 0) ? 0.0f : -1.0f / a;
  /* n1 = perp(p0 - p2) * i2a, n2 = perp(p1 - p0) * i2a, :215-216 */
  vec2f n1, n2;
  n1.x = -(p[0].y - p[2].y) * i2a;
  n1.y = (p[0].x - p[2].x) * i2a;
  n2.x = -(p[1].y - p[0].y) * i2a;
  n2.y = (p[1].x - p[0].x) * i2a;

  pmin.x = aclamp(pmin.x, 0.0f, w - 1);
  pmax.x = aclamp(pmax.x, 0.0f, w - 1);
  pmin.y = aclamp(pmin.y, 0.0f, h - 1);
  pmax.y = aclamp(pmax.y, 0.0f, h - 1);

  float zz[4] = {ndc[0].z, ndc[1].z, ndc[2].z, 1};
  float iz[4] = {-1.0f / vv[0]->p.z, -1.0f / vv[1]->p.z, -1.0f / vv[2]->p.z, 1};

  const int persp = s->persp;
  const int hasspecular = s->shininess != 0;
  const mro_texture* tex = s->tex;
  vec3f color = s->diffuse;
  float k[4] = {0, 0, 0, 0};

  for (float y = floorf(pmin.y) + 0.5f; y <= pmax.y + 0.5f; y++) {
    float ptx = floorf(pmin.x) + 0.5f;
    float e1 = n1.x * (ptx - p[2].x) + n1.y * (y - p[2].y);
    float e2 = n2.x * (ptx - p[0].x) + n2.y * (y - p[0].y);
    for (float x = ptx; x <= pmax.x + 0.5f; x++, e1 += n1.x, e2 += n2.x) {
      if (e1 < 0 || e2 < 0 || 1.0f - e1 - e2 < 0) continue;
      k[0] = 1.0f - e1 - e2;
      k[1] = e1;
      k[2] = e2;
      float z;
      if (persp) {
        z = 1.0f / (k[0] * iz[0] + k[1] * iz[1] + k[2] * iz[2]);
        k[0] *= iz[0] * z;
        k[1] *= iz[1] * z;
        k[2] *= iz[2] * z;
      } else {
        z = k[0] * zz[0] + k[1] * zz[1] + k[2] * zz[2] + k[3] * zz[3];
      }
      int i = (int)y, j = (int)x;
      size_t pix = (size_t)i * (size_t)s->W + (size_t)j;
      if (!(z < s->depth[pix])) continue;
      s->depth[pix] = z;
      if (s->winner) s->winner[pix] = order;

      if (tex) { /* :271-275 */
        float u = k[0] * vv[0]->uv.x + k[1] * vv[1]->uv.x + k[2] * vv[2]->uv.x;
        float v = k[0] * vv[0]->uv.y + k[1] * vv[1]->uv.y + k[2] * vv[2]->uv.y;
        color = texel(tex, tex_index(v, tex->rows), tex_index(u, tex->cols));
      }
      vec3f value = s->emissive;
      if (s->lighting) { /* :279-304 */
        vec3f position = v3_add(v3_add(v3_scale(k[0], vv[0]->p), v3_scale(k[1], vv[1]->p)),
                             v3_scale(k[2], vv[2]->p));
        vec3f lightdir = s->light_is_point ? v3_norm(v3_sub(s->light, position)) : s->light;
        vec3f normal = v3_add(v3_add(v3_scale(k[0], vv[0]->n), v3_scale(k[1], vv[1]->n)),
                           v3_scale(k[2], vv[2]->n));
        float diff = amax(0.0f, v3_dot(normal, lightdir)) / v3_len(normal) + s->ambient;
        value = v3_add(value, v3_scale(diff, color));
        if (hasspecular) {
          vec3f view = v3_norm(position);
          vec3f hv = v3_sub(lightdir, view);
          float sp = powf(amax(v3_dot(hv, normal), 0.0f) / (v3_len(hv) * v3_len(normal)),
                          s->shininess);
          value = v3_add(value, v3_scale(sp, s->specular));
        }
        if (s->save_normals && s->normals) {
          s->normals[pix * 3 + 0] = normal.x;
          s->normals[pix * 3 + 1] = normal.y;
          s->normals[pix * 3 + 2] = normal.z;
        }
      }
      s->image[pix * 3 + 0] = value.x;
      s->image[pix * 3 + 1] = value.y;
      s->image[pix * 3 + 2] = value.z;
    }
  }
}

/* savePPM quantization, io.cpp:356-361. */
static uint8_t quantize(float v) { return (uint8_t)aclamp(v * 255.0f, 0.0f, 255.0f); }

static _Thread_local int64_t g_last_clipped;
int64_t mro_last_clipped(void) { return g_last_clipped; }

static void setup_frame(ostate* s, const mr_frame* f) {
  memset(s, 0, sizeof(*s));
  s->W = f->width;
  s->H = f->height;
  s->w = (float)f->width;
  s->h = (float)f->height;
  s->P = f->projection;
  s->persp = f->projection[15] == 0.0f;
  s->znear = mro_znear(f->projection);
  s->light.x = f->light[0];
  s->light.y = f->light[1];
  s->light.z = f->light[2];
  s->light_is_point = f->light_is_point != 0;
  s->ambient = f->ambient;
  s->lighting = f->lighting != 0;
  s->save_normals = f->save_normals != 0;
}

static void set_material(ostate* s, const mr_draw* d, const mro_texture* tex,
                         int texturing) {
  s->diffuse.x = d->diffuse[0]; s->diffuse.y = d->diffuse[1]; s->diffuse.z = d->diffuse[2];
  s->specular.x = d->specular[0]; s->specular.y = d->specular[1]; s->specular.z = d->specular[2];
  s->emissive.x = d->emissive[0]; s->emissive.y = d->emissive[1]; s->emissive.z = d->emissive[2];
  s->shininess = d->shininess;
  s->tex = (texturing && tex && tex->rows > 0 && tex->cols > 0) ? tex : NULL;
}

/* Renderer::render, Renderer.cpp:311-332, with paintMesh :334-381. */
/* Renderer::getRangeImage (Renderer.cpp:388-415): each pixel's eye-space
 * point back-projected from the depth buffer through the projection's
 * diagonal terms; (0,0,0) where the depth lies past the far depth. */
static void range_image(const mr_frame* f, const float* depth, float* pts) {
  const float* P = f->projection;
  const int persp = P[15] == 0.0f;
  const float zfar = persp ? P[11] / (P[10] + 1) : (P[11] - 1) / P[10];
  const float fardepth = persp ? zfar : 1.0f;
  const float w = (float)f->width, h = (float)f->height;
  for (int i = 0; i < f->height; i++)
    for (int j = 0; j < f->width; j++) {
      const size_t k = (size_t)i * f->width + j;
      float x = 0.0f, y = 0.0f, z = 0.0f;
      if (!(depth[k] > fardepth)) {
        const float u = (j + 0.5f) / (w / 2) - 1;
        const float v = -(i + 0.5f) / (h / 2) + 1;
        z = -depth[k];
        x = -(u + P[2]) * z / P[0];
        y = -(v + P[6]) * z / P[5];
      }
      pts[k * 3 + 0] = x;
      pts[k * 3 + 1] = y;
      pts[k * 3 + 2] = z;
    }
}

int mro_render(const mr_frame* f, const mro_mesh* meshes, int32_t nmeshes,
               const mro_texture* textures, int32_t ntextures,
               const mr_outputs* out) {
  if (!f || !out || out->on_device || f->width <= 0 || f->height <= 0 ||
      f->ndraws < 0 || (f->ndraws > 0 && !f->draws))
    return MR_ERR_ARG;
  const size_t npix = (size_t)f->width * (size_t)f->height;
  uint64_t total = 0;
  for (int d = 0; d < f->ndraws; d++) {
    const mr_draw* dr = &f->draws[d];
    if (dr->mesh < 0 || dr->mesh >= nmeshes) return MR_ERR_HANDLE;
    if (dr->texture >= ntextures) return MR_ERR_HANDLE;
    total += 2u * (uint64_t)meshes[dr->mesh].ntriangles;
  }
  if (total >= 0xFFFFFFFFull) return MR_ERR_ARG;

  ostate s;
  setup_frame(&s, f);
  float* image = out->rgb ? out->rgb : (float*)malloc(npix * 3 * sizeof(float));
  float* depth = out->depth ? out->depth : (float*)malloc(npix * sizeof(float));
  if (!image || !depth) return MR_ERR_NOMEM;
  s.image = image;
  s.depth = depth;
  s.winner = out->winner;
  s.normals = out->normals;

  /* clear(), :113-119 */
  for (size_t i = 0; i < npix; i++) {
    image[i * 3 + 0] = f->background[0];
    image[i * 3 + 1] = f->background[1];
    image[i * 3 + 2] = f->background[2];
    depth[i] = 1e11f;
    if (s.winner) s.winner[i] = 0xFFFFFFFFu;
    if (s.normals && s.save_normals) {
      s.normals[i * 3 + 0] = 0.0f;
      s.normals[i * 3 + 1] = 0.0f;
      s.normals[i * 3 + 2] = 1.0f;
    }
  }

  uint32_t base = 0;
  int rc = MR_OK;
  for (int d = 0; d < f->ndraws && rc == MR_OK; d++) {
    const mr_draw* dr = &f->draws[d];
    const mro_mesh* m = &meshes[dr->mesh];
    const mro_texture* tex = dr->texture >= 0 ? &textures[dr->texture] : NULL;
    set_material(&s, dr, tex, f->texturing);
    /* PREMULT transforms, :340-349 */
    vec3f* ev = (vec3f*)malloc((size_t)(m->npositions > 0 ? m->npositions : 1) * sizeof(vec3f));
    vec3f* en = (vec3f*)malloc((size_t)(m->nnormals > 0 ? m->nnormals : 1) * sizeof(vec3f));
    if (!ev || !en) { free(ev); free(en); rc = MR_ERR_NOMEM; break; }
    for (int i = 0; i < m->npositions; i++) {
      vec3f p = {m->positions[i * 3], m->positions[i * 3 + 1], m->positions[i * 3 + 2]};
      ev[i] = xform(dr->modelview, p);
    }
    for (int i = 0; i < m->nnormals; i++) {
      vec3f n = {m->normals[i * 3], m->normals[i * 3 + 1], m->normals[i * 3 + 2]};
      en[i] = xform(dr->normalmat, n);
    }
    const int has_uv = m->nuvs > 0 && m->uvs && m->uv_indices;
    for (int t = 0; t < m->ntriangles; t++) { /* :351-380 */
      vert v[3];
      for (int c = 0; c < 3; c++) {
        int ip = m->indices[t * 3 + c], in = m->normal_indices[t * 3 + c];
        if (ip < 0 || ip >= m->npositions || in < 0 || in >= m->nnormals) {
          rc = MR_ERR_INDEX;
          break;
        }
        v[c].p = ev[ip];
        v[c].n = en[in];
        v[c].uv.x = 0.0f;
        v[c].uv.y = 0.0f;
        if (has_uv) {
          int iu = m->uv_indices[t * 3 + c];
          if (iu < 0 || iu >= m->nuvs) { rc = MR_ERR_INDEX; break; }
          v[c].uv.x = m->uvs[iu * 2];
          v[c].uv.y = m->uvs[iu * 2 + 1];
        }
      }
      if (rc != MR_OK) break;
      /* paintMesh draws with world = true (:373-379); a draw flagged
         MR_DRAW_NO_NEAR_CLIP holds paintTriangle(a, b, c, false) triangles */
      paint(&s, &v[0], &v[1], &v[2], (dr->flags & MR_DRAW_NO_NEAR_CLIP) ? 0 : 1,
            base + 2u * (uint32_t)t);
    }
    free(ev);
    free(en);
    base += 2u * (uint32_t)m->ntriangles;
  }
  g_last_clipped = s.clipped;

  if (rc == MR_OK && out->rgb8)
    for (size_t i = 0; i < npix * 3; i++) out->rgb8[i] = quantize(image[i]);
  if (rc == MR_OK && out->points) range_image(f, depth, out->points);
  if (!out->rgb) free(image);
  if (!out->depth) free(depth);
  return rc;
}

int mro_paint_triangle(const mr_frame* f, const mr_draw* dr,
                       const mro_texture* tex, const float vtx[24],
                       uint32_t order, float* image, float* depth,
                       uint32_t* winner, float* normals) {
  if (!f || !dr || !vtx || !image || !depth) return MR_ERR_ARG;
  ostate s;
  setup_frame(&s, f);
  set_material(&s, dr, tex, f->texturing);
  s.image = image;
  s.depth = depth;
  s.winner = winner;
  s.normals = normals;
  vert v[3];
  for (int c = 0; c < 3; c++) {
    const float* q = vtx + c * 8;
    v[c].p.x = q[0]; v[c].p.y = q[1]; v[c].p.z = q[2];
    v[c].n.x = q[3]; v[c].n.y = q[4]; v[c].n.z = q[5];
    v[c].uv.x = q[6]; v[c].uv.y = q[7];
  }
  paint(&s, &v[0], &v[1], &v[2], 1, order);
  return (int)s.clipped;
}

typedef struct { float* out; int n; } collect_ctx;

static void emit_collect(void* user, const vert* a, const vert* b,
                         const vert* c, int sub) {
  collect_ctx* cc = (collect_ctx*)user;
  const vert* vs[3] = {a, b, c};
  float* o = cc->out + sub * 24;
  for (int i = 0; i < 3; i++) {
    o[i * 8 + 0] = vs[i]->p.x; o[i * 8 + 1] = vs[i]->p.y; o[i * 8 + 2] = vs[i]->p.z;
    o[i * 8 + 3] = vs[i]->n.x; o[i * 8 + 4] = vs[i]->n.y; o[i * 8 + 5] = vs[i]->n.z;
    o[i * 8 + 6] = vs[i]->uv.x; o[i * 8 + 7] = vs[i]->uv.y;
  }
  cc->n = sub + 1;
}

int mro_clip_triangle(float znear, const float vtx[24], float out[48]) {
  vert v[3];
  for (int c = 0; c < 3; c++) {
    const float* q = vtx + c * 8;
    v[c].p.x = q[0]; v[c].p.y = q[1]; v[c].p.z = q[2];
    v[c].n.x = q[3]; v[c].n.y = q[4]; v[c].n.z = q[5];
    v[c].uv.x = q[6]; v[c].uv.y = q[7];
  }
  collect_ctx cc = {out, 0};
  return clip_triangle(znear, v, emit_collect, &cc);
}
