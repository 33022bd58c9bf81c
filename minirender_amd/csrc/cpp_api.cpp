// cpp_api.cpp -- the minirender C++ API (include/minirender/*.h) over the
// C-ABI of this library.  Host code only (the HIP runtime API is used for the
// Renderer's device framebuffers).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "minirender/Renderer.h"
#include "minirender/io.h"
#include "minirender/primitives.h"

namespace minirender {

static_assert(sizeof(Vec3) == 12 && sizeof(Vec2) == 8, "Vec types must be packed floats");

// ---------------------------------------------------------------------------
// Matrix4::inverse: cofactor expansion in double, rounded to float.
Matrix4 Matrix4::inverse() const {
  double m[16], inv[16];
  for (int i = 0; i < 16; i++) m[i] = a[i / 4][i % 4];
  inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] +
           m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] -
           m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] +
           m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] -
            m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] -
           m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] +
           m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] -
           m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] +
            m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] +
           m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] -
           m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] +
            m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] -
            m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] -
           m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] +
           m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] -
            m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] +
            m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  const double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
  Matrix4 r;
  const double id = det != 0 ? 1.0 / det : 0.0;
  for (int i = 0; i < 16; i++) r.a[i / 4][i % 4] = (float)(inv[i] * id);
  return r;
}

// ---------------------------------------------------------------------------
// Scene graph (Scene.cpp:13-63).
void SceneNode::collectShapes(Array<Renderable>& list, const Matrix4& xform) {
  const Matrix4 tr = xform * transform;
  for (auto& n : children) n->collectShapes(list, tr);
}

BBox SceneNode::getBbox(const Matrix4& xform) const {
  BBox b;
  for (auto& n : children) b += n->getBbox(xform * transform);
  return b;
}

void TriMesh::collectShapes(Array<Renderable>& list, const Matrix4& xform) {
  const Matrix4 tr = xform * transform;
  list.push_back(Renderable(this, tr));
  for (auto& n : children) n->collectShapes(list, tr);
}

BBox TriMesh::getBbox(const Matrix4& xform) const {
  BBox b;
  const Matrix4 m = xform * transform;
  for (const Vec3& p : vertices) b += m * p;
  for (auto& n : children) b += n->getBbox(m);
  return b;
}

void TriMesh::applyTransform() {
  for (Vec3& p : vertices) p = transform * p;
  const Matrix4 it = transform.inverse().transposed();
  for (Vec3& n : normals) n = (it % n).normalized();
  transform = Matrix4::identity();
}

// ---------------------------------------------------------------------------
// Procedural meshes with the reference's vertex and index order
// (primitives.cpp:7-230).
static const float kPi = 3.14159265358979323846f;

Shared<TriMesh> createCube(float size) {
  auto m = std::make_shared<TriMesh>();
  const float h = size * 0.5f;
  struct Face { Vec3 v[4], n; };
  const Face faces[6] = {
      {{Vec3(h, -h, -h), Vec3(h, -h, h), Vec3(h, h, h), Vec3(h, h, -h)}, Vec3(1, 0, 0)},
      {{Vec3(-h, -h, h), Vec3(-h, -h, -h), Vec3(-h, h, -h), Vec3(-h, h, h)}, Vec3(-1, 0, 0)},
      {{Vec3(-h, h, -h), Vec3(h, h, -h), Vec3(h, h, h), Vec3(-h, h, h)}, Vec3(0, 1, 0)},
      {{Vec3(-h, -h, h), Vec3(h, -h, h), Vec3(h, -h, -h), Vec3(-h, -h, -h)}, Vec3(0, -1, 0)},
      {{Vec3(h, -h, h), Vec3(-h, -h, h), Vec3(-h, h, h), Vec3(h, h, h)}, Vec3(0, 0, 1)},
      {{Vec3(-h, -h, -h), Vec3(h, -h, -h), Vec3(h, h, -h), Vec3(-h, h, -h)}, Vec3(0, 0, -1)}};
  const Vec2 uv[4] = {Vec2(0, 0), Vec2(1, 0), Vec2(1, 1), Vec2(0, 1)};
  for (const Face& f : faces) {
    const int base = (int)m->vertices.size();
    for (int k = 0; k < 4; k++) {
      m->vertices.push_back(f.v[k]);
      m->normals.push_back(f.n);
      m->texcoords.push_back(uv[k]);
    }
    const int tri[6] = {base, base + 2, base + 1, base, base + 3, base + 2};
    for (int k : tri) {
      m->indices.push_back(k);
      m->normalsI.push_back(k);
    }
  }
  m->material = std::make_shared<Material>();
  return m;
}

Shared<TriMesh> createCylinder(float radius, float height, int segments, int heightSegments,
                               bool caps) {
  auto m = std::make_shared<TriMesh>();
  if (segments < 3) segments = 3;
  if (heightSegments < 1) heightSegments = 1;
  const float halfH = height * 0.5f, twoPi = 2.0f * kPi;
  for (int hh = 0; hh <= heightSegments; ++hh) {
    const float v = (float)hh / (float)heightSegments;
    const float z = -halfH + v * height;
    for (int r = 0; r < segments; ++r) {
      const float u = (float)r / (float)segments;
      const float ang = u * twoPi;
      const float x = radius * std::cos(ang), y = radius * std::sin(ang);
      m->vertices.push_back(Vec3(x, y, z));
      m->normals.push_back(Vec3(x, y, 0).normalized());
      m->texcoords.push_back(Vec2(u, v));
    }
  }
  auto tri = [&](int a, int b, int c) {
    for (int k : {a, b, c}) {
      m->indices.push_back(k);
      m->normalsI.push_back(k);
    }
  };
  for (int hh = 0; hh < heightSegments; ++hh)
    for (int r = 0; r < segments; ++r) {
      const int i0 = hh * segments + r, i1 = hh * segments + (r + 1) % segments;
      const int i2 = (hh + 1) * segments + r, i3 = (hh + 1) * segments + (r + 1) % segments;
      tri(i0, i1, i2);
      tri(i1, i3, i2);
    }
  if (caps) {
    for (int side = 0; side < 2; side++) {  // bottom, then top
      const int ring = (int)m->vertices.size();
      const int orig = side == 0 ? 0 : heightSegments * segments;
      const float nz = side == 0 ? -1.0f : 1.0f;
      for (int r = 0; r < segments; ++r) {
        const Vec3 p = m->vertices[orig + r];
        m->vertices.push_back(p);
        m->normals.push_back(Vec3(0, 0, nz));
        m->texcoords.push_back(Vec2((p.x / radius + 1.0f) * 0.5f, (p.y / radius + 1.0f) * 0.5f));
      }
      const int center = (int)m->vertices.size();
      m->vertices.push_back(Vec3(0, 0, side == 0 ? -halfH : halfH));
      m->normals.push_back(Vec3(0, 0, nz));
      m->texcoords.push_back(Vec2(0.5f, 0.5f));
      for (int r = 0; r < segments; ++r) {
        const int a = ring + r, b = ring + (r + 1) % segments;
        if (side == 0) tri(center, b, a); else tri(center, a, b);
      }
    }
  }
  m->material = std::make_shared<Material>();
  return m;
}

Shared<TriMesh> createSphere(float radius, int latSegments, int longSegments) {
  auto m = std::make_shared<TriMesh>();
  if (latSegments < 2) latSegments = 2;
  if (longSegments < 3) longSegments = 3;
  m->vertices.push_back(Vec3(0, radius, 0));
  m->normals.push_back(Vec3(0, 1, 0));
  m->texcoords.push_back(Vec2(0.5f, 0));
  const int first = 1;
  for (int lat = 1; lat < latSegments; ++lat) {
    const float v = (float)lat / (float)latSegments;
    const float phi = v * kPi, sp = std::sin(phi), cp = std::cos(phi);
    for (int lon = 0; lon < longSegments; ++lon) {
      const float u = (float)lon / (float)longSegments;
      const float th = u * 2.0f * kPi;
      const Vec3 p(radius * sp * std::cos(th), radius * cp, radius * sp * std::sin(th));
      m->vertices.push_back(p);
      m->normals.push_back(p.normalized());
      m->texcoords.push_back(Vec2(u, 1.0f - v));
    }
  }
  const int bottom = (int)m->vertices.size();
  m->vertices.push_back(Vec3(0, -radius, 0));
  m->normals.push_back(Vec3(0, -1, 0));
  m->texcoords.push_back(Vec2(0.5f, 1));
  auto tri = [&](int a, int b, int c) {
    for (int k : {a, b, c}) {
      m->indices.push_back(k);
      m->normalsI.push_back(k);
    }
  };
  for (int lon = 0; lon < longSegments; ++lon)
    tri(0, first + lon, first + (lon + 1) % longSegments);
  const int rings = latSegments - 1;
  for (int lat = 0; lat + 1 < rings; ++lat) {
    const int r1 = first + lat * longSegments, r2 = r1 + longSegments;
    for (int lon = 0; lon < longSegments; ++lon) {
      const int a = r1 + lon, b = r1 + (lon + 1) % longSegments;
      const int c = r2 + lon, d = r2 + (lon + 1) % longSegments;
      tri(a, b, c);
      tri(b, d, c);
    }
  }
  const int last = first + (rings - 1) * longSegments;
  for (int lon = 0; lon < longSegments; ++lon)
    tri(bottom, last + lon, last + (lon + 1) % longSegments);
  m->material = std::make_shared<Material>();
  return m;
}

// ---------------------------------------------------------------------------
// PPM (io.cpp:337-415).
bool savePPM(const Array2<Vec3>& image, const std::string& filename) {
  FILE* f = std::fopen(filename.c_str(), "wb");
  if (!f) return false;
  std::fprintf(f, "P6\n%d %d\n255\n", image.cols(), image.rows());
  std::vector<unsigned char> row((size_t)image.cols() * 3);
  auto q = [](float v) {
    const float s = v * 255.0f;
    const float c = s > 0.0f ? (s < 255.0f ? s : 255.0f) : 0.0f;
    return (unsigned char)c;
  };
  for (int i = 0; i < image.rows(); i++) {
    for (int j = 0; j < image.cols(); j++) {
      const Vec3& v = image(i, j);
      row[j * 3] = q(v.x);
      row[j * 3 + 1] = q(v.y);
      row[j * 3 + 2] = q(v.z);
    }
    std::fwrite(row.data(), 1, row.size(), f);
  }
  return std::fclose(f) == 0;
}

Array2<Vec3> loadPPM(const std::string& filename) {
  Array2<Vec3> img;
  FILE* f = std::fopen(filename.c_str(), "rb");
  if (!f) return img;
  char magic[3] = {0};
  int w = 0, h = 0, maxv = 0;
  auto skip = [&]() {  // whitespace and comments
    int c;
    while ((c = std::fgetc(f)) != EOF) {
      if (c == '#') { while ((c = std::fgetc(f)) != EOF && c != '\n') {} }
      else if (c != ' ' && c != '\n' && c != '\r' && c != '\t') { std::ungetc(c, f); break; }
    }
  };
  if (std::fscanf(f, "%2s", magic) != 1 || std::strcmp(magic, "P6") != 0) { std::fclose(f); return img; }
  skip(); if (std::fscanf(f, "%d", &w) != 1) { std::fclose(f); return img; }
  skip(); if (std::fscanf(f, "%d", &h) != 1) { std::fclose(f); return img; }
  skip(); if (std::fscanf(f, "%d", &maxv) != 1 || maxv != 255) { std::fclose(f); return img; }
  std::fgetc(f);
  img.resize(h, w);
  std::vector<unsigned char> row((size_t)w * 3);
  for (int i = 0; i < h; i++) {
    if (std::fread(row.data(), 1, row.size(), f) != row.size()) break;
    for (int j = 0; j < w; j++)
      img(i, j) = Vec3((float)row[j * 3], (float)row[j * 3 + 1], (float)row[j * 3 + 2]) / 255.0f;
  }
  std::fclose(f);
  return img;
}

// ---------------------------------------------------------------------------
// Projections through the C-ABI helpers (one implementation).
static Matrix4 fromRows(const float* p) {
  return Matrix4(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], p[10], p[11],
                 p[12], p[13], p[14], p[15]);
}
Matrix4 projectionOrtho(float l, float r, float b, float t, float n, float f) {
  float o[16];
  mr_projection_ortho(l, r, b, t, n, f, o);
  return fromRows(o);
}
Matrix4 projectionPerspective(float l, float r, float b, float t, float n, float f) {
  float o[16];
  mr_projection_perspective(l, r, b, t, n, f, o);
  return fromRows(o);
}
Matrix4 projectionFrustum(float fov, float aspect, float n, float f) {
  float o[16];
  mr_projection_frustum(fov, aspect, n, f, o);
  return fromRows(o);
}
Matrix4 projectionFrustumH(float fov, float aspect, float n, float f) {
  float o[16];
  mr_projection_frustum_h(fov, aspect, n, f, o);
  return fromRows(o);
}
Matrix4 projectionOrtho(float fov, float aspect, float n, float f) {
  float o[16];
  mr_projection_ortho_fov(fov, aspect, n, f, o);
  return fromRows(o);
}
Matrix4 projectionCV(const Matrix4& K, float w, float h, float n, float f) {
  float o[16];
  mr_projection_cv(K.data(), w, h, n, f, o);
  return fromRows(o);
}

// ---------------------------------------------------------------------------
// Renderer.
Renderer::Renderer() {  // Renderer.cpp:85-98
  setSize(800, 600);
  _projection = projectionOrtho(-40, 40, -30, 30, 50, 120);
  _light = Vec3(-0.15f, 0.6f, 1).normalized();
  _ambient = 0.1f;
  _defmaterial = std::make_shared<Material>();
  _material = _defmaterial;
  _bgcolor = Vec3(0, 0, 0);
}

Renderer::~Renderer() {
  freeDeviceOutputs();
  if (_ctx) mr_ctx_destroy(_ctx);
}

// Device framebuffers of render(): the frame stays on the GPU and each
// getter copies back only its own image.
void Renderer::freeDeviceOutputs() {
  for (void*& p : _devOut)
    if (p) {
      hipFree(p);
      p = nullptr;
    }
  _devPixels = 0;
}

bool Renderer::ensureDeviceOutputs() {
  const size_t n = (size_t)_image.cols() * _image.rows();
  if (_devPixels == n && _devOut[0]) return true;
  freeDeviceOutputs();
  const size_t bytes[kOutputs] = {n * 3, n * 4, n * 4, n * 12, n * 12};
  for (int k = 0; k < kOutputs; k++)
    if (hipMalloc(&_devOut[k], bytes[k] ? bytes[k] : 16) != hipSuccess) {
      (void)hipGetLastError();
      freeDeviceOutputs();
      _status = MR_ERR_NOMEM;
      return false;
    }
  _devPixels = n;
  return true;
}

// Copy device output k to the host image if the host copy is stale.
void Renderer::fetch(int k) {
  if (_fresh[k] || !_devOut[k] || _devPixels == 0) return;
  const size_t n = _devPixels;
  void* dst[kOutputs] = {_rgb8.data(), _depth.data(), _winners.data(), _image.data(),
                         _pnormals.data()};
  const size_t bytes[kOutputs] = {n * 3, n * 4, n * 4, n * 12, n * 12};
  if (hipMemcpy(dst[k], _devOut[k], bytes[k], hipMemcpyDeviceToHost) != hipSuccess) {
    (void)hipGetLastError();
    _status = MR_ERR_HIP;
    return;
  }
  _fresh[k] = true;
}

mr_ctx* Renderer::context() {
  if (!_ctx) _status = mr_ctx_create(0, 0, &_ctx);
  return _ctx;
}

void Renderer::setSize(int w, int h) {  // Renderer.cpp:100-106
  _image.resize(h, w);
  _depth.resize(h, w);
  _pnormals.resize(h, w);
  _rgb8.assign((size_t)w * h * 3, 0);
  _winners.assign((size_t)w * h, 0xFFFFFFFFu);
  clear();
}

void Renderer::clear() {  // Renderer.cpp:113-119
  for (bool& f : _fresh) f = true;  // the host images below are the frame
  _immediate.clear();
  _immediateMeshes.clear();
  _image.set(_bgcolor);
  _depth.set(1e11f);
  if (_saveNormals) _pnormals.set(Vec3(0, 0, 1));
  auto q = [](float v) {
    const float s = v * 255.0f;
    return (uint8_t)(s > 0.0f ? (s < 255.0f ? s : 255.0f) : 0.0f);
  };
  for (size_t i = 0; i < _winners.size(); i++) {
    _rgb8[i * 3] = q(_bgcolor.x);
    _rgb8[i * 3 + 1] = q(_bgcolor.y);
    _rgb8[i * 3 + 2] = q(_bgcolor.z);
    _winners[i] = 0xFFFFFFFFu;
  }
}

// Hash of an array's size and every byte, eight bytes per step.
template <class T>
static void hashArray(uint64_t& h, const Array<T>& v) {
  const size_t bytes = v.size() * sizeof(T);
  const unsigned char* p = (const unsigned char*)v.data();
  auto mix = [&](uint64_t w) {
    h = (h ^ w) * 0x9E3779B97F4A7C15ull;
    h ^= h >> 29;
  };
  mix(bytes);
  size_t i = 0;
  for (; i + 8 <= bytes; i += 8) {
    uint64_t w;
    std::memcpy(&w, p + i, 8);
    mix(w);
  }
  if (i < bytes) {
    uint64_t w = 0;
    std::memcpy(&w, p + i, bytes - i);
    mix(w);
  }
}

template <class T>
static bool sameBytes(const Array<T>& a, const Array<T>& b) {
  if (a.shares(b)) return true;
  return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0);
}

static bool usesTexcoords(const TriMesh* m) {
  return !m->texcoords.empty() && m->texcoordsI.size() == m->indices.size();
}

// Storage ids and modification versions of a mesh's six arrays.
static void meshStamps(const TriMesh* m, uint64_t out[12]) {
  const uint64_t v[12] = {m->vertices.id(),  m->vertices.version(), m->normals.id(),
                          m->normals.version(), m->texcoords.id(), m->texcoords.version(),
                          m->indices.id(),   m->indices.version(), m->normalsI.id(),
                          m->normalsI.version(), m->texcoordsI.id(), m->texcoordsI.version()};
  for (int i = 0; i < 12; i++) out[i] = v[i];
}

int32_t Renderer::meshHandle(TriMesh* mesh) {
  // All reads below go through a const view: they leave the arrays'
  // modification versions alone.
  const TriMesh* m = mesh;
  MeshEntry& e = _meshes[m];
  uint64_t stamps[12];
  meshStamps(m, stamps);
  if (e.handle >= 0 && std::memcmp(stamps, e.stamps, sizeof(stamps)) == 0) {
    if (_sharing) _frameMeshes.push_back(m);  // untouched since its upload
    return e.handle;
  }
  uint64_t fp = 1469598103934665603ull;
  hashArray(fp, m->vertices);
  hashArray(fp, m->normals);
  hashArray(fp, m->texcoords);
  hashArray(fp, m->indices);
  hashArray(fp, m->normalsI);
  hashArray(fp, m->texcoordsI);
  std::memcpy(e.stamps, stamps, sizeof(stamps));
  if (e.handle >= 0 && e.fingerprint == fp) {
    if (_sharing) _frameMeshes.push_back(m);
    return e.handle;
  }
  mr_ctx* c = context();
  if (!c) return -1;
  if (e.handle >= 0) {
    mr_mesh_release(c, e.handle);
    _handleMesh.erase(e.handle);
  }
  const int ntri = (int)(m->indices.size() / 3);
  const bool nI = m->normalsI.size() == m->indices.size();
  const bool uv = usesTexcoords(m);
  // A mesh of the scene being described whose attribute arrays hold the
  // same bytes (the per-material meshes of one model, io.cpp:313-315)
  // lends them: one resident copy and one vertex transform per frame.
  int32_t base = -1;
  if (_sharing)
    for (const TriMesh* o : _frameMeshes) {
      auto it = _meshes.find(o);
      if (o == m || it == _meshes.end() || it->second.handle < 0) continue;
      if ((!uv || usesTexcoords(o)) && sameBytes(o->vertices, m->vertices) &&
          sameBytes(o->normals, m->normals) && sameBytes(o->texcoords, m->texcoords)) {
        base = it->second.handle;
        break;
      }
    }
  int32_t h = -1;
  if (base >= 0) {
    _status = mr_mesh_upload_shared(c, base, m->indices.data(),
                                    nI ? m->normalsI.data() : m->indices.data(),
                                    uv ? m->texcoordsI.data() : nullptr, ntri, &h);
  } else {
    _status = mr_mesh_upload(
        c, (const float*)m->vertices.data(), (int32_t)m->vertices.size(),
        (const float*)m->normals.data(), (int32_t)m->normals.size(),
        uv ? (const float*)m->texcoords.data() : nullptr, uv ? (int32_t)m->texcoords.size() : 0,
        m->indices.data(), nI ? m->normalsI.data() : m->indices.data(),
        uv ? m->texcoordsI.data() : nullptr, ntri, &h);
  }
  e.handle = _status == MR_OK ? h : -1;
  e.fingerprint = fp;
  e.mesh = m;
  if (e.handle >= 0) {
    _handleMesh[e.handle] = m;
    if (_sharing) _frameMeshes.push_back(m);
  }
  return e.handle;
}

int32_t Renderer::textureHandle(const Material* mat) {
  if (mat->texture.rows() <= 0) return -1;
  uint64_t fp = 1469598103934665603ull;
  const size_t n = mat->texture.size();
  fp ^= (uint64_t)n * 1099511628211ull ^ (uint64_t)(uintptr_t)mat->texture.data();
  const size_t step = n > 4096 ? n / 4096 : 1;
  for (size_t i = 0; i < n; i += step) {
    const Vec3& v = mat->texture.data()[i];
    uint32_t b[3];
    std::memcpy(b, &v, 12);
    fp = (fp ^ b[0] ^ ((uint64_t)b[1] << 21) ^ ((uint64_t)b[2] << 42)) * 1099511628211ull;
  }
  TexEntry& e = _textures[mat];
  if (e.handle >= 0 && e.fingerprint == fp) return e.handle;
  mr_ctx* c = context();
  if (!c) return -1;
  if (e.handle >= 0) mr_texture_release(c, e.handle);
  int32_t h = -1;
  _status = mr_texture_upload(c, mat->texture.data(), MR_TEX_RGB32F, mat->texture.rows(),
                              mat->texture.cols(), &h);
  e.handle = _status == MR_OK ? h : -1;
  e.fingerprint = fp;
  return e.handle;
}

void Renderer::appendDraw(TriMesh* mesh, const Matrix4& transform, const Material& mat,
                          std::vector<mr_draw>& out) {
  const int32_t h = meshHandle(mesh);
  if (h < 0) return;
  mr_draw d{};
  d.mesh = h;
  d.texture = textureHandle(&mat);
  const Matrix4 mv = _view * transform;       // Renderer.cpp:337
  const Matrix4 nm = mv.inverse().t();        // :338
  std::memcpy(d.modelview, mv.data(), sizeof(d.modelview));
  std::memcpy(d.normalmat, nm.data(), sizeof(d.normalmat));
  const float* dif = &mat.diffuse.x;
  const float* spe = &mat.specular.x;
  const float* emi = &mat.emissive.x;
  for (int i = 0; i < 3; i++) {
    d.diffuse[i] = dif[i];
    d.specular[i] = spe[i];
    d.emissive[i] = emi[i];
  }
  d.shininess = mat.shininess;
  out.push_back(d);
}

void Renderer::fillFrame(mr_frame& f, const std::vector<mr_draw>& draws) {
  f = mr_frame{};
  f.width = _image.cols();
  f.height = _image.rows();
  std::memcpy(f.projection, _projection.data(), sizeof(f.projection));
  f.background[0] = _bgcolor.x;
  f.background[1] = _bgcolor.y;
  f.background[2] = _bgcolor.z;
  const Vec3 L = _lightIsPoint ? _view * _light : _light.normalized();  // :317-320
  f.light[0] = L.x;
  f.light[1] = L.y;
  f.light[2] = L.z;
  f.light_is_point = _lightIsPoint;
  f.ambient = _ambient;
  f.lighting = _lighting;
  f.texturing = _texturing;
  f.save_normals = _saveNormals;
  f.ndraws = (int32_t)draws.size();
  f.draws = draws.data();
}

const mr_frame& Renderer::describeFrame() {
  _draws.clear();
  if (_scene) {
    _ambient = _scene->ambientLight;  // :322
    Array<Renderable> list;
    _scene->collectShapes(list, Matrix4::identity());  // :315
    _sharing = true;  // the scene's meshes stay alive while it is described
    _frameMeshes.clear();
    for (const Renderable& r : list) {
      const Material& mat = r.mesh->material ? *r.mesh->material : *_defmaterial;  // :336
      appendDraw(r.mesh, r.transform, mat, _draws);
    }
    _sharing = false;
    _frameMeshes.clear();
  }
  fillFrame(_frame, _draws);
  return _frame;
}

const TriMesh* Renderer::meshOfHandle(int handle) const {
  auto it = _handleMesh.find(handle);
  return it == _handleMesh.end() ? nullptr : it->second;
}

void Renderer::render() {
  _immediate.clear();
  _immediateMeshes.clear();
  const mr_frame& f = describeFrame();
  mr_ctx* c = context();
  if (!c || !ensureDeviceOutputs()) return;
  mr_outputs o{};
  o.rgb8 = (uint8_t*)_devOut[kRGB8];
  o.depth = (float*)_devOut[kDepth];
  o.winner = (uint32_t*)_devOut[kWinner];
  o.rgb = (float*)_devOut[kRGB];
  o.normals = _saveNormals ? (float*)_devOut[kNormals] : nullptr;
  o.on_device = 1;
  _status = mr_render(c, &f, &o);
  for (int k = 0; k < kOutputs; k++) _fresh[k] = false;
  if (!_saveNormals) _fresh[kNormals] = true;  // not produced: keeps its last contents
}

void Renderer::paintMesh(TriMesh* mesh, const Matrix4& transform) {
  _material = mesh->material ? mesh->material : _defmaterial;  // :336
  appendDraw(mesh, transform, *_material, _immediate);
}

void Renderer::paintTriangle(const Vertex& a, const Vertex& b, const Vertex& c, bool world) {
  // Eye-space vertices: drawn with identity modelview through the same
  // pipeline; world = false skips the near-plane test and clip
  // (Renderer.cpp:169).
  auto m = std::make_shared<TriMesh>();
  for (const Vertex* v : {&a, &b, &c}) {
    m->vertices.push_back(v->position);
    m->normals.push_back(v->normal);
    m->texcoords.push_back(v->uv);
  }
  m->indices = {0, 1, 2};
  m->normalsI = {0, 1, 2};
  m->texcoordsI = {0, 1, 2};
  _immediateMeshes.push_back(m);
  const Matrix4 saved = _view;
  _view = Matrix4::identity();
  const size_t before = _immediate.size();
  appendDraw(m.get(), Matrix4::identity(), *_material, _immediate);
  if (!world && _immediate.size() > before) _immediate.back().flags = MR_DRAW_NO_NEAR_CLIP;
  _view = saved;
}

void Renderer::flushImmediate() {
  if (_immediate.empty()) return;
  mr_ctx* c = context();
  if (!c) return;
  for (int k = 0; k < kOutputs; k++) fetch(k);  // composited on the host
  const int W = _image.cols(), H = _image.rows();
  const size_t n = (size_t)W * H;
  std::vector<Vec3> rgb(n), nrm(_saveNormals ? n : 0);
  std::vector<float> depth(n);
  std::vector<uint8_t> rgb8(n * 3);
  std::vector<uint32_t> win(n);
  mr_frame f;
  fillFrame(f, _immediate);
  mr_outputs o{};
  o.rgb = (float*)rgb.data();
  o.depth = depth.data();
  o.rgb8 = rgb8.data();
  o.winner = win.data();
  o.normals = _saveNormals ? (float*)nrm.data() : nullptr;
  _status = mr_render(c, &f, &o);
  if (_status == MR_OK) {
    // Immediate draws come after the frame's: they win only where strictly
    // closer (Renderer.cpp:267).  Their draw-order indices continue the
    // frame's.
    uint32_t base = 0;
    for (const mr_draw& d : _draws) {
      const TriMesh* m = meshOfHandle(d.mesh);
      if (m) base += 2u * (uint32_t)(m->indices.size() / 3);
    }
    for (size_t i = 0; i < n; i++) {
      if (!(depth[i] < _depth.data()[i])) continue;
      _depth.data()[i] = depth[i];
      _image.data()[i] = rgb[i];
      std::memcpy(&_rgb8[i * 3], &rgb8[i * 3], 3);
      _winners[i] = base + win[i];
      if (_saveNormals) _pnormals.data()[i] = nrm[i];
    }
  }
  _immediate.clear();
  _immediateMeshes.clear();
}

Array2<float> Renderer::getDepth() { flushImmediate(); fetch(kDepth); return _depth; }
Array2<Vec3> Renderer::getImage() { flushImmediate(); fetch(kRGB); return _image; }
Array2<Vec3> Renderer::getNormalsImage() { flushImmediate(); fetch(kNormals); return _pnormals; }
std::vector<uint8_t> Renderer::getImageRGB8() { flushImmediate(); fetch(kRGB8); return _rgb8; }
std::vector<uint32_t> Renderer::getWinners() { flushImmediate(); fetch(kWinner); return _winners; }

Array2<Vec3> Renderer::getRangeImage() {  // Renderer.cpp:388-415
  flushImmediate();
  fetch(kDepth);
  const Matrix4& P = _projection;
  const bool persp = P(3, 3) == 0;
  float zn = 0;
  mr_znear(P.data(), &zn);
  const float zfar = persp ? P(2, 3) / (P(2, 2) + 1) : (P(2, 3) - 1) / P(2, 2);
  const float fardepth = persp ? zfar : 1.0f;
  (void)zn;
  Array2<Vec3> pts(_depth.rows(), _depth.cols());
  const float w = (float)_image.cols(), h = (float)_image.rows();
  for (int i = 0; i < _depth.rows(); i++)
    for (int j = 0; j < _depth.cols(); j++) {
      if (_depth(i, j) > fardepth) {
        pts(i, j) = Vec3(0, 0, 0);
      } else {
        const float u = (j + 0.5f) / (w / 2) - 1;
        const float v = -(i + 0.5f) / (h / 2) + 1;
        const float z = -_depth(i, j);
        const float x = -(u + P(0, 2)) * z / P(0, 0);
        const float y = -(v + P(1, 2)) * z / P(1, 1);
        pts(i, j) = Vec3(x, y, z);
      }
    }
  return pts;
}

}  // namespace minirender
