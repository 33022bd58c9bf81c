// ref_idiom.cpp -- a caller written the way the reference's samples drive
// minirender (samples/bench.cpp:15-62 and 84-135, samples/main.cpp:108-134):
// `Shared<T> x = new T()`, `<<` appends, index arrays shared by assignment,
// `scene->children << shape`, a raw `Scene*` handed to setScene, and a frame
// loop of setView + render.  Against the reference only the include path and
// the namespace differ (and the random numbers, which come from a small
// generator here instead of asl::Random).
//
//   ref_idiom cpu            container semantics the idiom relies on (shared
//                            storage, dup / clone, one object wrapped twice),
//                            the scene it builds and the console painter
//   ref_idiom gpu            frames of such a scene through Renderer::render()
//                            against the CPU oracle, and paintTriangle with
//                            world = false (no near clip, Renderer.cpp:169)
//   ref_idiom bench [n]      the reference's bench workload: 20 surfaces of
//                            revolution of 200 x 200 vertices (1,584,040
//                            triangles), 1920x1080, an orbiting camera; n
//                            frames (default 10) of render() alone, of
//                            render() + getImage() and of render() +
//                            getImageRGB8(); prints one JSON line
// Prints "ok" (cpu, gpu) and exits 0, or the first failure and exits 1.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "minirender/Renderer.h"
#include "minirender/Scene.h"
#include "minirender/io.h"
#include "oracle_compare.h"

using namespace minirender;
using namespace mrtest;

static const float kPi = 3.14159265f;

// Uniform floats from a splitmix64 sequence: random(a) in [0, a),
// random(a, b) in [a, b).
struct Random {
  uint64_t s;
  explicit Random(uint64_t seed) : s(seed) {}
  float operator()(float a) { return a * unit(); }
  float operator()(float a, float b) { return a + (b - a) * unit(); }
  float unit() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (float)((z ^ (z >> 31)) >> 40) / 16777216.0f;
  }
};

static double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Radius of the surface of revolution at height z (a bulging vase).
static float profile(float z, float h) {
  const float t = z / h;
  return 22 + 14 * std::sin(2.6f * kPi * t + 0.4f) * (1 - 0.5f * t);
}

// m rings of n vertices around the z axis: 2 (m-1)(n-1) triangles.
Shared<TriMesh> createObject(int m, int n, bool usetex, Random& random) {
  Shared<TriMesh> mesh = new TriMesh();
  const float da = 2 * kPi / n;
  const float h = 70.f, dz = h / m;
  for (int i = 0; i < m; i++) {
    const float z = i * dz;
    const float r = profile(z, h);
    const Vec2 nor = Vec2(1, -(profile(z + 0.001f, h) - r) / 0.001f).normalized();
    for (int j = 0; j < n; j++) {
      mesh->vertices << Vec3(r * std::cos(j * da), r * std::sin(j * da), z);
      mesh->normals << Vec3(nor.x * std::cos(j * da), nor.x * std::sin(j * da), nor.y).normalized();
      if (usetex) mesh->texcoords << Vec2((float)j / n, (float)i / m);
      if (i > 0 && j > 0)
        mesh->indices << n * (i - 1) + j - 1 << n * (i - 1) + j << n * i + j
                      << n * (i - 1) + j - 1 << n * i + j << n * i + j - 1;
    }
  }
  mesh->normalsI = mesh->indices;
  if (usetex) mesh->texcoordsI = mesh->indices;
  mesh->material = new Material();
  mesh->material->shininess = 15;
  if (usetex) {
    Vec3 color = {random(1.f), random(1.f), random(1.f)};
    Array2<Vec3> tex(64, 64);
    for (int i = 0; i < tex.rows(); i++)
      for (int j = 0; j < tex.cols(); j++)
        tex(i, j) = color * (0.7f + 0.3f * std::cos(i * 0.5f) * std::sin(j * 0.3f));
    mesh->material->texture = tex;
  }
  return mesh;
}

// The bench scene: `count` objects scattered and rotated at random.
static Shared<Scene> createScene(int count, int m, int n, bool usetex, Random& random) {
  Shared<Scene> scene = new Scene();
  for (int i = 0; i < count; i++) {
    auto shape = createObject(m, n, usetex, random);
    shape->material->diffuse = {random(1.f), random(1.f), random(1.f)};
    shape->transform = Matrix4::translate(random(-180.f, 180.f), random(-180.f, 180.f),
                                          random(-100.f, 100.f)) *
                       Matrix4::rotate({random(1.f), random(1.f), random(1.f)});
    scene->children << shape;
  }
  scene->ambientLight = 0.2f;
  return scene;
}

static int cpu_checks() {
  // Arrays share storage on copy; dup() detaches, clone() copies.
  Array<int> a;
  a << 1 << 2 << 3;
  Array<int> b = a;
  b << 4;
  CHECK(a.length() == 4 && a.shares(b), "copies share storage");
  Array<int> c = a.clone();
  c << 5;
  CHECK(a.length() == 4 && c.length() == 5 && !c.shares(a), "clone is independent");
  b.dup();
  b << 6;
  CHECK(a.length() == 4 && b.length() == 5, "dup detaches");
  CHECK(!Array<int>() && (bool)a, "empty test");
  Array<double> times;
  times << 0.5 << 0.25;
  double tp = 0;
  for (auto t : times) tp += t;
  CHECK(tp == 0.75 && times.length() == 2, "range for");

  // One object wrapped twice from its raw pointer: one owner.
  Scene* raw = new Scene();
  Shared<Scene> s1 = raw;
  {
    Shared<Scene> s2 = raw;
    CHECK(s1 == s2 && s1.shared().use_count() == 2, "double wrap shares the count");
  }
  CHECK(s1.shared().use_count() == 1, "double wrap released");

  // The scene of the idiom.
  Random random(7);
  Shared<Scene> scene = createScene(3, 20, 30, true, random);
  CHECK(scene->children.length() == 3, "children");
  auto* m0 = dynamic_cast<TriMesh*>(scene->children[0].get());
  CHECK(m0 && m0->indices.length() == 3 * 2 * 19 * 29, "indices %d", m0 ? m0->indices.length() : -1);
  CHECK(m0->normalsI.shares(m0->indices) && m0->texcoordsI.shares(m0->indices),
        "normalsI = indices shares the array");
  CHECK(m0->vertices.length() == 600 && m0->texcoords.length() == 600, "vertices");
  CHECK(m0->material && m0->material->shininess == 15 && m0->material->texture.rows() == 64,
        "material");
  Array<Renderable> list;
  scene->collectShapes(list, Matrix4::identity());
  CHECK(list.length() == 3 && list[1].mesh == scene->children[1].get(), "collectShapes");
  auto box = scene->getBbox();
  auto size = box.size();
  auto center = box.center();
  CHECK(size.x > 0 && size.y > 0 && size.z > 0 && std::fabs(center.x) < 400, "bbox");

  // Console output of the render sample (samples/main.cpp:13-28).
  Array2<Vec3> img(1, 2);
  img(0, 0) = Vec3(1, 0.5f, 0);
  img(0, 1) = Vec3(2, -1, 0.1f);
  CHECK(consoleImage(img) == "\x1b[H\x1b[48;2;255;127;0m \x1b[48;2;255;0;25m \x1b[0m\n",
        "console image");
  CHECK(consoleImage(img, true) == "\x1b[H\x1b[48;5;208m \x1b[48;5;196m \x1b[0m\n",
        "console image, 256 colours");
  return 0;
}

// Oracle comparison of the renderer's current frame (draws of the scene).
static int check_frame(Renderer& r, const Shared<Scene>& scene, const char* what) {
  const mr_frame f = r.describeFrame();
  std::vector<mr_draw> draws(f.draws, f.draws + f.ndraws);
  std::vector<const Array2<Vec3>*> texs;
  Array<Renderable> list;
  scene->collectShapes(list, Matrix4::identity());
  for (auto& it : list) texs.push_back(&it.mesh->material->texture);
  std::vector<float> depth;
  std::vector<uint32_t> win;
  std::vector<uint8_t> rgb8;
  CHECK(oracle_render(r, f, draws, texs, depth, win, rgb8) == 0, "%s: oracle", what);
  return compare(r, depth, win, rgb8, what);
}

static int gpu_checks() {
  Random random(11);
  Scene* scene = new Scene();  // as samples/main.cpp:108-112
  Shared<Scene> objects = createScene(4, 24, 32, true, random);
  for (auto& c : objects->children) scene->children << c;
  scene->ambientLight = 0.2f;
  Renderer renderer;
  renderer.setLight(Vec3(-0.4f, .6f, 1.f));
  renderer.setScene(scene);
  renderer.setSize(320, 180);
  renderer.setProjection(projectionFrustum(35 * kPi / 180, renderer.aspect(), 10, 7000));
  renderer.setLighting(true);
  renderer.setTexturing(true);
  renderer.setSaveNormals(false);
  float rx = -kPi / 2 + 0.35f, rz = 0;
  for (int i = 0; i < 3; i++) {
    rz += 0.7f;
    renderer.setView(Matrix4::translate(0, 0, -500) * Matrix4::rotateX(rx) * Matrix4::rotateZ(rz));
    renderer.render();
    char what[32];
    std::snprintf(what, sizeof(what), "frame %d", i);
    if (check_frame(renderer, Shared<Scene>(scene), what)) return 1;
  }

  // paintTriangle(a, b, c, false): eye-space vertices, one between the eye
  // and the near plane, drawn without the near test (Renderer.cpp:169),
  // against the oracle drawing the same triangle unclipped.
  Renderer r2;
  r2.setSize(96, 64);
  r2.setProjection(projectionFrustum(1.0f, r2.aspect(), 0.5f, 50.0f));
  r2.setLighting(false);
  Shared<Scene> empty = new Scene();
  r2.setScene(empty);
  r2.render();
  Shared<Material> mat = new Material();
  mat->emissive = Vec3(0.9f, 0.5f, 0.1f);
  r2.setMaterial(mat);
  const Vertex va(Vec3(-1.0f, -0.8f, -2.0f)), vb(Vec3(1.5f, -0.6f, -3.0f)),
      vc(Vec3(0.2f, 1.2f, -0.3f));
  r2.paintTriangle(va, vb, vc, false);
  std::vector<uint32_t> gw = r2.getWinners();
  Array2<float> gd = r2.getDepth();
  std::vector<uint8_t> g8 = r2.getImageRGB8();
  CHECK(r2.lastStatus() == 0, "paintTriangle status %d", r2.lastStatus());
  // oracle: the same triangle as a one-triangle mesh drawn unclipped
  const float pos[9] = {va.position.x, va.position.y, va.position.z, vb.position.x,
                        vb.position.y, vb.position.z, vc.position.x, vc.position.y,
                        vc.position.z};
  const float nrm[9] = {0, 0, 1, 0, 0, 1, 0, 0, 1};
  const float uv[6] = {0, 0, 0, 0, 0, 0};
  const int32_t idx[3] = {0, 1, 2};
  mro_mesh mm{pos, 3, nrm, 3, uv, 3, idx, idx, idx, 1};
  mr_draw d{};
  const Matrix4 I = Matrix4::identity();
  std::memcpy(d.modelview, I.data(), 64);
  std::memcpy(d.normalmat, I.data(), 64);
  d.diffuse[0] = 0.7f; d.diffuse[1] = 0.7f; d.diffuse[2] = 0.9f;
  d.specular[0] = d.specular[1] = d.specular[2] = 0.8f;
  d.emissive[0] = 0.9f; d.emissive[1] = 0.5f; d.emissive[2] = 0.1f;
  d.shininess = 12;
  d.texture = -1;
  d.mesh = 0;
  d.flags = MR_DRAW_NO_NEAR_CLIP;
  mr_frame f = r2.describeFrame();
  f.draws = &d;
  f.ndraws = 1;
  const size_t n = 96 * 64;
  std::vector<float> depth(n);
  std::vector<uint32_t> win(n);
  std::vector<uint8_t> rgb8(n * 3);
  mr_outputs o{};
  o.depth = depth.data();
  o.winner = win.data();
  o.rgb8 = rgb8.data();
  CHECK(mro_render(&f, &mm, 1, nullptr, 0, &o) == 0, "oracle unclipped triangle");
  size_t covered = 0;
  for (size_t i = 0; i < n; i++) {
    CHECK(gw[i] == win[i], "unclipped: winner at %zu: gpu %u oracle %u", i, gw[i], win[i]);
    CHECK(std::memcmp(&gd.data()[i], &depth[i], 4) == 0, "unclipped: depth at %zu", i);
    for (int k = 0; k < 3; k++)
      CHECK(std::abs((int)g8[i * 3 + k] - (int)rgb8[i * 3 + k]) <= 1, "unclipped: rgb8 at %zu", i);
    covered += win[i] != 0xFFFFFFFFu;
  }
  // the clipped drawing of the same triangle covers other pixels
  d.flags = 0;
  CHECK(mro_render(&f, &mm, 1, nullptr, 0, &o) == 0, "oracle clipped triangle");
  size_t differ = 0;
  for (size_t i = 0; i < n; i++) differ += gw[i] != win[i];
  CHECK(covered > 50 && differ > 0, "unclipped covered %zu, differs from clipped at %zu",
        covered, differ);
  return 0;
}

static int bench(int frames) {
  Random random(5);
  Shared<Scene> scene = createScene(20, 200, 200, false, random);
  int64_t tris = 0;
  for (auto& c : scene->children) tris += dynamic_cast<TriMesh*>(c.get())->indices.length() / 3;
  const int sizew = 1920, sizeh = 1080;
  const float d = 700, fov = 35 * kPi / 180;
  Renderer renderer;
  renderer.setLight(Vec3(-0.4f, .6f, 1.f));
  renderer.setScene(scene);
  renderer.setSize(sizew, sizeh);
  renderer.setProjection(projectionFrustum(fov, renderer.aspect(), 10, 7000));
  renderer.setLighting(true);
  renderer.setTexturing(false);
  renderer.setSaveNormals(false);
  const float wz = 40 * kPi / 180, tilt = 20 * kPi / 180;
  double res[3] = {0, 0, 0};
  for (int mode = 0; mode < 3; mode++) {
    float rx = -kPi / 2 + tilt, rz = 0;
    auto frame = [&]() {
      rz += wz * 0.1f;
      renderer.setView(Matrix4::translate(0, 0, -d) * Matrix4::rotateX(rx) * Matrix4::rotateZ(rz));
      renderer.render();
      if (mode == 1) {
        Array2<Vec3> img = renderer.getImage();
        if (img.rows() != sizeh) std::abort();
      } else if (mode == 2) {
        std::vector<uint8_t> img = renderer.getImageRGB8();
        if (img.size() != (size_t)sizew * sizeh * 3) std::abort();
      }
    };
    frame();  // warm-up (uploads the meshes on the first call)
    const double t0 = now();
    for (int i = 0; i < frames; i++) frame();
    res[mode] = (now() - t0) / frames;
    if (renderer.lastStatus() != 0) {
      std::printf("render status %d\n", renderer.lastStatus());
      return 1;
    }
  }
  std::printf("{\"triangles\": %lld, \"width\": %d, \"height\": %d, \"frames\": %d, "
              "\"ms_render\": %.4f, \"ms_render_getImage\": %.4f, "
              "\"ms_render_getImageRGB8\": %.4f}\n",
              (long long)tris, sizew, sizeh, frames, res[0] * 1e3, res[1] * 1e3, res[2] * 1e3);
  return 0;
}

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "";
  if (mode == "cpu") {
    if (cpu_checks()) return 1;
    std::printf("ok\n");
    return 0;
  }
  if (mode == "gpu") {
    if (gpu_checks()) return 1;
    std::printf("ok\n");
    return 0;
  }
  if (mode == "bench") return bench(argc > 2 ? std::atoi(argv[2]) : 10);
  std::printf("usage: ref_idiom cpu|gpu|bench [frames]\n");
  return 2;
}
