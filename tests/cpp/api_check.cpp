// Checks of the minirender C++ API (include/minirender/*.h), written the way
// a user of the reference library drives it (samples/main.cpp:128-182).
//
//   api_check cpu   host-only: primitive counts and order, scene flattening
//                   order, matrix inverse, projection helpers.
//   api_check gpu   Renderer::render() on the GPU against the CPU oracle for
//                   scenes built with the API (nested nodes, transforms,
//                   materials, a float texture, point light, immediate-mode
//                   paintMesh after render, a two-material OBJ whose meshes
//                   share one attribute upload): winner and depth bit-equal,
//                   RGB8 within 1 LSB.
// Prints "ok" and exits 0, or prints the first failure and exits 1.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "minirender/Renderer.h"
#include "minirender/io.h"
#include "minirender/primitives.h"
#include "oracle.h"
#include "oracle_compare.h"

using namespace minirender;
using namespace mrtest;

// STL / OBJ loaders on small files written here, against the reference's
// parsing rules (io.cpp:15-335, x3d.cpp:17-33).
static int loader_checks() {
  // fan triangulation with -1 separators
  const std::vector<int> fan = triangulateIndices({0, 1, 2, 3, -1, 4, 5, 6, -1});
  const std::vector<int> fan_ref = {0, 1, 2, 0, 2, 3, 4, 5, 6};
  CHECK(fan == fan_ref, "triangulateIndices");
  // OBJ: a quad and a triangle with "v//vn" in material "b", a triangle
  // without material (""), a material "a" never used; vt v flipped.
  {
    FILE* m = std::fopen("/tmp/mr_api_check.mtl", "w");
    std::fprintf(m, "newmtl a\nKd 0.1 0.2 0.3\nnewmtl b\nKd 0.5 0.6 0.7\nNs 0\nKs 1 1 1\n");
    std::fclose(m);
    FILE* o = std::fopen("/tmp/mr_api_check.obj", "w");
    std::fprintf(o,
                 "# test\nmtllib mr_api_check.mtl\n"
                 "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 2 2 2\n"
                 "vt 0.25 0.25\nvn 0 0 1\n"
                 "f 1/1/1 2/1/1 3/1/1\n"
                 "usemtl b\nf 1/1/1 2/1/1 3/1/1 4/1/1\nf 1//1 3//1 5//1\n");
    std::fclose(o);
    Shared<SceneNode> node = loadOBJ("/tmp/mr_api_check.obj");
    CHECK(node && node->children.size() == 2, "OBJ meshes");
    auto* m0 = dynamic_cast<TriMesh*>(node->children[0].get());
    auto* m1 = dynamic_cast<TriMesh*>(node->children[1].get());
    CHECK(m0 && m1, "OBJ mesh types");
    CHECK(m0->indices == std::vector<int>({0, 1, 2}), "OBJ default mesh indices");
    CHECK(m1->indices == std::vector<int>({0, 1, 2, 0, 2, 3, 0, 2, 4}), "OBJ quad + tri");
    // "v//vn" leaves texcoord index -1, which closes the polygon: only the
    // quad keeps texcoord indices
    CHECK(m1->texcoordsI == std::vector<int>({0, 0, 0, 0, 0, 0}), "OBJ texcoord indices");
    CHECK(m1->normalsI.size() == 9 && m1->normals.size() == 1, "OBJ normals");
    CHECK(m1->texcoords.size() == 1 && m1->texcoords[0].y == 0.75f, "OBJ vt flip");
    CHECK(m1->material && m1->material->diffuse.x == 0.5f && m1->material->shininess == 10.0f,
          "OBJ material b");
    CHECK(m0->material && m0->material->diffuse.x == 0.7f, "OBJ default material");
    CHECK(m0->vertices.size() == 5, "OBJ shared vertices");
  }
  // OBJ without normals: one normal per face
  {
    FILE* o = std::fopen("/tmp/mr_api_check2.obj", "w");
    std::fprintf(o, "v 0 0 0\nv 2 0 0\nv 0 2 0\nf 1 2 3\n");
    std::fclose(o);
    Shared<SceneNode> node = loadOBJ("/tmp/mr_api_check2.obj");
    auto* m = dynamic_cast<TriMesh*>(node->children[0].get());
    CHECK(m && m->normals.size() == 1 && m->normals[0].z == 1.0f &&
              m->normalsI == std::vector<int>({0, 0, 0}),
          "OBJ face normals");
  }
  // STL: binary round trip of a cube, then an ASCII file
  {
    Shared<TriMesh> cube = createCube(1.0f);
    CHECK(saveSTL(cube, "/tmp/mr_api_check.stl"), "saveSTL");
    Shared<TriMesh> back = loadSTL("/tmp/mr_api_check.stl");
    CHECK(back && back->indices.size() == cube->indices.size() &&
              back->vertices.size() == cube->indices.size() && back->normals.size() == 12,
          "STL binary round trip sizes");
    for (size_t i = 0; i < cube->indices.size(); i++) {
      const Vec3 a = cube->vertices[cube->indices[i]], b = back->vertices[back->indices[i]];
      CHECK(a.x == b.x && a.y == b.y && a.z == b.z, "STL vertex %zu", i);
    }
    FILE* s = std::fopen("/tmp/mr_api_check_a.stl", "w");
    std::fprintf(s,
                 "solid t\nfacet normal 0 0 1\nouter loop\nvertex 0 0 0\nvertex 1 0 0\n"
                 "vertex 0 1 0\nendloop\nendfacet\nendsolid t\n");
    std::fclose(s);
    Shared<TriMesh> a = loadSTL("/tmp/mr_api_check_a.stl");
    CHECK(a && a->indices == std::vector<int>({0, 1, 2}) && a->normals.size() == 1 &&
              a->normalsI == std::vector<int>({0, 0, 0}) && a->vertices[1].x == 1.0f,
          "STL ascii");
    Shared<SceneNode> node = loadMesh("/tmp/mr_api_check_a.stl");
    CHECK(node && node->children.size() == 1, "loadMesh stl");
  }
  // X3D: a Transform holding a Shape (IndexedFaceSet: a quad and a
  // triangle, commas, no normals or texcoords, DEF appearance), a Group
  // with a Shape that USEs the appearance (IndexedTriangleSet with normals
  // and texcoords), and an Inline scene; comments, prolog and DOCTYPE.
  {
    FILE* in = std::fopen("/tmp/mr_api_check_inl.x3d", "w");
    std::fprintf(in,
                 "<X3D><Scene><Shape><IndexedTriangleSet index='0 1 2'>"
                 "<Coordinate point='0 0 0 1 0 0 0 1 0'/></IndexedTriangleSet></Shape>"
                 "</Scene></X3D>\n");
    std::fclose(in);
    FILE* x = std::fopen("/tmp/mr_api_check.x3d", "w");
    std::fprintf(x,
                 "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n"
                 "<!DOCTYPE X3D PUBLIC \"ISO//Web3D//DTD X3D 3.3//EN\" \"x3d-3.3.dtd\">\n"
                 "<X3D profile='Interchange'><head><meta name='t' content='a &amp; b'/></head>\n"
                 "<Scene><!-- a comment <Shape/> -->\n"
                 "<Transform translation='1 2 3' rotation='0 0 2 1.5707963' scale='2 2 2'>\n"
                 " <Shape><Appearance DEF='look'><Material diffuseColor='0.1 0.2 0.3' "
                 "shininess='0.25'/></Appearance>\n"
                 "  <IndexedFaceSet coordIndex='0 1 2 3 -1, 0 2 4 -1'>"
                 "<Coordinate point='0 0 0, 1 0 0, 1 1 0, 0 1 0, 2 2 2'/></IndexedFaceSet>\n"
                 " </Shape>\n"
                 "</Transform>\n"
                 "<Group><Shape><Appearance USE='look'/>\n"
                 "  <IndexedTriangleSet index='0 1 2'><Coordinate point='0 0 0 1 0 0 0 1 0'/>"
                 "<Normal vector='0 0 1 0 0 1 0 0 1'/>"
                 "<TextureCoordinate point='0 0 1 0 0.25 0.75'/></IndexedTriangleSet>\n"
                 "</Shape></Group>\n"
                 "<Inline url='\"mr_api_check_inl.x3d\"'/>\n"
                 "</Scene></X3D>\n");
    std::fclose(x);
    Shared<SceneNode> root = loadMesh("/tmp/mr_api_check.x3d");
    CHECK(root && root->children.size() == 3, "X3D root children");
    SceneNode* tr = root->children[0].get();
    CHECK(tr && tr->children.size() == 1, "X3D transform children");
    // translate(1,2,3) * rotate(z, pi/2) * scale(2): (1,0,0) -> (1, 4, 3)
    const Vec3 p = tr->transform * Vec3(1, 0, 0);
    CHECK(std::fabs(p.x - 1) < 1e-5f && std::fabs(p.y - 4) < 1e-5f && std::fabs(p.z - 3) < 1e-5f,
          "X3D transform %f %f %f", p.x, p.y, p.z);
    auto* m0 = dynamic_cast<TriMesh*>(tr->children[0].get());
    CHECK(m0 && m0->indices == std::vector<int>({0, 1, 2, 0, 2, 3, 0, 2, 4}), "X3D faces");
    CHECK(m0->normalsI == std::vector<int>({0, 0, 0, 1, 1, 1, 2, 2, 2}) &&
              m0->normals.size() == 3 && m0->normals[0].z == 1.0f,
          "X3D face normals");
    CHECK(m0->texcoords.size() == 1 && m0->texcoordsI == std::vector<int>(9, 0),
          "X3D default texcoords");
    CHECK(m0->material && m0->material->diffuse.y == 0.2f && m0->material->shininess == 2.0f &&
              m0->material->specular.x == 0.4f,
          "X3D material");
    auto* m1 = dynamic_cast<TriMesh*>(root->children[1]->children[0].get());
    CHECK(m1 && m1->material && m1->material->diffuse.z == 0.3f, "X3D USE appearance");
    CHECK(m1->indices == std::vector<int>({0, 1, 2}) && m1->normalsI == m1->indices &&
              m1->normals.size() == 3 && m1->texcoords.size() == 3 &&
              m1->texcoords[2].y == 0.25f,
          "X3D triangle set");
    SceneNode* inl = root->children[2].get();
    CHECK(inl && inl->children.size() == 1 && dynamic_cast<TriMesh*>(inl->children[0].get()),
          "X3D inline");
    CHECK(!loadX3D("/tmp/mr_api_check_a.stl"), "X3D rejects a non-X3D file");
  }
  return 0;
}

static int cpu_checks() {
  // createSphere(r, lat, long): (lat-1)*long + 2 vertices, 2*long*(lat-1) triangles
  auto s = createSphere(1.0f, 32, 64);
  CHECK(s->vertices.size() == 31 * 64 + 2, "sphere vertices %zu", s->vertices.size());
  CHECK(s->indices.size() == 3 * 2 * 64 * 31, "sphere indices %zu", s->indices.size());
  CHECK(s->indices[0] == 0 && s->indices[1] == 1 && s->indices[2] == 2, "sphere top fan order");
  CHECK(s->normalsI == s->indices && s->texcoordsI.empty(), "sphere index arrays");
  auto c = createCube(2.0f);
  CHECK(c->vertices.size() == 24 && c->indices.size() == 36, "cube sizes");
  CHECK(c->indices[0] == 0 && c->indices[1] == 2 && c->indices[2] == 1, "cube winding");
  auto y = createCylinder(0.5f, 1.0f, 16, 2, true);
  CHECK(y->vertices.size() == 3 * 16 + 2 * 17, "cylinder vertices %zu", y->vertices.size());
  CHECK(y->indices.size() == 3 * (2 * 16 * 2 + 2 * 16), "cylinder indices");
  // every cube face is counter-clockwise seen from outside
  for (size_t t = 0; t < c->indices.size(); t += 3) {
    const Vec3 a = c->vertices[c->indices[t]], b = c->vertices[c->indices[t + 1]],
               d = c->vertices[c->indices[t + 2]];
    const Vec3 n = (b - a) ^ (d - a);
    CHECK(n * c->normals[c->normalsI[t]] > 0, "cube face %zu winding", t / 3);
  }
  // scene flattening: pre-order, a TriMesh before its children (Scene.cpp:13-36)
  auto scene = std::make_shared<Scene>();
  auto node = std::make_shared<SceneNode>();
  auto m1 = createCube(), m2 = createCube(), m3 = createCube();
  m1->children.push_back(m2);
  node->children.push_back(m1);
  scene->add(node);
  scene->add(m3);
  node->transform = Matrix4::translate(1, 2, 3);
  m1->transform = Matrix4::scale(2.0f);
  Array<Renderable> list;
  scene->collectShapes(list, Matrix4::identity());
  CHECK(list.size() == 3 && list[0].mesh == m1.get() && list[1].mesh == m2.get() &&
            list[2].mesh == m3.get(), "flatten order");
  CHECK(list[1].transform(0, 3) == 1 && list[1].transform(0, 0) == 2, "flatten transforms");
  // inverse
  const Matrix4 M = Matrix4::translate(1, -2, 3) * Matrix4::rotate(Vec3(0.3f, 0.5f, -0.2f)) *
                    Matrix4::scale(1.7f);
  const Matrix4 I = M * M.inverse();
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      CHECK(std::fabs(I(i, j) - (i == j ? 1.0f : 0.0f)) < 1e-5f, "inverse %d %d", i, j);
  // projection helpers agree with the C-ABI
  const Matrix4 P = projectionFrustum(0.8f, 1.5f, 0.1f, 100.0f);
  float ref[16];
  mr_projection_frustum(0.8f, 1.5f, 0.1f, 100.0f, ref);
  CHECK(std::memcmp(P.data(), ref, sizeof(ref)) == 0, "projectionFrustum");
  // PPM round trip with the reference quantization
  Array2<Vec3> img(3, 4);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 4; j++) img(i, j) = Vec3(j / 3.0f, i / 2.0f, 0.5f);
  const std::string path = "/tmp/mr_api_check.ppm";
  CHECK(savePPM(img, path), "savePPM");
  Array2<Vec3> back = loadPPM(path);
  CHECK(back.rows() == 3 && back.cols() == 4, "loadPPM size");
  CHECK(back(0, 0).z == 127.0f / 255.0f && back(2, 3).x == 1.0f, "PPM values");
  if (int rc = loader_checks()) return rc;
  std::printf("ok\n");
  return 0;
}

// Oracle rendering of draws (from describeFrame) whose mesh handles map to
// the renderer's TriMeshes; textures given per draw as float arrays.
static int gpu_checks() {
  // Scene as in samples/main.cpp: a loaded-style mesh under a node, camera
  // orbit via setView, frustum projection.
  auto scene = std::make_shared<Scene>();
  scene->ambientLight = 0.2f;
  auto sphere = createSphere(1.0f, 24, 48);
  sphere->material->diffuse = Vec3(0.7f, 0.5f, 0.3f);
  sphere->material->shininess = 32;
  auto cube = createCube(0.8f);
  cube->transform = Matrix4::translate(1.2f, 0.3f, -0.5f) * Matrix4::rotate(Vec3(0.2f, 0.7f, 0.1f));
  cube->material->diffuse = Vec3(0.2f, 0.6f, 0.9f);
  // a textured quad (float texture, uv indices given)
  auto quad = std::make_shared<TriMesh>();
  quad->vertices = {Vec3(-2, -1.2f, -1), Vec3(2, -1.2f, -1), Vec3(2, -1.2f, 1.5f), Vec3(-2, -1.2f, 1.5f)};
  quad->normals = {Vec3(0, 1, 0)};
  quad->texcoords = {Vec2(0, 0), Vec2(3, 0), Vec2(3, 2), Vec2(0, 2)};
  quad->indices = {0, 2, 1, 0, 3, 2};
  quad->normalsI = {0, 0, 0, 0, 0, 0};
  quad->texcoordsI = {0, 2, 1, 0, 3, 2};
  quad->material = std::make_shared<Material>();
  quad->material->texture.resize(16, 16);
  for (int i = 0; i < 16; i++)
    for (int j = 0; j < 16; j++)
      quad->material->texture(i, j) = ((i / 4 + j / 4) & 1) ? Vec3(0.9f, 0.9f, 0.8f) : Vec3(0.2f, 0.3f, 0.5f);
  auto node = std::make_shared<SceneNode>();
  node->transform = Matrix4::rotateY(0.3f);
  node->children.push_back(sphere);
  sphere->children.push_back(cube);  // cube inherits the sphere's transform
  scene->add(node);
  scene->add(quad);

  Renderer r;
  r.setBackground(Vec3(0.1f, 0.1f, 0.15f));
  r.setLight(Vec3(2, 3, 4), true);
  r.setScene(scene);
  r.setSize(200, 150);
  r.setProjection(projectionFrustum(0.8f, r.aspect(), 0.1f, 100.0f));
  for (int frame = 0; frame < 3; frame++) {
    const float rz = 0.4f * frame;
    r.setView(Matrix4::translate(0, 0, -5) * Matrix4::rotateX(0.35f) * Matrix4::rotateZ(rz));
    r.render();
    const mr_frame f = r.describeFrame();
    std::vector<mr_draw> draws(f.draws, f.draws + f.ndraws);
    std::vector<const Array2<Vec3>*> texs;
    Array<Renderable> list;
    scene->collectShapes(list, Matrix4::identity());
    for (auto& it : list) texs.push_back(&it.mesh->material->texture);
    std::vector<float> depth;
    std::vector<uint32_t> win;
    std::vector<uint8_t> rgb8;
    CHECK(oracle_render(r, f, draws, texs, depth, win, rgb8) == 0, "oracle");
    if (compare(r, depth, win, rgb8, "frame")) return 1;
    // immediate mode: paintMesh after render() composites like sequential drawing
    if (frame == 2) {
      auto extra = createSphere(0.6f, 12, 24);
      extra->material->diffuse = Vec3(1, 0.2f, 0.2f);
      const Matrix4 xf = Matrix4::translate(0.4f, 0.2f, 0.9f);
      r.paintMesh(extra.get(), xf);
      r.paintMesh(extra.get(), xf);  // drawn twice: the copy never wins (strict test)
      // oracle: the frame's draws followed by the two extra draws, one sequence
      const mr_frame f2 = r.describeFrame();  // re-describes the scene part
      std::vector<mr_draw> all(f2.draws, f2.draws + f2.ndraws);
      std::vector<const Array2<Vec3>*> texs2 = texs;
      // the extra draw as paintMesh builds it (Renderer.cpp:336-338)
      const Matrix4 view = Matrix4::translate(0, 0, -5) * Matrix4::rotateX(0.35f) * Matrix4::rotateZ(rz);
      const Matrix4 m4 = view * xf;
      const Matrix4 nm = m4.inverse().t();
      mr_draw d{};
      std::memcpy(d.modelview, m4.data(), 64);
      std::memcpy(d.normalmat, nm.data(), 64);
      d.diffuse[0] = 1; d.diffuse[1] = 0.2f; d.diffuse[2] = 0.2f;
      d.specular[0] = d.specular[1] = d.specular[2] = 0.8f;
      d.shininess = 12;
      d.texture = -1;
      d.mesh = -1;
      for (int h = 0; h < 4096; h++)
        if (r.meshOfHandle(h) == extra.get()) { d.mesh = h; break; }
      CHECK(d.mesh >= 0, "immediate mesh handle");
      all.push_back(d);
      all.push_back(d);
      texs2.push_back(nullptr);
      texs2.push_back(nullptr);
      CHECK(oracle_render(r, f2, all, texs2, depth, win, rgb8) == 0, "oracle immediate");
      if (compare(r, depth, win, rgb8, "immediate")) return 1;
    }
  }
  // A two-material OBJ: loadOBJ's meshes share one vertex and normal array
  // (io.cpp:313-315); the renderer uploads it once and the image matches.
  {
    FILE* mt = std::fopen("/tmp/mr_api_check_grid.mtl", "w");
    std::fprintf(mt, "newmtl a\nKd 0.1 0.6 0.3\nnewmtl b\nKd 0.7 0.4 0.2\nNs 20\n");
    std::fclose(mt);
    FILE* o = std::fopen("/tmp/mr_api_check_grid.obj", "w");
    std::fprintf(o, "mtllib mr_api_check_grid.mtl\nvn 0 0 1\n");
    const int n = 9;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++)
        std::fprintf(o, "v %g %g %g\n", -1.0 + 2.0 * j / (n - 1), -1.0 + 2.0 * i / (n - 1),
                     0.15 * ((i * 7 + j * 3) % 5));
    for (int i = 0; i + 1 < n; i++) {
      std::fprintf(o, "usemtl %s\n", i < n / 2 ? "a" : "b");
      for (int j = 0; j + 1 < n; j++) {
        const int v = i * n + j + 1;
        std::fprintf(o, "f %d//1 %d//1 %d//1 %d//1\n", v, v + 1, v + n + 1, v + n);
      }
    }
    std::fclose(o);
    Shared<SceneNode> grid = loadOBJ("/tmp/mr_api_check_grid.obj");
    // meshes "" (no faces), "a", "b", each with all 81 vertices
    CHECK(grid && grid->children.size() == 3, "grid OBJ meshes");
    auto* g0 = dynamic_cast<TriMesh*>(grid->children[1].get());
    auto* g1 = dynamic_cast<TriMesh*>(grid->children[2].get());
    CHECK(g0 && g1 && g0->vertices.size() == (size_t)(n * n) &&
              g1->vertices.size() == g0->vertices.size(), "grid OBJ vertices");
    auto gscene = std::make_shared<Scene>();
    gscene->add(grid);
    auto one = std::make_shared<Scene>();
    one->add(grid->children[0]);
    one->add(grid->children[1]);
    Renderer rg, r0;
    for (Renderer* q : {&rg, &r0}) {
      q->setLight(Vec3(1, 2, 3), true);
      q->setSize(160, 120);
      q->setProjection(projectionFrustum(0.9f, q->aspect(), 0.1f, 50.0f));
      q->setView(Matrix4::translate(0, 0, -3.5f) * Matrix4::rotateX(-0.4f));
    }
    rg.setScene(gscene);
    r0.setScene(one);
    rg.render();
    r0.render();
    CHECK(rg.lastStatus() == 0 && r0.lastStatus() == 0, "grid render");
    uint64_t res_g = 0, res_0 = 0, scr = 0;
    CHECK(mr_ctx_memory(rg.context(), &res_g, &scr) == 0 &&
              mr_ctx_memory(r0.context(), &res_0, &scr) == 0, "grid memory");
    // Mesh "b" adds only its index arrays (positions and normal indices;
    // device allocations are rounded up to 256 bytes).
    const uint64_t ibytes = (g1->indices.size() * sizeof(int) + 255) & ~(uint64_t)255;
    CHECK(res_g - res_0 == 2 * ibytes, "grid arrays shared: %llu %llu",
          (unsigned long long)res_g, (unsigned long long)res_0);
    const mr_frame f = rg.describeFrame();
    std::vector<mr_draw> draws(f.draws, f.draws + f.ndraws);
    std::vector<const Array2<Vec3>*> texs(draws.size(), nullptr);
    std::vector<float> depth;
    std::vector<uint32_t> win;
    std::vector<uint8_t> rgb8;
    CHECK(oracle_render(rg, f, draws, texs, depth, win, rgb8) == 0, "grid oracle");
    if (compare(rg, depth, win, rgb8, "grid")) return 1;
    size_t covered = 0;
    for (uint32_t w : win) covered += w != 0xFFFFFFFFu;
    CHECK(covered > 4000, "grid coverage %zu", covered);
  }
  std::printf("ok\n");
  return 0;
}

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "cpu";
  if (mode == "cpu") return cpu_checks();
  if (mode == "gpu") return gpu_checks();
  std::printf("usage: api_check cpu|gpu\n");
  return 2;
}
