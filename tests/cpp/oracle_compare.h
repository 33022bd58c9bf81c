// oracle_compare.h -- shared by the C++ API checks: render a Renderer's
// frame with the CPU oracle (test infrastructure) and compare it with the
// GPU's (winner and depth bit-equal, RGB8 within 1 LSB).
#ifndef MR_TESTS_ORACLE_COMPARE_H
#define MR_TESTS_ORACLE_COMPARE_H

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "minirender/Renderer.h"
#include "oracle.h"

namespace mrtest {

using namespace minirender;

#define CHECK(cond, ...)                      \
  do {                                        \
    if (!(cond)) {                            \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
      std::printf(__VA_ARGS__);               \
      std::printf("\n");                      \
      return 1;                               \
    }                                         \
  } while (0)

struct OracleInput {
  std::vector<mro_mesh> meshes;
  std::vector<mro_texture> textures;
};

inline int oracle_render(Renderer& r, const mr_frame& f0, const std::vector<mr_draw>& draws,
                         const std::vector<const Array2<Vec3>*>& texs, std::vector<float>& depth,
                         std::vector<uint32_t>& win, std::vector<uint8_t>& rgb8) {
  OracleInput in;
  std::vector<mr_draw> d2 = draws;
  for (size_t i = 0; i < draws.size(); i++) {
    const TriMesh* m = r.meshOfHandle(draws[i].mesh);
    if (!m) return MR_ERR_HANDLE;
    mro_mesh mm{};
    mm.positions = (const float*)m->vertices.data();
    mm.npositions = (int32_t)m->vertices.size();
    mm.normals = (const float*)m->normals.data();
    mm.nnormals = (int32_t)m->normals.size();
    mm.indices = m->indices.data();
    mm.normal_indices = m->normalsI.size() == m->indices.size() ? m->normalsI.data() : m->indices.data();
    if (!m->texcoords.empty() && m->texcoordsI.size() == m->indices.size()) {
      mm.uvs = (const float*)m->texcoords.data();
      mm.nuvs = (int32_t)m->texcoords.size();
      mm.uv_indices = m->texcoordsI.data();
    }
    mm.ntriangles = (int32_t)(m->indices.size() / 3);
    d2[i].mesh = (int32_t)in.meshes.size();
    in.meshes.push_back(mm);
    d2[i].texture = -1;
    if (texs[i] && texs[i]->rows() > 0) {
      mro_texture t{texs[i]->data(), MR_TEX_RGB32F, texs[i]->rows(), texs[i]->cols()};
      d2[i].texture = (int32_t)in.textures.size();
      in.textures.push_back(t);
    }
  }
  mr_frame f = f0;
  f.draws = d2.data();
  f.ndraws = (int32_t)d2.size();
  const size_t n = (size_t)f.width * f.height;
  depth.assign(n, 0);
  win.assign(n, 0);
  rgb8.assign(n * 3, 0);
  mr_outputs o{};
  o.depth = depth.data();
  o.winner = win.data();
  o.rgb8 = rgb8.data();
  return mro_render(&f, in.meshes.data(), (int32_t)in.meshes.size(), in.textures.data(),
                    (int32_t)in.textures.size(), &o);
}

inline int compare(Renderer& r, const std::vector<float>& depth, const std::vector<uint32_t>& win,
                   const std::vector<uint8_t>& rgb8, const char* what) {
  Array2<float> gd = r.getDepth();
  std::vector<uint32_t> gw = r.getWinners();
  std::vector<uint8_t> g8 = r.getImageRGB8();
  CHECK(r.lastStatus() == 0, "%s: render status %d", what, r.lastStatus());
  size_t covered = 0;
  for (size_t i = 0; i < win.size(); i++) {
    CHECK(gw[i] == win[i], "%s: winner at %zu: gpu %u oracle %u", what, i, gw[i], win[i]);
    CHECK(std::memcmp(&gd.data()[i], &depth[i], 4) == 0, "%s: depth at %zu", what, i);
    for (int k = 0; k < 3; k++)
      CHECK(std::abs((int)g8[i * 3 + k] - (int)rgb8[i * 3 + k]) <= 1, "%s: rgb8 at %zu", what, i);
    covered += win[i] != 0xFFFFFFFFu;
  }
  CHECK(covered > 100, "%s: only %zu pixels covered", what, covered);
  return 0;
}

}  // namespace mrtest

#endif
