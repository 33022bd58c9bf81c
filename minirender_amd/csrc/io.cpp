// io.cpp -- mesh files of the minirender API: STL (ASCII and binary) and
// OBJ (+ MTL) loaders, STL and XYZ writers (reference: src/io.cpp:15-335,
// 417-431).  Loaded meshes feed the renderer's resident mesh pool
// (mr_mesh_upload) on their first render.
//
// Behaviour kept from the reference because it decides what is drawn and in
// which order: binary STL is recognised by its exact size (84 + 50 n); OBJ
// faces are fan-triangulated per polygon; a texcoord field left empty
// ("v//vn") yields index -1, which ends the polygon in the triangulation, so
// such faces carry no texcoord indices; one mesh per material, visited in
// material-name order ("" first); meshes without normals get one normal per
// face, ((b - a) x (c - a)) normalised; `vt` v is flipped (1 - v); Ns below
// 1e-4 becomes 10.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "minirender/io.h"

namespace minirender {

namespace {

std::vector<std::string> split_ws(const std::string& line) {
  std::vector<std::string> out;
  std::istringstream is(line);
  std::string t;
  while (is >> t) out.push_back(t);
  return out;
}

// Split on one character keeping empty fields ("1//2" -> "1", "", "2").
std::vector<std::string> split_char(const std::string& s, char c) {
  std::vector<std::string> out;
  size_t i = 0;
  for (;;) {
    const size_t j = s.find(c, i);
    out.push_back(s.substr(i, j == std::string::npos ? std::string::npos : j - i));
    if (j == std::string::npos) break;
    i = j + 1;
  }
  return out;
}

float to_f(const std::string& s) { return std::strtof(s.c_str(), nullptr); }
int to_i(const std::string& s) { return std::atoi(s.c_str()); }

Vec3 vec3_of(const std::vector<std::string>& p, size_t i) {
  return Vec3(p.size() > i ? to_f(p[i]) : 0.0f, p.size() > i + 1 ? to_f(p[i + 1]) : 0.0f,
              p.size() > i + 2 ? to_f(p[i + 2]) : 0.0f);
}

std::string directory_of(const std::string& path) {
  const size_t k = path.find_last_of('/');
  return k == std::string::npos ? std::string(".") : path.substr(0, k);
}

bool has_extension(const std::string& path, const char* ext) {
  const size_t k = path.find_last_of('.');
  if (k == std::string::npos) return false;
  std::string e = path.substr(k + 1);
  for (char& ch : e) ch = (char)std::tolower((unsigned char)ch);
  return e == ext;
}

Shared<TriMesh> load_stl_ascii(const std::string& filename) {
  std::ifstream f(filename);
  if (!f) return nullptr;
  std::string tag;
  if (!(f >> tag) || tag != "solid") return nullptr;
  auto obj = std::make_shared<TriMesh>();
  int np = 0, iv = 0, in = 0;
  while (f >> tag) {
    if (tag == "endfacet") {
      if (np == 3) {
        obj->indices << iv - 3 << iv - 2 << iv - 1;
        obj->normalsI << in - 1 << in - 1 << in - 1;
      }
      np = 0;
    } else if (tag == "normal" || tag == "vertex") {
      std::string rest;
      std::getline(f, rest);
      const Vec3 v = vec3_of(split_ws(rest), 0);
      if (tag == "normal") {
        obj->normals.push_back(v);
        in++;
      } else {
        obj->vertices.push_back(v);
        np++;
        iv++;
      }
    }
  }
  return obj;
}

Shared<TriMesh> load_stl_binary(const std::string& filename) {
  FILE* f = std::fopen(filename.c_str(), "rb");
  if (!f) return nullptr;
  char header[80];
  int32_t nf = 0;
  if (std::fread(header, 1, 80, f) != 80 || std::fread(&nf, 4, 1, f) != 1 || nf < 0 ||
      nf > 100000000) {
    std::fclose(f);
    return nullptr;
  }
  auto obj = std::make_shared<TriMesh>();
  obj->normals.reserve(nf);
  obj->vertices.reserve((size_t)nf * 3);
  obj->indices.reserve((size_t)nf * 3);
  obj->normalsI.reserve((size_t)nf * 3);
  unsigned char rec[50];
  for (int i = 0; i < nf; i++) {
    if (std::fread(rec, 1, 50, f) != 50) break;
    float v[12];
    std::memcpy(v, rec, sizeof(v));
    obj->normals.push_back(Vec3(v[0], v[1], v[2]));
    for (int j = 0; j < 3; j++) obj->vertices.push_back(Vec3(v[3 + 3 * j], v[4 + 3 * j], v[5 + 3 * j]));
    const int b = 3 * i;
    obj->indices << b << b + 1 << b + 2;
    obj->normalsI << i << i << i;
  }
  std::fclose(f);
  return obj;
}

}  // namespace

// Fan triangulation of -1-terminated polygons (reference x3d.cpp:17-33).
Array<int> triangulateIndices(const Array<int>& idx) {
  Array<int> tris;
  tris.reserve(idx.size());
  int i = 0, j = 0;
  const int n = (int)idx.size();
  while (i < n) {
    for (j = i + 1; j < n - 1; j++) {
      if (idx[j] == -1 || idx[j + 1] == -1) break;
      tris << idx[i] << idx[j] << idx[j + 1];
    }
    i = j + 2;
  }
  return tris;
}

Shared<TriMesh> loadSTL(const std::string& filename) {
  FILE* f = std::fopen(filename.c_str(), "rb");
  if (!f) return nullptr;
  std::fseek(f, 0, SEEK_END);
  const long size = std::ftell(f);
  bool binary = false;
  if (size > 84) {
    std::fseek(f, 80, SEEK_SET);
    int32_t nf = 0;
    if (std::fread(&nf, 4, 1, f) == 1 && (long)80 + 4 + (long)nf * 50 == size) binary = true;
  }
  std::fclose(f);
  if (size < 5) return nullptr;
  return binary ? load_stl_binary(filename) : load_stl_ascii(filename);
}

bool saveSTL(const Shared<TriMesh>& mesh, const std::string& name) {
  FILE* f = std::fopen(name.c_str(), "wb");
  if (!f || !mesh) {
    if (f) std::fclose(f);
    return false;
  }
  char header[80];
  std::memset(header, ' ', sizeof(header));
  std::fwrite(header, 1, 80, f);
  const int32_t nf = (int32_t)(mesh->indices.size() / 3);
  std::fwrite(&nf, 4, 1, f);
  for (size_t i = 0; i + 2 < mesh->indices.size(); i += 3) {
    const Vec3 a = mesh->transform * mesh->vertices[mesh->indices[i]];
    const Vec3 b = mesh->transform * mesh->vertices[mesh->indices[i + 1]];
    const Vec3 c = mesh->transform * mesh->vertices[mesh->indices[i + 2]];
    const Vec3 n = ((c - a) ^ (b - a)).normalized();
    const float rec[12] = {n.x, n.y, n.z, a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z};
    const uint16_t attr = 0;
    std::fwrite(rec, sizeof(rec), 1, f);
    std::fwrite(&attr, 2, 1, f);
  }
  std::fclose(f);
  return true;
}

Shared<SceneNode> loadOBJ(const std::string& filename) {
  std::ifstream file(filename);
  if (!file) return nullptr;
  std::map<std::string, Shared<TriMesh>> meshes;
  std::map<std::string, Shared<Material>> materials;
  materials[""] = std::make_shared<Material>();
  Shared<TriMesh> mesh = std::make_shared<TriMesh>();
  mesh->material = materials[""];
  meshes[""] = mesh;
  // One set of attribute arrays, shared by the per-material meshes
  // (io.cpp:313-315: assignment shares the storage).
  Array<Vec3> vertices, normals;
  Array<Vec2> texcoords;
  std::string line;
  while (std::getline(file, line)) {
    if (!line.empty() && line[0] == '#') continue;
    const std::vector<std::string> parts = split_ws(line);
    if (parts.empty()) continue;
    const std::string& k = parts[0];
    if (k == "v") {
      vertices.push_back(vec3_of(parts, 1));
    } else if (k == "vn") {
      normals.push_back(vec3_of(parts, 1));
    } else if (k == "vt") {
      texcoords.push_back(Vec2(parts.size() > 1 ? to_f(parts[1]) : 0.0f,
                               1.0f - (parts.size() > 2 ? to_f(parts[2]) : 0.0f)));
    } else if (k == "f") {
      for (size_t i = 1; i < parts.size(); i++) {
        const std::vector<std::string> ix = split_char(parts[i], '/');
        mesh->indices.push_back(to_i(ix[0]) - 1);
        if (ix.size() > 1) mesh->texcoordsI.push_back(to_i(ix[1]) - 1);
        if (ix.size() > 2) mesh->normalsI.push_back(to_i(ix[2]) - 1);
      }
      mesh->indices.push_back(-1);
      mesh->texcoordsI.push_back(-1);
      mesh->normalsI.push_back(-1);
    } else if (k == "usemtl" && parts.size() > 1) {
      const std::string& name = parts[1];
      auto it = meshes.find(name);
      if (it == meshes.end()) {
        auto m = std::make_shared<TriMesh>();
        auto mt = materials.find(name);
        m->material = mt != materials.end() ? mt->second : materials[""];
        meshes[name] = m;
        mesh = m;
      } else {
        mesh = it->second;
      }
    } else if (k == "mtllib" && parts.size() > 1) {
      std::ifstream mf(directory_of(filename) + "/" + parts[1]);
      Shared<Material> mat = materials[""];
      std::string ml;
      while (std::getline(mf, ml)) {
        const std::vector<std::string> p = split_ws(ml);
        if (p.empty()) continue;
        if (p[0] == "newmtl" && p.size() > 1) {
          mat = std::make_shared<Material>();
          materials[p[1]] = mat;
        } else if (p[0] == "Kd") {
          mat->diffuse = vec3_of(p, 1);
        } else if (p[0] == "Ks") {
          mat->specular = vec3_of(p, 1);
        } else if (p[0] == "Ke") {
          mat->emissive = vec3_of(p, 1);
        } else if (p[0] == "Ns" && p.size() > 1) {
          mat->shininess = to_f(p[1]);
          if (mat->shininess < 0.0001f) mat->shininess = 10;
        } else if (p[0] == "d" && p.size() > 1) {
          mat->opacity = to_f(p[1]);
        } else if (p[0] == "map_Kd" && p.size() > 1) {
          mat->textureName = p[1];
        }
      }
    }
  }
  for (auto& e : materials)
    if (!e.second->textureName.empty())
      e.second->texture = loadPPM(directory_of(filename) + "/" + e.second->textureName);
  auto node = std::make_shared<SceneNode>();
  for (auto& e : meshes) {
    TriMesh& m = *e.second;
    m.indices = triangulateIndices(m.indices);
    m.texcoordsI = triangulateIndices(m.texcoordsI);
    m.normalsI = triangulateIndices(m.normalsI);
    m.vertices = vertices;
    m.normals = normals;
    m.texcoords = texcoords;
    if (m.normals.empty()) {  // one normal per face, in arrays of this mesh
      m.normals.dup();
      m.normalsI.clear();
      for (size_t i = 0, j = 0; i + 2 < m.indices.size(); i += 3, j++) {
        const Vec3 a = m.vertices[m.indices[i]];
        const Vec3 b = m.vertices[m.indices[i + 1]];
        const Vec3 c = m.vertices[m.indices[i + 2]];
        m.normals.push_back(((b - a) ^ (c - a)).normalized());
        m.normalsI << (int)j << (int)j << (int)j;
      }
    }
    node->children.push_back(e.second);
  }
  return node;
}

Shared<SceneNode> loadMesh(const std::string& filename) {
  if (has_extension(filename, "stl")) {
    Shared<TriMesh> mesh = loadSTL(filename);
    if (!mesh) return nullptr;
    auto node = std::make_shared<SceneNode>();
    node->children.push_back(mesh);
    return node;
  }
  if (has_extension(filename, "obj")) return loadOBJ(filename);
  if (has_extension(filename, "x3d")) return loadX3D(filename);
  return std::make_shared<SceneNode>();
}

bool saveXYZ(const Array2<Vec3>& points, const std::string& filename, const Matrix4& m) {
  FILE* f = std::fopen(filename.c_str(), "w");
  if (!f) return false;
  for (int i = 0; i < points.rows(); i++)
    for (int j = 0; j < points.cols(); j++) {
      Vec3 p = points(i, j);
      if (p.z < -1e-5f) {
        p = m * p;
        std::fprintf(f, "%f %f %f\n", p.x, p.y, p.z);
      }
    }
  std::fclose(f);
  return true;
}

}  // namespace minirender

namespace minirender {

// samples/main.cpp:13-28: home the cursor, then per pixel a background-colour
// escape and a space.  The channel value is int(v * 255) (truncation, no
// clamping, as Vec3::with<int>()); escapes take it clamped to 0..255.
std::string consoleImage(const Array2<Vec3>& image, bool colors256) {
  auto chan = [](float v) {
    const int c = (int)(v * 255);
    return c < 0 ? 0 : (c > 255 ? 255 : c);
  };
  std::string out = "\x1b[H";
  out.reserve((size_t)image.rows() * image.cols() * 20 + 16);
  char buf[48];
  for (int i = 0; i < image.rows(); i++) {
    for (int j = 0; j < image.cols(); j++) {
      const Vec3& p = image(i, j);
      const int r = chan(p.x), g = chan(p.y), b = chan(p.z);
      if (colors256)
        std::snprintf(buf, sizeof(buf), "\x1b[48;5;%dm ",
                      16 + 36 * (r * 6 / 256) + 6 * (g * 6 / 256) + b * 6 / 256);
      else
        std::snprintf(buf, sizeof(buf), "\x1b[48;2;%d;%d;%dm ", r, g, b);
      out += buf;
    }
    out += "\x1b[0m\n";
  }
  return out;
}

void consolePaint(const Array2<Vec3>& image, bool colors256) {
  const std::string s = consoleImage(image, colors256);
  std::fwrite(s.data(), 1, s.size(), stdout);
  std::fflush(stdout);
}

}  // namespace minirender
