// x3d.cpp -- X3D scene loader of the minirender API (reference:
// src/x3d.cpp:35-196), with the small XML reader it needs.
//
// What is read, as in the reference: the <Scene> children of an <X3D>
// document; <Transform>/<Group> nodes (translation * rotation * scale,
// attribute defaults "0 0 0", "0 0 1 0", "1 1 1"); <Shape> with its
// <Appearance> (<Material> colours, shininess x 8, <ImageTexture> url with
// the extension replaced by .ppm, read next to the scene file) and an
// <IndexedFaceSet> (polygons fan-triangulated, -1 separators) or
// <IndexedTriangleSet>; <Coordinate point>, <Normal vector>,
// <TextureCoordinate point> (v flipped to 1 - v; commas count as spaces);
// <Inline url> (a scene file read relative to the current one, after which
// the reader continues with the inlined document, as the reference's reader
// object does); DEF/USE references, resolved by a depth-first search of the
// current document.  Meshes without normals get one normal per face,
// ((b - a) x (c - a)) normalised; meshes without texcoords get one (0, 0)
// texcoord for every index.
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "minirender/io.h"

namespace minirender {

namespace {

// A parsed XML element: tag, attributes in document order, child elements.
// Text content, comments, processing instructions and DOCTYPE are skipped.
struct XmlNode {
  std::string tag;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<std::shared_ptr<XmlNode>> children;

  const std::string* attr(const std::string& name) const {
    for (const auto& a : attrs)
      if (a.first == name) return &a.second;
    return nullptr;
  }
  std::string get(const std::string& name, const char* dflt = "") const {
    const std::string* v = attr(name);
    return v ? *v : std::string(dflt);
  }
  // First child element with this tag.
  std::shared_ptr<XmlNode> child(const std::string& t) const {
    for (const auto& c : children)
      if (c->tag == t) return c;
    return nullptr;
  }
};
using XmlPtr = std::shared_ptr<XmlNode>;

std::string decode_entities(const std::string& s) {
  std::string out;
  out.reserve(s.size());
  for (size_t i = 0; i < s.size(); i++) {
    if (s[i] != '&') {
      out += s[i];
      continue;
    }
    const size_t semi = s.find(';', i);
    if (semi == std::string::npos) {
      out += s[i];
      continue;
    }
    const std::string ent = s.substr(i + 1, semi - i - 1);
    if (ent == "quot") out += '"';
    else if (ent == "amp") out += '&';
    else if (ent == "lt") out += '<';
    else if (ent == "gt") out += '>';
    else if (ent == "apos") out += '\'';
    else if (!ent.empty() && ent[0] == '#') {
      const long code = ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X')
                            ? std::strtol(ent.c_str() + 2, nullptr, 16)
                            : std::strtol(ent.c_str() + 1, nullptr, 10);
      out += code > 0 && code < 128 ? (char)code : '?';
    } else {
      out += s.substr(i, semi - i + 1);
    }
    i = semi;
  }
  return out;
}

// Recursive-descent reader over the whole file text.
class XmlReader {
 public:
  explicit XmlReader(std::string text) : s_(std::move(text)) {}

  // The document element, or null on malformed input.
  XmlPtr document() {
    XmlPtr root;
    while (skip_misc()) {
      if (s_.compare(p_, 1, "<") != 0) return nullptr;
      XmlPtr e = element();
      if (!e) return nullptr;
      if (!root) root = e;
    }
    return root;
  }

 private:
  std::string s_;
  size_t p_ = 0;

  void skip_ws() {
    while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) p_++;
  }
  // Skips whitespace, comments, <?...?> and <!...> outside elements;
  // returns whether input remains.
  bool skip_misc() {
    for (;;) {
      skip_ws();
      if (p_ >= s_.size()) return false;
      if (s_.compare(p_, 4, "<!--") == 0) {
        const size_t e = s_.find("-->", p_ + 4);
        p_ = e == std::string::npos ? s_.size() : e + 3;
      } else if (s_.compare(p_, 2, "<?") == 0) {
        const size_t e = s_.find("?>", p_ + 2);
        p_ = e == std::string::npos ? s_.size() : e + 2;
      } else if (s_.compare(p_, 2, "<!") == 0) {
        int depth = 0;  // DOCTYPE may hold an internal subset [...]
        for (; p_ < s_.size(); p_++) {
          if (s_[p_] == '[') depth++;
          else if (s_[p_] == ']') depth--;
          else if (s_[p_] == '>' && depth <= 0) {
            p_++;
            break;
          }
        }
      } else {
        return true;
      }
    }
  }
  std::string name() {
    const size_t b = p_;
    while (p_ < s_.size() && !std::isspace((unsigned char)s_[p_]) && s_[p_] != '>' &&
           s_[p_] != '/' && s_[p_] != '=')
      p_++;
    return s_.substr(b, p_ - b);
  }
  // At '<' of a start tag.
  XmlPtr element() {
    p_++;
    auto e = std::make_shared<XmlNode>();
    e->tag = name();
    if (e->tag.empty()) return nullptr;
    for (;;) {  // attributes
      skip_ws();
      if (p_ >= s_.size()) return nullptr;
      if (s_.compare(p_, 2, "/>") == 0) {
        p_ += 2;
        return e;
      }
      if (s_[p_] == '>') {
        p_++;
        break;
      }
      std::string key = name();
      if (key.empty()) return nullptr;
      skip_ws();
      if (p_ >= s_.size() || s_[p_] != '=') return nullptr;
      p_++;
      skip_ws();
      if (p_ >= s_.size() || (s_[p_] != '"' && s_[p_] != '\'')) return nullptr;
      const char q = s_[p_++];
      const size_t end = s_.find(q, p_);
      if (end == std::string::npos) return nullptr;
      e->attrs.emplace_back(key, decode_entities(s_.substr(p_, end - p_)));
      p_ = end + 1;
    }
    for (;;) {  // content
      const size_t lt = s_.find('<', p_);
      if (lt == std::string::npos) return nullptr;
      p_ = lt;
      if (s_.compare(p_, 2, "</") == 0) {
        const size_t gt = s_.find('>', p_);
        if (gt == std::string::npos) return nullptr;
        p_ = gt + 1;
        return e;
      }
      if (s_.compare(p_, 4, "<!--") == 0 || s_.compare(p_, 2, "<?") == 0 ||
          s_.compare(p_, 2, "<!") == 0) {
        if (s_.compare(p_, 9, "<![CDATA[") == 0) {
          const size_t ce = s_.find("]]>", p_);
          if (ce == std::string::npos) return nullptr;
          p_ = ce + 3;
        } else if (!skip_misc()) {
          return nullptr;
        }
        continue;
      }
      XmlPtr c = element();
      if (!c) return nullptr;
      e->children.push_back(c);
    }
  }
};

XmlPtr read_xml(const std::string& filename) {
  std::ifstream f(filename, std::ios::binary);
  if (!f) return nullptr;
  std::stringstream ss;
  ss << f.rdbuf();
  return XmlReader(ss.str()).document();
}

XmlPtr find_def(const XmlPtr& e, const std::string& def) {
  if (!e) return nullptr;
  const std::string* d = e->attr("DEF");
  if (d && *d == def) return e;
  for (const auto& c : e->children)
    if (XmlPtr r = find_def(c, def)) return r;
  return nullptr;
}

std::vector<float> floats_of(std::string s) {
  for (char& ch : s)
    if (ch == ',') ch = ' ';
  std::vector<float> v;
  std::istringstream in(s);
  std::string tok;
  while (in >> tok) v.push_back(std::strtof(tok.c_str(), nullptr));
  return v;
}

std::vector<int> ints_of(std::string s) {
  for (char& ch : s)
    if (ch == ',') ch = ' ';
  std::vector<int> v;
  std::istringstream in(s);
  std::string tok;
  while (in >> tok) v.push_back(std::atoi(tok.c_str()));
  return v;
}

Vec3 vec3_of(const std::string& s) {
  std::vector<float> v = floats_of(s);
  v.resize(3, 0.0f);
  return Vec3(v[0], v[1], v[2]);
}

std::string directory_of(const std::string& path) {
  const size_t k = path.find_last_of('/');
  return k == std::string::npos ? std::string(".") : path.substr(0, k);
}

std::string without_extension(const std::string& path) {
  const size_t slash = path.find_last_of('/');
  const size_t dot = path.find_last_of('.');
  if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return path;
  return path.substr(0, dot);
}

std::string without_quotes(std::string s) {
  std::string out;
  for (char ch : s)
    if (ch != '"') out += ch;
  return out;
}

class X3dReader {
 public:
  Shared<SceneNode> load(const std::string& filename) {
    doc_ = read_xml(filename);
    if (!doc_ || doc_->tag != "X3D") return nullptr;
    XmlPtr scene = doc_->child("Scene");
    if (!scene) return nullptr;
    auto root = std::make_shared<SceneNode>();
    filename_ = filename;
    for (const auto& e : scene->children)
      if (Shared<SceneNode> n = item(e)) root->children.push_back(n);
    return root;
  }

 private:
  XmlPtr doc_;
  std::string filename_;

  // The element itself, or the DEF it USEs.
  XmlPtr resolve(const XmlPtr& e) const {
    if (!e) return nullptr;
    if (const std::string* use = e->attr("USE")) return find_def(doc_, *use);
    return e;
  }

  Shared<SceneNode> item(const XmlPtr& e) {
    if (e->tag == "Transform" || e->tag == "Group") return group(*e);
    if (e->tag == "Shape") return shape(*e);
    if (e->tag == "Inline")
      return load(directory_of(filename_) + "/" + without_quotes(e->get("url")));
    return nullptr;
  }

  Shared<SceneNode> group(const XmlNode& e) {
    std::vector<float> r = floats_of(e.get("rotation", "0 0 1 0"));
    std::vector<float> t = floats_of(e.get("translation", "0 0 0"));
    std::vector<float> s = floats_of(e.get("scale", "1 1 1"));
    r.resize(4, 0.0f);
    t.resize(3, 0.0f);
    s.resize(3, 0.0f);
    Vec3 axis(r[0], r[1], r[2]);
    const float len = axis.length();
    const Matrix4 rot = len > 0 ? Matrix4::rotate(axis / len, r[3]) : Matrix4::identity();
    auto node = std::make_shared<SceneNode>();
    node->transform = Matrix4::translate(t[0], t[1], t[2]) * rot *
                      Matrix4::scale(Vec3(s[0], s[1], s[2]));
    for (const auto& c : e.children)
      if (Shared<SceneNode> n = item(c)) node->children.push_back(n);
    return node;
  }

  Shared<SceneNode> shape(const XmlNode& e) {
    auto mesh = std::make_shared<TriMesh>();
    mesh->material = std::make_shared<Material>();
    XmlPtr app = resolve(e.child("Appearance"));
    if (XmlPtr mat = app ? resolve(app->child("Material")) : nullptr) {
      mesh->material->diffuse = vec3_of(mat->get("diffuseColor", "0.7 0.75 0.8"));
      mesh->material->specular = vec3_of(mat->get("specularColor", "0.4 0.4 0.4"));
      mesh->material->emissive = vec3_of(mat->get("emissiveColor", "0 0 0"));
      mesh->material->shininess = std::strtof(mat->get("shininess", "0.5").c_str(), nullptr) * 8;
    }
    if (XmlPtr tex = app ? resolve(app->child("ImageTexture")) : nullptr) {
      mesh->material->textureName = without_extension(tex->get("url")) + ".ppm";
      mesh->material->texture =
          loadPPM(directory_of(filename_) + "/" + mesh->material->textureName);
    }
    XmlPtr ifs = resolve(e.child("IndexedFaceSet"));
    XmlPtr its = resolve(e.child("IndexedTriangleSet"));
    const XmlPtr g = ifs ? ifs : its;
    if (!g) return mesh;
    auto field = [&](const char* tag, const char* attr) {
      XmlPtr c = resolve(g->child(tag));
      return c ? c->get(attr) : std::string();
    };
    const std::vector<float> verts = floats_of(field("Coordinate", "point"));
    const std::vector<float> normals = floats_of(field("Normal", "vector"));
    const std::vector<float> uvs = floats_of(field("TextureCoordinate", "point"));
    for (size_t i = 0; i + 2 < verts.size(); i += 3)
      mesh->vertices.push_back(Vec3(verts[i], verts[i + 1], verts[i + 2]));
    for (size_t i = 0; i + 2 < normals.size(); i += 3)
      mesh->normals.push_back(Vec3(normals[i], normals[i + 1], normals[i + 2]));
    for (size_t i = 0; i + 1 < uvs.size(); i += 2)
      mesh->texcoords.push_back(Vec2(uvs[i], 1 - uvs[i + 1]));
    if (ifs) {
      mesh->indices = triangulateIndices(ints_of(g->get("coordIndex")));
      const Array<int> ti = ints_of(g->get("texCoordIndex"));
      const Array<int> ni = ints_of(g->get("normalIndex"));
      mesh->texcoordsI = ti.empty() ? mesh->indices.clone() : triangulateIndices(ti);
      mesh->normalsI = ni.empty() ? mesh->indices.clone() : triangulateIndices(ni);
    } else {
      mesh->indices = ints_of(g->get("index"));
      mesh->texcoordsI = mesh->indices.clone();
      mesh->normalsI = mesh->indices.clone();
    }
    const int nv = (int)mesh->vertices.size();
    for (int v : mesh->indices)
      if (v < 0 || v >= nv) return nullptr;  // the reference would read out of bounds
    if (mesh->normals.empty()) {  // one normal per face
      mesh->normalsI.clear();
      for (size_t i = 0, j = 0; i + 2 < mesh->indices.size(); i += 3, j++) {
        const Vec3 a = mesh->vertices[mesh->indices[i]];
        const Vec3 b = mesh->vertices[mesh->indices[i + 1]];
        const Vec3 c = mesh->vertices[mesh->indices[i + 2]];
        mesh->normals.push_back(((b - a) ^ (c - a)).normalized());
        mesh->normalsI << (int)j << (int)j << (int)j;
      }
    }
    if (mesh->texcoords.empty()) {
      mesh->texcoords.push_back(Vec2(0, 0));
      mesh->texcoordsI = Array<int>(mesh->indices.length());  // zeros
    }
    return mesh;
  }
};

}  // namespace

Shared<SceneNode> loadX3D(const std::string& filename) { return X3dReader().load(filename); }

}  // namespace minirender
