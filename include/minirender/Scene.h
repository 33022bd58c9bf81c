// Scene.h -- scene data model of the minirender API (reference:
// include/minirender/Scene.h:14-95): TriMesh with separate position / normal
// / texcoord arrays and index arrays, materials, a transform hierarchy.
#ifndef MINIRENDER_SCENE_H
#define MINIRENDER_SCENE_H

#include <string>

#include "vecmath.h"

namespace minirender {

struct Vertex {
  Vec3 position;
  Vec3 normal;
  Vec2 uv;
  Vertex() {}
  Vertex(const Vec3& p) : position(p), normal(0, 0, 1), uv(0, 0) {}
  Vertex(const Vec3& p, const Vec3& n) : position(p), normal(n), uv(0, 0) {}
  Vertex(const Vec3& p, const Vec3& n, const Vec2& t) : position(p), normal(n), uv(t) {}
};

struct BBox {
  Vec3 pmin, pmax;
  BBox() : pmin(INFINITY, INFINITY, INFINITY), pmax(-INFINITY, -INFINITY, -INFINITY) {}
  BBox& operator+=(const BBox& b) { pmin = min(pmin, b.pmin); pmax = max(pmax, b.pmax); return *this; }
  BBox& operator+=(const Vec3& p) { pmin = min(pmin, p); pmax = max(pmax, p); return *this; }
  Vec3 size() const { return max(pmax - pmin, Vec3(0, 0, 0)); }
  Vec3 center() const { return (pmax + pmin) / 2; }
};

struct TriMesh;

// Defaults of Scene.cpp:65-71.  `opacity` is carried but not rendered, as
// in the reference.
struct Material : public RefCounted {
  Vec3 diffuse{0.7f, 0.7f, 0.9f};
  Vec3 specular{0.8f, 0.8f, 0.8f};
  Vec3 emissive{0, 0, 0};
  float shininess = 12.0f;
  float opacity = 1.0f;
  Array2<Vec3> texture;
  std::string textureName;
};

struct Renderable {
  TriMesh* mesh = nullptr;
  Matrix4 transform;
  Renderable() {}
  Renderable(TriMesh* m, const Matrix4& t) : mesh(m), transform(t) {}
};

struct SceneNode : public RefCounted {
  bool visible = true;  // not read by the renderer (as in the reference)
  Matrix4 transform = Matrix4::identity();
  Array<Shared<SceneNode>> children;
  virtual ~SceneNode() {}
  // Depth-first flattening, Scene.cpp:13-20.
  virtual void collectShapes(Array<Renderable>& list, const Matrix4& xform);
  virtual BBox getBbox(const Matrix4& xform = Matrix4::identity()) const;
};

struct Shape : public SceneNode {
  Shared<Material> material;
  virtual void applyTransform() {}
};

struct TriMesh : public Shape {
  Array<Vec3> vertices;
  Array<Vec3> normals;
  Array<Vec2> texcoords;
  Array<int> indices;
  Array<int> normalsI;
  Array<int> texcoordsI;
  // A TriMesh emits itself, then its children (Scene.cpp:30-36).
  void collectShapes(Array<Renderable>& list, const Matrix4& xform) override;
  BBox getBbox(const Matrix4& xform = Matrix4::identity()) const override;
  void applyTransform() override;  // Scene.cpp:53-63
};

struct Scene : public SceneNode {
  float ambientLight = 0.1f;  // Scene.cpp:73-76
  Vec3 light;                 // not read by the renderer (as in the reference)
  Scene& add(const Shared<SceneNode>& node) { children << node; return *this; }
};

}  // namespace minirender

#endif
