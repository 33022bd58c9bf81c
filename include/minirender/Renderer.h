// Renderer.h -- minirender's Renderer on MI355X (reference:
// include/minirender/Renderer.h:9-64).  Same methods and meaning; render()
// runs the HIP pipeline of minirender_amd through its C-ABI
// (include/minirender_amd.h).  Meshes and textures stay resident on the GPU
// between frames, and so does the last frame: each getter copies back only
// the image it returns.
#ifndef MINIRENDER_RENDERER_H
#define MINIRENDER_RENDERER_H

#include <cstdint>
#include <unordered_map>
#include <vector>

#include "../minirender_amd.h"
#include "Scene.h"

namespace minirender {

// Projection matrices, Renderer.cpp:40-83 (row-major, OpenGL conventions).
Matrix4 projectionOrtho(float l, float r, float b, float t, float n, float f);
Matrix4 projectionPerspective(float l, float r, float b, float t, float n, float f);
Matrix4 projectionFrustum(float fov, float aspect, float n, float f);
Matrix4 projectionFrustumH(float fov, float aspect, float n, float f);
Matrix4 projectionOrtho(float fov, float aspect, float n, float f);
Matrix4 projectionCV(const Matrix4& K, float w, float h, float n, float f);

class Renderer {
 public:
  Renderer();
  ~Renderer();
  Renderer(const Renderer&) = delete;
  Renderer& operator=(const Renderer&) = delete;

  void setSize(int w, int h);
  float aspect() const { return (float)_image.cols() / _image.rows(); }
  void setScene(Shared<Scene> scene) { _scene = scene; }
  void setProjection(const Matrix4& m) { _projection = m; }
  void setView(const Matrix4& m) { _view = m; }
  void setLight(const Vec3& v, bool point = false) { _light = v; _lightIsPoint = point; }
  void setMaterial(Shared<Material> material) { _material = material; }
  void setLighting(bool on) { _lighting = on; }
  void setTexturing(bool on) { _texturing = on; }
  void setSaveNormals(bool on) { _saveNormals = on; }
  void setBackground(const Vec3& color) { _bgcolor = color; }
  void clear();
  // Renderer::render (Renderer.cpp:311-332) on the GPU.
  void render();
  // Immediate mode (Renderer.cpp:334-381, 163-309): draws after the last
  // render() in call order, composited with the strict depth test when the
  // frame is read.  Vertices of paintTriangle are in eye space.
  void paintMesh(TriMesh* mesh, const Matrix4& transform = Matrix4::identity());
  void paintTriangle(const Vertex& a, const Vertex& b, const Vertex& c, bool world = true);
  Array2<float> getDepth();
  Array2<Vec3> getImage();
  Array2<Vec3> getRangeImage();   // Renderer.cpp:388-415, on the host
  Array2<Vec3> getNormalsImage();
  // RGB8 of the last frame, quantized as savePPM (io.cpp:358-361), rows x cols x 3.
  std::vector<uint8_t> getImageRGB8();
  // Draw-order index of the visible triangle per pixel (0xFFFFFFFF: none).
  std::vector<uint32_t> getWinners();

  // The C-ABI frame render() submits for the current scene and settings:
  // draws reference mesh / texture handles of this renderer's context.
  // Valid until the next call on this renderer.
  const mr_frame& describeFrame();
  // The TriMesh behind a mesh handle of describeFrame().
  const TriMesh* meshOfHandle(int handle) const;
  mr_ctx* context();
  // Status of the last GPU call (0 = ok, MR_ERR_*).
  int lastStatus() const { return _status; }

 private:
  struct MeshEntry {
    int32_t handle = -1;
    uint64_t fingerprint = 0;       // content hash at the last upload
    uint64_t stamps[12] = {};       // (storage id, version) of the six arrays then
    const TriMesh* mesh = nullptr;
  };
  struct TexEntry {
    int32_t handle = -1;
    uint64_t fingerprint = 0;
  };
  int32_t meshHandle(TriMesh* mesh);
  int32_t textureHandle(const Material* m);
  void appendDraw(TriMesh* mesh, const Matrix4& transform, const Material& mat,
                  std::vector<mr_draw>& out);
  void fillFrame(mr_frame& f, const std::vector<mr_draw>& draws);
  void flushImmediate();
  bool ensureDeviceOutputs();
  void freeDeviceOutputs();
  void fetch(int kind);

  Array2<Vec3> _image;
  Array2<float> _depth;
  Array2<Vec3> _pnormals;
  std::vector<uint8_t> _rgb8;
  std::vector<uint32_t> _winners;
  // render() writes into device framebuffers; a getter copies back its own
  // image when the host copy is stale (_fresh false).
  enum { kRGB8, kDepth, kWinner, kRGB, kNormals, kOutputs };
  void* _devOut[kOutputs] = {};
  size_t _devPixels = 0;
  bool _fresh[kOutputs] = {true, true, true, true, true};
  Matrix4 _view, _projection;
  Vec3 _light, _bgcolor;
  float _ambient = 0.1f;
  Shared<Scene> _scene;
  Shared<Material> _material, _defmaterial;
  bool _lighting = true, _texturing = true, _lightIsPoint = false;
  bool _saveNormals = false;  // uninitialized in the reference; off here

  mr_ctx* _ctx = nullptr;
  int _status = 0;
  std::unordered_map<const TriMesh*, MeshEntry> _meshes;
  std::unordered_map<const Material*, TexEntry> _textures;
  std::unordered_map<int32_t, const TriMesh*> _handleMesh;
  // meshes of the scene being described, which may lend attribute arrays
  std::vector<const TriMesh*> _frameMeshes;
  bool _sharing = false;
  std::vector<mr_draw> _draws;
  mr_frame _frame{};
  // immediate-mode draws since the last render()
  std::vector<mr_draw> _immediate;
  std::vector<Shared<TriMesh>> _immediateMeshes;
};

}  // namespace minirender

#endif
