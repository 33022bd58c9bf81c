// primitives.h -- procedural meshes of the minirender API (reference:
// include/minirender/primitives.h, src/primitives.cpp), same vertex and index
// order.
#ifndef MINIRENDER_PRIMITIVES_H
#define MINIRENDER_PRIMITIVES_H

#include "Scene.h"

namespace minirender {

Shared<TriMesh> createCube(float size = 1.0f);
Shared<TriMesh> createCylinder(float radius, float height, int segments = 32,
                               int heightSegments = 1, bool caps = true);
Shared<TriMesh> createSphere(float radius, int latSegments = 16, int longSegments = 32);

}  // namespace minirender

#endif
