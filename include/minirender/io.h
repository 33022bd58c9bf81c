// io.h -- files of the minirender API (reference: include/minirender/io.h:
// 10-22): PPM (P6) save with the reference's RGB8 quantization and load with
// texel = byte / 255; STL (ASCII / binary), OBJ (+ MTL) and X3D loaders,
// STL and XYZ point writers.
#ifndef MINIRENDER_IO_H
#define MINIRENDER_IO_H

#include <string>
#include <vector>

#include "Scene.h"
#include "vecmath.h"

namespace minirender {

bool savePPM(const Array2<Vec3>& image, const std::string& filename);

// The console output of the reference's render sample (samples/main.cpp:
// 13-28): cursor to the top left, then one space per pixel on a background
// of the pixel's colour, int(colour * 255) per channel, a line per row.
// consoleImage returns that text (24-bit colour escapes, or the nearest of
// the 6x6x6 cube of 256-colour terminals with colors256); consolePaint
// writes it to stdout.
std::string consoleImage(const Array2<Vec3>& image, bool colors256 = false);
void consolePaint(const Array2<Vec3>& image, bool colors256 = false);
Array2<Vec3> loadPPM(const std::string& filename);

// io.cpp:114-133: .stl -> a node holding the mesh, .obj -> loadOBJ,
// .x3d -> loadX3D, anything else -> an empty node.
Shared<SceneNode> loadMesh(const std::string& filename);
// io.cpp:135-152 (binary when the size is 84 + 50 * facets).
Shared<TriMesh> loadSTL(const std::string& filename);
// io.cpp:154-181.
bool saveSTL(const Shared<TriMesh>& mesh, const std::string& name);
// io.cpp:185-335: one mesh per material, fan-triangulated faces.
Shared<SceneNode> loadOBJ(const std::string& filename);
// io.cpp:417-431: points with z < -1e-5, transformed by m, "%f %f %f" lines.
bool saveXYZ(const Array2<Vec3>& points, const std::string& filename,
             const Matrix4& m = Matrix4::identity());
// x3d.cpp:35-196: Transform/Group/Shape/Inline scene of an X3D file (null if
// it is not an X3D document with a Scene).
Shared<SceneNode> loadX3D(const std::string& filename);
// x3d.cpp:17-33: fan triangulation of -1-terminated polygons.
Array<int> triangulateIndices(const Array<int>& indices);

}  // namespace minirender

#endif
