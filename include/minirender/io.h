// io.h -- image files of the minirender API (reference:
// include/minirender/io.h:10-22): PPM (P6) save with the reference's RGB8
// quantization and load with texel = byte / 255.  Mesh and scene file
// loaders (STL / OBJ / X3D) are not part of this library.
#ifndef MINIRENDER_IO_H
#define MINIRENDER_IO_H

#include <string>

#include "vecmath.h"

namespace minirender {

bool savePPM(const Array2<Vec3>& image, const std::string& filename);
Array2<Vec3> loadPPM(const std::string& filename);

}  // namespace minirender

#endif
