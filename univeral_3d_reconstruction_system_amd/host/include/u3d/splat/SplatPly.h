// Drop-in declaration of the reference's model and point-cloud files
// (engine/include/u3d/splat/SplatPly.h): same functions and defaults.
#pragma once

#include "u3d/splat/GaussianModel.h"

#include <string>
#include <vector>

namespace u3d {

// The 3D Gaussian Splatting .ply layout, binary little endian: x y z nx ny nz f_dc_0..2 opacity
// scale_0..2 rot_0..3.
bool writeSplatPly(const std::string &path, const GaussianModel &model, std::string *error = nullptr);

// A splat .ply, or a coloured point cloud (x y z [red green blue]) turned into small opaque
// Gaussians; ASCII or binary of either byte order. isSplat (optional) tells which it was.
bool readSplatPly(const std::string &path, GaussianModel &model, std::string *error = nullptr,
                  bool *isSplat = nullptr);

// Coloured point cloud (rgb 0..1) as x y z red green blue (uchar).
bool writePointsPly(const std::string &path, const std::vector<float> &xyz,
                    const std::vector<float> &rgb, std::string *error = nullptr);
bool readPointsPly(const std::string &path, std::vector<float> &xyz, std::vector<float> &rgb,
                   std::string *error = nullptr);

} // namespace u3d
