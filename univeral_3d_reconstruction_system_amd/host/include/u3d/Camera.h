// Drop-in declaration of the reference's camera types (engine/include/u3d/Camera.h), unchanged
// in names, members and meaning so callers of the trainer compile against this library as is.
#pragma once

#include <array>
#include <functional>
#include <string>

namespace u3d {

// Pinhole camera, world-to-camera pose in the OpenCV convention (x right, y down, z forward):
// p_cam = R * p_world + t, R row-major.
struct Camera {
    int width = 0;
    int height = 0;
    double fx = 0, fy = 0, cx = 0, cy = 0;
    std::array<double, 9> R{1, 0, 0, 0, 1, 0, 0, 0, 1};
    std::array<double, 3> t{0, 0, 0};

    std::array<double, 3> center() const;                  // -R^T t
    Camera scaled(double scale) const;                      // camera of a resized image
    bool project(const std::array<double, 3> &world, double &u, double &v) const;
};

struct Progress {
    std::string stage;
    double fraction = 0;
    std::string message;
};
using ProgressCallback = std::function<void(const Progress &)>;
using LogCallback = std::function<void(const std::string &)>;

} // namespace u3d
