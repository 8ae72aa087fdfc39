// Camera helpers of the drop-in API (engine/src/Camera.cpp semantics).
#include "u3d/Camera.h"

namespace u3d {

std::array<double, 3> Camera::center() const
{
    std::array<double, 3> c{};
    for (int k = 0; k < 3; ++k)
        c[k] = -(R[k] * t[0] + R[3 + k] * t[1] + R[6 + k] * t[2]);
    return c;
}

Camera Camera::scaled(double scale) const
{
    // Pixel centres are kept: continuous x maps to (x + 0.5) * s - 0.5 (Camera.cpp:12-23).
    Camera c = *this;
    c.width = int(width * scale + 0.5);
    c.height = int(height * scale + 0.5);
    c.fx = fx * scale;
    c.fy = fy * scale;
    c.cx = (cx + 0.5) * scale - 0.5;
    c.cy = (cy + 0.5) * scale - 0.5;
    return c;
}

bool Camera::project(const std::array<double, 3> &w, double &u, double &v) const
{
    double p[3];
    for (int r = 0; r < 3; ++r)
        p[r] = R[r * 3] * w[0] + R[r * 3 + 1] * w[1] + R[r * 3 + 2] * w[2] + t[r];
    if (p[2] <= 1e-9)
        return false;
    u = fx * p[0] / p[2] + cx;
    v = fy * p[1] / p[2] + cy;
    return true;
}

} // namespace u3d
