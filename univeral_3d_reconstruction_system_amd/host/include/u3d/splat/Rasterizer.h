// Drop-in declaration of the reference's differentiable rasterizer
// (engine/include/u3d/splat/Rasterizer.h). The public interface is unchanged; the work runs on
// the GPU through libu3d_hip (include/u3d_hip.h) and host vectors are copied in and out.
#pragma once

#include "u3d/Camera.h"
#include "u3d/ThreadPool.h"
#include "u3d/splat/GaussianModel.h"

#include <memory>
#include <vector>

namespace u3d {

struct RenderCamera {
    int width = 0;
    int height = 0;
    float fx = 0, fy = 0, cx = 0, cy = 0;
    float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}; // world-to-camera, row-major
    float t[3] = {0, 0, 0};

    static RenderCamera fromCamera(const Camera &camera);
};

struct RasterSettings {
    int tileSize = 16;             // any size; multiples of 16 take the fastest kernels
    float nearPlane = 0.01f;
    float eps2d = 0.3f;
    float background[3] = {0.0f, 0.0f, 0.0f};
    float minAlpha = 1.0f / 255.0f;
    float minTransmittance = 1e-4f;
};

class Rasterizer
{
public:
    explicit Rasterizer(ThreadPool &pool);
    ~Rasterizer();
    Rasterizer(const Rasterizer &) = delete;
    Rasterizer &operator=(const Rasterizer &) = delete;

    // Renders height*width*3 interleaved RGB and keeps the state for backward().
    void forward(const GaussianModel &model, const RenderCamera &camera, const RasterSettings &settings,
                 std::vector<float> &rgb);
    // ADDS the gradients of the last forward() to `grads` (model-sized); if mean2dGradNdc is
    // non-null (size N) adds |dL/d mean2d| in NDC, the densification statistic.
    void backward(const GaussianModel &model, const float *dLdRgb, GaussianModel &grads,
                  float *mean2dGradNdc = nullptr);
    // Screen-space radius of every Gaussian in the last forward(); 0 = not visible.
    const std::vector<int> &radii() const { return m_radius; }

private:
    struct Device;
    std::unique_ptr<Device> m_dev;
    std::vector<int> m_radius;
};

} // namespace u3d
