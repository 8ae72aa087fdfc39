// Drop-in declaration of the reference's optimiser (engine/include/u3d/splat/Adam.h). Moments
// live on the GPU.
#pragma once

#include "u3d/ThreadPool.h"
#include "u3d/splat/GaussianModel.h"

#include <memory>

namespace u3d {

struct LearningRates {
    float means = 1.6e-4f;
    float scales = 5e-3f;
    float quats = 1e-3f;
    float opacities = 5e-2f;
    float colors = 2.5e-3f;
};

// Adam (beta1 0.9, beta2 0.999, eps 1e-15) with one rate per parameter group; the moments
// follow densification (compact / grow).
class Adam
{
public:
    Adam();
    ~Adam();
    Adam(const Adam &) = delete;
    Adam &operator=(const Adam &) = delete;
    void reset(std::size_t n);
    void compact(const std::vector<std::uint8_t> &keep);
    void grow(std::size_t count); // appends zero moments
    void step(GaussianModel &params, const GaussianModel &grads, const LearningRates &lr, ThreadPool &pool);
    void resetOpacityMoments();

private:
    struct Device;
    std::unique_ptr<Device> m_dev;
};

} // namespace u3d
