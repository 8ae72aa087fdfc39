// Drop-in declaration of the reference's Gaussian parameter set
// (engine/include/u3d/splat/GaussianModel.h): same fields, layout and helpers.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace u3d {

constexpr float kShC0 = 0.28209479177387814f; // degree-0 spherical harmonic constant

inline float sigmoid(float x) { return 1.0f / (1.0f + std::exp(-x)); }
inline float logit(float p) { return std::log(p / (1.0f - p)); }

// Structure of arrays: means xyz, scales (log std-dev) xyz, quats wxyz (not necessarily
// normalised), opacities (logits), colors (SH degree-0 coefficients, rgb = 0.5 + C0 * c).
// Also used for gradients and optimiser moments.
struct GaussianModel {
    std::vector<float> means;
    std::vector<float> scales;
    std::vector<float> quats;
    std::vector<float> opacities;
    std::vector<float> colors;

    std::size_t size() const { return opacities.size(); }
    void resize(std::size_t n);
    void setZero();
    void compact(const std::vector<std::uint8_t> &keep); // keeps keep[i] != 0, in order
    void appendFrom(const GaussianModel &other, std::size_t index);
};

// Isotropic Gaussians from a coloured point cloud: scale from the mean squared distance to the
// three nearest neighbours on a hash grid, identity rotation, opacity 0.1.
GaussianModel initGaussiansFromPoints(const std::vector<float> &xyz, const std::vector<float> &rgb);

} // namespace u3d
