// Parameter container of the drop-in API (GaussianModel.cpp:9-66) and its kNN-distance
// initialisation (GaussianModel.cpp:68-144), which runs on the device through the C ABI.
#include "u3d/splat/GaussianModel.h"

#include "device.h"

#include <algorithm>

namespace u3d {

void GaussianModel::resize(std::size_t n)
{
    means.resize(n * 3);
    scales.resize(n * 3);
    quats.resize(n * 4);
    opacities.resize(n);
    colors.resize(n * 3);
}

void GaussianModel::setZero()
{
    for (std::vector<float> *v : {&means, &scales, &quats, &opacities, &colors})
        std::fill(v->begin(), v->end(), 0.0f);
}

namespace {
void keepRows(std::vector<float> &v, const std::vector<std::uint8_t> &keep, std::size_t width)
{
    std::size_t dst = 0;
    for (std::size_t i = 0; i < keep.size(); ++i) {
        if (!keep[i])
            continue;
        if (dst != i)
            std::copy_n(v.begin() + std::ptrdiff_t(i * width), width, v.begin() + std::ptrdiff_t(dst * width));
        ++dst;
    }
    v.resize(dst * width);
}

void appendRow(std::vector<float> &dst, const std::vector<float> &src, std::size_t i, std::size_t width)
{
    dst.insert(dst.end(), src.begin() + std::ptrdiff_t(i * width), src.begin() + std::ptrdiff_t((i + 1) * width));
}
} // namespace

void GaussianModel::compact(const std::vector<std::uint8_t> &keep)
{
    keepRows(means, keep, 3);
    keepRows(scales, keep, 3);
    keepRows(quats, keep, 4);
    keepRows(opacities, keep, 1);
    keepRows(colors, keep, 3);
}

void GaussianModel::appendFrom(const GaussianModel &other, std::size_t index)
{
    appendRow(means, other.means, index, 3);
    appendRow(scales, other.scales, index, 3);
    appendRow(quats, other.quats, index, 4);
    appendRow(opacities, other.opacities, index, 1);
    appendRow(colors, other.colors, index, 3);
}

GaussianModel initGaussiansFromPoints(const std::vector<float> &xyz, const std::vector<float> &rgb)
{
    // The kNN-distance initialisation runs on the device (u3d_points_knn_init): the points go up
    // once, the hash-grid 3-NN search and the parameter fill run there, the model comes back.
    const std::size_t n = xyz.size() / 3;
    GaussianModel m;
    m.resize(n);
    if (n == 0)
        return m;
    u3d_ctx *ctx = dev::context();
    // rgb entries beyond 3n are never read (the reference reads rgb[i*3+k] for i < n only)
    const std::size_t rgbCount = std::min(rgb.size(), n * 3);
    dev::Buffer points, colours;
    float *dXyz = static_cast<float *>(points.ensure(n * 3 * sizeof(float)));
    float *dRgb = rgbCount ? static_cast<float *>(colours.ensure(rgbCount * sizeof(float))) : nullptr;
    dev::check(u3d_memcpy_h2d(ctx, dXyz, xyz.data(), n * 3 * sizeof(float)), "upload points");
    if (rgbCount)
        dev::check(u3d_memcpy_h2d(ctx, dRgb, rgb.data(), rgbCount * sizeof(float)), "upload colours");
    dev::Model out;
    out.zero(n);
    u3d_gaussians ov = out.view();
    dev::check(u3d_points_knn_init(ctx, dXyz, dRgb, int(rgbCount), int64_t(n), &ov), "u3d_points_knn_init");
    out.download(m);
    return m;
}

} // namespace u3d
