// Parameter container and kNN-distance initialisation of the drop-in API
// (engine/src/splat/GaussianModel.cpp semantics: GaussianModel.cpp:9-66, 68-144).
#include "u3d/splat/GaussianModel.h"

#include <algorithm>
#include <limits>
#include <unordered_map>

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
    const std::size_t n = xyz.size() / 3;
    GaussianModel m;
    m.resize(n);
    if (n == 0)
        return m;
    const float big = std::numeric_limits<float>::max();
    float lo[3] = {big, big, big}, hi[3] = {-big, -big, -big};
    for (std::size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], xyz[i * 3 + k]);
            hi[k] = std::max(hi[k], xyz[i * 3 + k]);
        }
    // Grid cell of twice the mean spacing; cells keyed by an XOR of prime multiples, so
    // colliding cells share a bucket exactly as in the reference.
    const float span = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-6f});
    const float cell = std::max(span / std::cbrt(float(n)) * 2.0f, 1e-6f);
    auto cellKey = [](int x, int y, int z) {
        return (std::int64_t(x) * 73856093) ^ (std::int64_t(y) * 19349663) ^ (std::int64_t(z) * 83492791);
    };
    auto cellOf = [&](std::size_t i, int c[3]) {
        for (int k = 0; k < 3; ++k)
            c[k] = int(std::floor((xyz[i * 3 + k] - lo[k]) / cell));
    };
    std::unordered_map<std::int64_t, std::vector<std::uint32_t>> buckets;
    buckets.reserve(n);
    for (std::size_t i = 0; i < n; ++i) {
        int c[3];
        cellOf(i, c);
        buckets[cellKey(c[0], c[1], c[2])].push_back(std::uint32_t(i));
    }
    const float logitTenth = logit(0.1f);
    for (std::size_t i = 0; i < n; ++i) {
        float best[3] = {big, big, big}; // three smallest squared distances, ascending
        int c[3];
        cellOf(i, c);
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    auto it = buckets.find(cellKey(c[0] + dx, c[1] + dy, c[2] + dz));
                    if (it == buckets.end())
                        continue;
                    for (std::uint32_t j : it->second) {
                        if (j == i)
                            continue;
                        const float ex = xyz[j * 3] - xyz[i * 3], ey = xyz[j * 3 + 1] - xyz[i * 3 + 1],
                                    ez = xyz[j * 3 + 2] - xyz[i * 3 + 2];
                        float d2 = ex * ex + ey * ey + ez * ez;
                        for (int k = 0; k < 3; ++k)
                            if (d2 < best[k])
                                std::swap(d2, best[k]);
                    }
                }
        float sum = 0;
        int found = 0;
        for (float b : best)
            if (b < big) {
                sum += b;
                ++found;
            }
        const float meanD2 = found ? sum / float(found) : cell * cell;
        const float logScale = std::log(std::sqrt(std::max(meanD2, 1e-14f)));
        for (int k = 0; k < 3; ++k) {
            m.means[i * 3 + k] = xyz[i * 3 + k];
            m.scales[i * 3 + k] = logScale;
            const float col = rgb.size() >= (i + 1) * 3 ? rgb[i * 3 + k] : 0.5f;
            m.colors[i * 3 + k] = (col - 0.5f) / kShC0;
        }
        m.quats[i * 4] = 1.0f;
        m.quats[i * 4 + 1] = m.quats[i * 4 + 2] = m.quats[i * 4 + 3] = 0.0f;
        m.opacities[i] = logitTenth;
    }
    return m;
}

} // namespace u3d
