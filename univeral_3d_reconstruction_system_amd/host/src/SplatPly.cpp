// Model and point-cloud files of the drop-in API (SplatPly.cpp:261-386 semantics) over the C ABI
// readers and writers; a point-cloud file given to readSplatPly becomes Gaussians through the
// device kNN initialisation.
#include "u3d/splat/SplatPly.h"

#include "u3d_hip.h"

#include <cmath>

namespace u3d {

namespace {
bool failed(std::string *error)
{
    if (error)
        *error = u3d_last_error();
    return false;
}
} // namespace

bool writeSplatPly(const std::string &path, const GaussianModel &m, std::string *error)
{
    return u3d_write_splat_ply(path.c_str(), m.means.data(), m.scales.data(), m.quats.data(),
                               m.opacities.data(), m.colors.data(), int64_t(m.size()), nullptr, 0) == U3D_OK
           || failed(error);
}

bool readSplatPly(const std::string &path, GaussianModel &model, std::string *error, bool *isSplat)
{
    int64_t vertices = 0;
    int splat = 0, degree = 0;
    if (u3d_ply_info(path.c_str(), &vertices, &splat, &degree) != U3D_OK)
        return failed(error);
    if (isSplat)
        *isSplat = splat != 0;
    if (!splat) {
        // a point cloud: kNN-initialised Gaussians, nearly opaque and half the kNN scale
        // (SplatPly.cpp:297-316)
        std::vector<float> xyz, rgb;
        if (!readPointsPly(path, xyz, rgb, error))
            return false;
        model = initGaussiansFromPoints(xyz, rgb);
        const float opaque = logit(0.95f), shrink = std::log(0.5f);
        for (float &o : model.opacities)
            o = opaque;
        for (float &s : model.scales)
            s += shrink;
        return true;
    }
    GaussianModel m;
    m.resize(std::size_t(vertices));
    std::vector<float> rest(degree > 0 ? std::size_t(vertices) * 45 : 0);
    int64_t n = 0;
    if (u3d_read_splat_ply(path.c_str(), vertices > 0 ? vertices : 1, m.means.data(), m.scales.data(),
                           m.quats.data(), m.opacities.data(), m.colors.data(),
                           degree > 0 ? rest.data() : nullptr, &n, nullptr) != U3D_OK)
        return failed(error);
    m.resize(std::size_t(n)); // rows with a non-finite position were dropped
    model = std::move(m);
    return true;
}

bool writePointsPly(const std::string &path, const std::vector<float> &xyz, const std::vector<float> &rgb,
                    std::string *error)
{
    return u3d_write_points_ply(path.c_str(), xyz.data(), rgb.data(), int64_t(rgb.size()),
                                int64_t(xyz.size() / 3)) == U3D_OK
           || failed(error);
}

bool readPointsPly(const std::string &path, std::vector<float> &xyz, std::vector<float> &rgb, std::string *error)
{
    int64_t n = 0;
    if (u3d_read_points_ply(path.c_str(), 0, nullptr, nullptr, &n) != U3D_OK)
        return failed(error);
    xyz.assign(std::size_t(n) * 3, 0.0f);
    rgb.assign(std::size_t(n) * 3, 0.0f);
    if (n > 0 && u3d_read_points_ply(path.c_str(), n, xyz.data(), rgb.data(), &n) != U3D_OK)
        return failed(error);
    return true;
}

} // namespace u3d
