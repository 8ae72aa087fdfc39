// Rasterizer of the drop-in API over libu3d_hip: host vectors in, GPU compositing, host out.
#include "u3d/splat/Rasterizer.h"

#include "device.h"

#include <cstring>

namespace u3d {

RenderCamera RenderCamera::fromCamera(const Camera &c)
{
    RenderCamera r;
    r.width = c.width;
    r.height = c.height;
    r.fx = float(c.fx);
    r.fy = float(c.fy);
    r.cx = float(c.cx);
    r.cy = float(c.cy);
    for (int k = 0; k < 9; ++k)
        r.R[k] = float(c.R[std::size_t(k)]);
    for (int k = 0; k < 3; ++k)
        r.t[k] = float(c.t[std::size_t(k)]);
    return r;
}

namespace detail {
u3d_camera toAbi(const RenderCamera &c)
{
    u3d_camera o;
    o.width = c.width;
    o.height = c.height;
    o.fx = c.fx;
    o.fy = c.fy;
    o.cx = c.cx;
    o.cy = c.cy;
    std::memcpy(o.R, c.R, sizeof o.R);
    std::memcpy(o.t, c.t, sizeof o.t);
    return o;
}

u3d_raster_settings toAbi(const RasterSettings &s)
{
    u3d_raster_settings o;
    o.tile_size = s.tileSize;
    o.near_plane = s.nearPlane;
    o.eps2d = s.eps2d;
    std::memcpy(o.background, s.background, sizeof o.background);
    o.min_alpha = s.minAlpha;
    o.min_transmittance = s.minTransmittance;
    return o;
}
} // namespace detail

struct Rasterizer::Device {
    u3d_raster *raster = nullptr;
    dev::Model model, grads;
    dev::Buffer rgb, dl, ndc;
    std::size_t pixels = 0;
};

Rasterizer::Rasterizer(ThreadPool &) : m_dev(new Device())
{
    dev::check(u3d_raster_create(dev::context(), &m_dev->raster), "u3d_raster_create");
}

Rasterizer::~Rasterizer()
{
    if (m_dev && m_dev->raster)
        u3d_raster_destroy(m_dev->raster);
}

void Rasterizer::forward(const GaussianModel &model, const RenderCamera &camera,
                         const RasterSettings &settings, std::vector<float> &rgb)
{
    u3d_ctx *ctx = dev::context();
    Device &d = *m_dev;
    d.model.upload(model);
    d.pixels = std::size_t(camera.width) * std::size_t(camera.height);
    rgb.resize(d.pixels * 3);
    float *out = static_cast<float *>(d.rgb.ensure(d.pixels * 3 * sizeof(float) + 4));
    const u3d_gaussians m = d.model.view();
    const u3d_camera cam = detail::toAbi(camera);
    const u3d_raster_settings st = detail::toAbi(settings);
    dev::check(u3d_raster_forward(d.raster, &m, &cam, &st, out), "u3d_raster_forward");
    dev::check(u3d_memcpy_d2h(ctx, rgb.data(), out, rgb.size() * sizeof(float)), "download image");
    const int *radii = nullptr;
    int64_t n = 0;
    dev::check(u3d_raster_radii(d.raster, &radii, &n), "u3d_raster_radii");
    m_radius.resize(std::size_t(n));
    dev::check(u3d_memcpy_d2h(ctx, m_radius.data(), radii, std::size_t(n) * sizeof(int)), "download radii");
}

void Rasterizer::backward(const GaussianModel &model, const float *dLdRgb, GaussianModel &grads,
                          float *mean2dGradNdc)
{
    u3d_ctx *ctx = dev::context();
    Device &d = *m_dev;
    const std::size_t n = model.size();
    float *dl = static_cast<float *>(d.dl.ensure(d.pixels * 3 * sizeof(float) + 4));
    dev::check(u3d_memcpy_h2d(ctx, dl, dLdRgb, d.pixels * 3 * sizeof(float)), "upload dL/drgb");
    d.grads.zero(n);
    float *ndc = nullptr;
    if (mean2dGradNdc) {
        ndc = static_cast<float *>(d.ndc.ensure(n * sizeof(float) + 4));
        dev::check(u3d_memset(ctx, ndc, 0, n * sizeof(float)), "memset");
    }
    const u3d_gaussians m = d.model.view();
    u3d_gaussians g = d.grads.view();
    dev::check(u3d_raster_backward(d.raster, &m, dl, &g, ndc), "u3d_raster_backward");
    d.grads.downloadAdd(grads);
    if (mean2dGradNdc) {
        std::vector<float> tmp(n);
        dev::check(u3d_memcpy_d2h(ctx, tmp.data(), ndc, n * sizeof(float)), "download ndc");
        for (std::size_t i = 0; i < n; ++i)
            mean2dGradNdc[i] += tmp[i];
    }
}

} // namespace u3d
