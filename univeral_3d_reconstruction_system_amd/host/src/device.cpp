// Shared device context and SoA transfer helpers of the drop-in API.
#include "device.h"

#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <vector>

namespace u3d {
namespace dev {

void fail(const char *what, int rc)
{
    std::fprintf(stderr, "u3d (MI355X trainer): %s failed (%d): %s\n", what, rc, u3d_last_error());
    std::abort();
}

u3d_ctx *context()
{
    static std::once_flag once;
    static u3d_ctx *ctx = nullptr;
    std::call_once(once, [] {
        const char *env = std::getenv("U3D_DEVICE");
        const int device = env ? std::atoi(env) : 0;
        check(u3d_ctx_create(device, &ctx), "u3d_ctx_create");
    });
    return ctx;
}

Model::~Model()
{
    if (m_buf)
        u3d_free(context(), m_buf);
}

void Model::reserve(std::size_t n)
{
    if (n <= m_cap && m_buf)
        return;
    const std::size_t cap = n + n / 4 + 16;
    void *p = nullptr;
    check(u3d_malloc(context(), cap * 14 * sizeof(float), &p), "u3d_malloc");
    if (m_buf)
        u3d_free(context(), m_buf);
    m_buf = static_cast<float *>(p);
    m_cap = cap;
}

u3d_gaussians Model::view() const
{
    float *b = m_buf;
    const std::size_t c = m_cap;
    return {b, b + 3 * c, b + 6 * c, b + 10 * c, b + 11 * c, int64_t(m_n)};
}

void Model::upload(const GaussianModel &m)
{
    reserve(m.size());
    m_n = m.size();
    const u3d_gaussians v = view();
    const std::size_t f = sizeof(float);
    check(u3d_memcpy_h2d(context(), v.means, m.means.data(), m_n * 3 * f), "upload");
    check(u3d_memcpy_h2d(context(), v.scales, m.scales.data(), m_n * 3 * f), "upload");
    check(u3d_memcpy_h2d(context(), v.quats, m.quats.data(), m_n * 4 * f), "upload");
    check(u3d_memcpy_h2d(context(), v.opacities, m.opacities.data(), m_n * f), "upload");
    check(u3d_memcpy_h2d(context(), v.colors, m.colors.data(), m_n * 3 * f), "upload");
}

void Model::zero(std::size_t n)
{
    reserve(n);
    m_n = n;
    check(u3d_memset(context(), m_buf, 0, m_cap * 14 * sizeof(float)), "memset");
}

void Model::download(GaussianModel &m) const
{
    m.resize(m_n);
    const u3d_gaussians v = view();
    const std::size_t f = sizeof(float);
    check(u3d_memcpy_d2h(context(), m.means.data(), v.means, m_n * 3 * f), "download");
    check(u3d_memcpy_d2h(context(), m.scales.data(), v.scales, m_n * 3 * f), "download");
    check(u3d_memcpy_d2h(context(), m.quats.data(), v.quats, m_n * 4 * f), "download");
    check(u3d_memcpy_d2h(context(), m.opacities.data(), v.opacities, m_n * f), "download");
    check(u3d_memcpy_d2h(context(), m.colors.data(), v.colors, m_n * 3 * f), "download");
}

void Model::downloadAdd(GaussianModel &m) const
{
    GaussianModel tmp;
    download(tmp);
    auto add = [](std::vector<float> &a, const std::vector<float> &b) {
        for (std::size_t k = 0; k < a.size() && k < b.size(); ++k)
            a[k] += b[k];
    };
    add(m.means, tmp.means);
    add(m.scales, tmp.scales);
    add(m.quats, tmp.quats);
    add(m.opacities, tmp.opacities);
    add(m.colors, tmp.colors);
}

Buffer::~Buffer()
{
    if (m_p)
        u3d_free(context(), m_p);
}

void *Buffer::ensure(std::size_t bytes)
{
    if (bytes <= m_cap && m_p)
        return m_p;
    if (m_p)
        u3d_free(context(), m_p);
    m_p = nullptr;
    const std::size_t cap = bytes + bytes / 4 + 64;
    check(u3d_malloc(context(), cap, &m_p), "u3d_malloc");
    m_cap = cap;
    return m_p;
}

} // namespace dev
} // namespace u3d
