// Adam of the drop-in API: moments on the GPU; parameters and gradients cross per step.
#include "u3d/splat/Adam.h"

#include "device.h"

namespace u3d {

struct Adam::Device {
    u3d_adam *adam = nullptr;
    dev::Model params, grads;
    dev::Buffer keep;
};

Adam::Adam() : m_dev(new Device())
{
    dev::check(u3d_adam_create(dev::context(), &m_dev->adam), "u3d_adam_create");
}

Adam::~Adam()
{
    if (m_dev && m_dev->adam)
        u3d_adam_destroy(m_dev->adam);
}

void Adam::reset(std::size_t n) { dev::check(u3d_adam_reset(m_dev->adam, int64_t(n)), "u3d_adam_reset"); }

void Adam::compact(const std::vector<std::uint8_t> &keep)
{
    void *k = m_dev->keep.ensure(keep.size() + 1);
    dev::check(u3d_memcpy_h2d(dev::context(), k, keep.data(), keep.size()), "upload keep");
    dev::check(u3d_adam_compact(m_dev->adam, static_cast<const uint8_t *>(k)), "u3d_adam_compact");
}

void Adam::grow(std::size_t count) { dev::check(u3d_adam_grow(m_dev->adam, int64_t(count)), "u3d_adam_grow"); }

void Adam::resetOpacityMoments()
{
    dev::check(u3d_adam_reset_opacity_moments(m_dev->adam), "u3d_adam_reset_opacity_moments");
}

void Adam::step(GaussianModel &params, const GaussianModel &grads, const LearningRates &lr, ThreadPool &)
{
    Device &d = *m_dev;
    d.params.upload(params);
    d.grads.upload(grads);
    u3d_gaussians p = d.params.view();
    const u3d_gaussians g = d.grads.view();
    const u3d_learning_rates r{lr.means, lr.scales, lr.quats, lr.opacities, lr.colors};
    dev::check(u3d_adam_step(d.adam, &p, &g, &r), "u3d_adam_step");
    d.params.download(params);
}

} // namespace u3d
