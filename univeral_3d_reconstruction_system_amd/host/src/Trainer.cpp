// Trainer of the drop-in API: a handle on the device-resident trainer of libu3d_hip.
#include "u3d/splat/Trainer.h"

#include "device.h"

#include <algorithm>
#include <cmath>

namespace u3d {

namespace detail {
u3d_camera toAbi(const RenderCamera &c);
u3d_raster_settings toAbi(const RasterSettings &s);
} // namespace detail

std::vector<std::uint8_t> TrainImage::quantize(const std::vector<float> &rgb01)
{
    std::vector<std::uint8_t> out(rgb01.size());
    for (std::size_t k = 0; k < rgb01.size(); ++k)
        out[k] = std::uint8_t(std::lround(std::min(1.0f, std::max(0.0f, rgb01[k])) * 255.0f));
    return out;
}

TrainOptions TrainOptions::forIterations(int iterations)
{
    // 3DGS schedule compressed to `iterations` (Trainer.cpp:26-35).
    TrainOptions o;
    o.iterations = std::max(1, iterations);
    o.refineStart = std::min(500, o.iterations / 10);
    o.refineStop = o.iterations * 6 / 10;
    o.refineInterval = 100;
    o.opacityResetInterval = o.iterations >= 6000 ? 3000 : 0;
    return o;
}

struct Trainer::Device {
    u3d_trainer *trainer = nullptr;
};

Trainer::Trainer(GaussianModel initial, std::vector<TrainImage> images, float sceneExtent,
                 const TrainOptions &options, ThreadPool &)
    : m_dev(new Device()), m_images(std::move(images))
{
    std::vector<u3d_camera> cams;
    std::vector<std::uint8_t> rgb;
    for (const TrainImage &img : m_images) {
        cams.push_back(detail::toAbi(img.camera));
        rgb.insert(rgb.end(), img.rgb.begin(), img.rgb.end());
    }
    u3d_train_options o{};
    o.iterations = options.iterations;
    o.max_gaussians = options.maxGaussians;
    o.ssim_lambda = options.ssimLambda;
    o.lr = {options.lr.means, options.lr.scales, options.lr.quats, options.lr.opacities, options.lr.colors};
    o.means_lr_final = options.meansLrFinal;
    o.densify_grad_threshold = options.densifyGradThreshold;
    o.percent_dense = options.percentDense;
    o.refine_start = options.refineStart;
    o.refine_stop = options.refineStop;
    o.refine_interval = options.refineInterval;
    o.opacity_reset_interval = options.opacityResetInterval;
    o.prune_opacity = options.pruneOpacity;
    o.seed = options.seed;
    o.raster = detail::toAbi(options.raster);
    dev::check(u3d_trainer_create(dev::context(), initial.means.data(), initial.scales.data(),
                                  initial.quats.data(), initial.opacities.data(), initial.colors.data(),
                                  int64_t(initial.size()), cams.data(), rgb.data(), int(cams.size()),
                                  sceneExtent, &o, &m_dev->trainer),
               "u3d_trainer_create");
    m_model = std::move(initial);
    m_modelIteration = 0;
    m_stats.gaussians = m_model.size();
}

Trainer::~Trainer()
{
    if (m_dev && m_dev->trainer)
        u3d_trainer_destroy(m_dev->trainer);
}

void Trainer::step() { dev::check(u3d_trainer_step(m_dev->trainer), "u3d_trainer_step"); }
bool Trainer::done() const { return u3d_trainer_done(m_dev->trainer) != 0; }
int Trainer::iteration() const { return u3d_trainer_iteration(m_dev->trainer); }

const GaussianModel &Trainer::model() const
{
    const int it = iteration();
    if (m_modelIteration != it) {
        m_model.resize(std::size_t(u3d_trainer_size(m_dev->trainer)));
        dev::check(u3d_trainer_get_model(m_dev->trainer, m_model.means.data(), m_model.scales.data(),
                                         m_model.quats.data(), m_model.opacities.data(), m_model.colors.data()),
                   "u3d_trainer_get_model");
        m_modelIteration = it;
    }
    return m_model;
}

const TrainStats &Trainer::stats() const
{
    u3d_train_stats s{};
    dev::check(u3d_trainer_stats(m_dev->trainer, &s), "u3d_trainer_stats");
    m_stats.iteration = s.iteration;
    m_stats.loss = s.loss;
    m_stats.psnr = s.psnr;
    m_stats.gaussians = std::size_t(s.gaussians);
    return m_stats;
}

float Trainer::evaluatePsnr()
{
    float v = 0;
    dev::check(u3d_trainer_evaluate_psnr(m_dev->trainer, &v), "u3d_trainer_evaluate_psnr");
    return v;
}

void Trainer::render(const RenderCamera &camera, std::vector<float> &rgb)
{
    rgb.resize(std::size_t(camera.width) * std::size_t(camera.height) * 3);
    const u3d_camera c = detail::toAbi(camera);
    dev::check(u3d_trainer_render(m_dev->trainer, &c, rgb.data()), "u3d_trainer_render");
}

} // namespace u3d
