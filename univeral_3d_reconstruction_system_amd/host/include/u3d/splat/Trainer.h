// Drop-in declaration of the reference's trainer (engine/include/u3d/splat/Trainer.h): same
// options, statistics and methods. All training state is resident on the GPU; model() copies
// the parameters back lazily.
#pragma once

#include "u3d/splat/Adam.h"
#include "u3d/splat/GaussianModel.h"
#include "u3d/splat/Rasterizer.h"

#include <cstdint>
#include <memory>
#include <random>
#include <vector>

namespace u3d {

struct TrainImage {
    RenderCamera camera;
    std::vector<std::uint8_t> rgb; // height * width * 3

    static std::vector<std::uint8_t> quantize(const std::vector<float> &rgb01);
};

struct TrainOptions {
    int iterations = 3000;
    int maxGaussians = 200000;
    float ssimLambda = 0.2f;
    LearningRates lr;               // lr.means is relative to the scene extent
    float meansLrFinal = 1.6e-6f;   // exponential decay target of lr.means
    float densifyGradThreshold = 2e-4f;
    float percentDense = 0.01f;     // clone (small) vs split (large) threshold x scene extent
    int refineStart = 500;
    int refineStop = 15000;
    int refineInterval = 100;
    int opacityResetInterval = 3000; // 0 = never
    float pruneOpacity = 0.005f;
    unsigned seed = 42;
    RasterSettings raster;

    static TrainOptions forIterations(int iterations);
};

struct TrainStats {
    int iteration = 0;
    float loss = 0;  // exponential moving average
    float psnr = 0;  // exponential moving average over training views
    std::size_t gaussians = 0;
};

class Trainer
{
public:
    Trainer(GaussianModel initial, std::vector<TrainImage> images, float sceneExtent,
            const TrainOptions &options, ThreadPool &pool);
    ~Trainer();
    Trainer(const Trainer &) = delete;
    Trainer &operator=(const Trainer &) = delete;

    void step();
    bool done() const;
    int iteration() const;
    const GaussianModel &model() const;
    const TrainStats &stats() const;
    const std::vector<TrainImage> &images() const { return m_images; }

    float evaluatePsnr();
    void render(const RenderCamera &camera, std::vector<float> &rgb);

private:
    struct Device;
    std::unique_ptr<Device> m_dev;
    std::vector<TrainImage> m_images;
    mutable GaussianModel m_model;
    mutable int m_modelIteration = -1;
    mutable TrainStats m_stats;
};

} // namespace u3d
