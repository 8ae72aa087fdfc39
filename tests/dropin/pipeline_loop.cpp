// The pipeline's training loop (engine/src/pipeline/Reconstructor.cpp:325-361) over the drop-in
// library: initGaussiansFromPoints, Trainer::step with stats() after every step, a progress
// report every 10 iterations and at the end, cancellation checked before each step, then
// writeSplatPly(trainer.model()) and readSplatPly of the file. Prints one JSON line.
//   pipeline_loop ITERATIONS CANCEL_AT OUT.ply     (CANCEL_AT 0: never)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "u3d/splat/GaussianModel.h"
#include "u3d/splat/Rasterizer.h"
#include "u3d/splat/SplatPly.h"
#include "u3d/splat/Trainer.h"
#include "u3d/ThreadPool.h"

using namespace u3d;

namespace {

// A camera on a ring of radius r at height y, looking at the origin (x right, y down, z forward).
RenderCamera ringCamera(int w, int h, float angle, float r, float y)
{
    RenderCamera c;
    c.width = w;
    c.height = h;
    c.fx = c.fy = 0.9f * float(w);
    c.cx = 0.5f * float(w);
    c.cy = 0.5f * float(h);
    const float eye[3] = {r * std::cos(angle), y, r * std::sin(angle)};
    float z[3] = {-eye[0], -eye[1], -eye[2]};
    const float zn = std::sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
    for (float &v : z)
        v /= zn;
    const float up[3] = {0.0f, -1.0f, 0.0f}; // image y runs down
    float x[3] = {up[1] * z[2] - up[2] * z[1], up[2] * z[0] - up[0] * z[2], up[0] * z[1] - up[1] * z[0]};
    const float xn = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    for (float &v : x)
        v /= xn;
    const float yv[3] = {z[1] * x[2] - z[2] * x[1], z[2] * x[0] - z[0] * x[2], z[0] * x[1] - z[1] * x[0]};
    for (int k = 0; k < 3; ++k) {
        c.R[k] = x[k];
        c.R[3 + k] = yv[k];
        c.R[6 + k] = z[k];
    }
    for (int row = 0; row < 3; ++row)
        c.t[row] = -(c.R[row * 3] * eye[0] + c.R[row * 3 + 1] * eye[1] + c.R[row * 3 + 2] * eye[2]);
    return c;
}

GaussianModel truthModel(std::mt19937 &rng, std::size_t n)
{
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    std::normal_distribution<float> g(0.0f, 1.0f);
    GaussianModel m;
    m.resize(n);
    for (std::size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            m.means[i * 3 + k] = 0.8f * u(rng);
            m.scales[i * 3 + k] = std::log(0.04f + 0.06f * (0.5f + 0.5f * u(rng)));
            m.colors[i * 3 + k] = 0.6f * u(rng);
        }
        float q[4], qn = 0.0f;
        for (float &v : q) {
            v = g(rng);
            qn += v * v;
        }
        for (int k = 0; k < 4; ++k)
            m.quats[i * 4 + k] = q[k] / std::sqrt(qn);
        m.opacities[i] = logit(0.8f);
    }
    return m;
}

} // namespace

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: pipeline_loop ITERATIONS CANCEL_AT OUT.ply\n");
        return 2;
    }
    const int iterations = std::atoi(argv[1]);
    const int cancelAt = std::atoi(argv[2]);
    const std::string path = argv[3];
    ThreadPool pool;
    std::mt19937 rng(5);
    const GaussianModel truth = truthModel(rng, 400);
    Rasterizer raster(pool);
    std::vector<TrainImage> images;
    for (int k = 0; k < 12; ++k) {
        TrainImage img;
        img.camera = ringCamera(128, 96, 6.2831853f * float(k) / 12.0f, 4.0f, -0.8f + 0.2f * float(k % 3));
        std::vector<float> rendered;
        raster.forward(truth, img.camera, RasterSettings(), rendered);
        img.rgb = TrainImage::quantize(rendered);
        images.push_back(std::move(img));
    }
    std::normal_distribution<float> noise(0.0f, 0.05f);
    std::vector<float> xyz, rgb;
    for (std::size_t i = 0; i < truth.size(); i += 2)
        for (int k = 0; k < 3; ++k) {
            xyz.push_back(truth.means[i * 3 + k] + noise(rng));
            rgb.push_back(0.5f);
        }
    // -- the loop of Reconstructor.cpp:326-348
    GaussianModel init = initGaussiansFromPoints(xyz, rgb);
    const std::size_t initSize = init.size();
    TrainOptions topt = TrainOptions::forIterations(iterations);
    topt.maxGaussians = 4000;
    Trainer trainer(std::move(init), std::move(images), 4.8f, topt, pool);
    const float before = trainer.evaluatePsnr();
    bool cancel = false, cancelled = false, consistent = true;
    int reports = 0, lastReport = 0, logs = 0;
    while (!trainer.done()) {
        if (cancel) {
            cancelled = true;
            break;
        }
        trainer.step();
        const TrainStats &st = trainer.stats();
        consistent = consistent && st.iteration == trainer.iteration() && std::isfinite(st.loss)
                     && std::isfinite(st.psnr) && st.gaussians == trainer.model().size();
        if (st.iteration % 10 == 0 || trainer.done()) {
            ++reports;
            lastReport = st.iteration;
            if (cancelAt > 0 && st.iteration >= cancelAt)
                cancel = true; // the caller's cancel flag, seen before the next step
        }
        if (st.iteration % 250 == 0)
            ++logs;
    }
    // -- export (Reconstructor.cpp:351-361) and read back
    std::string error;
    const GaussianModel &model = trainer.model();
    const bool wrote = writeSplatPly(path, model, &error);
    GaussianModel back;
    bool isSplat = false;
    const bool read = wrote && readSplatPly(path, back, &error, &isSplat);
    bool same = read && back.size() == model.size();
    for (std::size_t i = 0; same && i < model.means.size(); ++i)
        same = back.means[i] == model.means[i];
    const TrainStats &st = trainer.stats();
    std::printf("{\"iterations\": %d, \"cancelled\": %s, \"reports\": %d, \"last_report\": %d, "
                "\"logs\": %d, \"consistent\": %s, \"initial\": %zu, \"gaussians\": %zu, "
                "\"loss_ema\": %.6f, \"psnr_ema\": %.4f, \"psnr_before\": %.4f, \"psnr_after\": %.4f, "
                "\"file_rows\": %zu, \"is_splat\": %s, \"round_trip\": %s, \"error\": \"%s\"}\n",
                trainer.iteration(), cancelled ? "true" : "false", reports, lastReport, logs,
                consistent ? "true" : "false", initSize, model.size(), st.loss, st.psnr, before,
                trainer.evaluatePsnr(), back.size(), isSplat ? "true" : "false", same ? "true" : "false",
                error.c_str());
    return 0;
}
