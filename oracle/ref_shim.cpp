// C entry points over the compiled reference trainer (engine/src/splat), used only by the test
// suite and the CPU-baseline leg of the benchmark. Built by oracle/Makefile from the reference
// sources where they lie; nothing from the reference is copied into this repository.
//
// `private` is widened to `public` for the reference headers so the tests can read the
// rasterizer's per-forward state (tile lists, final transmittance, contributor counts:
// Rasterizer.h:60-89) and drive Trainer::densifyAndPrune with injected statistics
// (Trainer.h:266-291). No reference behaviour is changed.
#define private public
#include "u3d/ThreadPool.h"
#include "u3d/splat/Adam.h"
#include "u3d/splat/GaussianModel.h"
#include "u3d/splat/Loss.h"
#include "u3d/splat/Rasterizer.h"
#include "u3d/splat/SplatPly.h"
#include "u3d/splat/Trainer.h"
#undef private

#include "splat_oracle.h"

#include <algorithm>
#include <cstring>
#include <memory>

using namespace u3d;

namespace {

GaussianModel toModel(const orc_model *m)
{
    GaussianModel g;
    const size_t n = size_t(m->n);
    g.means.assign(m->means, m->means + n * 3);
    g.scales.assign(m->scales, m->scales + n * 3);
    g.quats.assign(m->quats, m->quats + n * 4);
    g.opacities.assign(m->opacities, m->opacities + n);
    g.colors.assign(m->colors, m->colors + n * 3);
    return g;
}

void fromModel(const GaussianModel &g, orc_model *m)
{
    std::copy(g.means.begin(), g.means.end(), m->means);
    std::copy(g.scales.begin(), g.scales.end(), m->scales);
    std::copy(g.quats.begin(), g.quats.end(), m->quats);
    std::copy(g.opacities.begin(), g.opacities.end(), m->opacities);
    std::copy(g.colors.begin(), g.colors.end(), m->colors);
}

RenderCamera toCam(const orc_camera *c)
{
    RenderCamera r;
    r.width = c->width;
    r.height = c->height;
    r.fx = c->fx; r.fy = c->fy; r.cx = c->cx; r.cy = c->cy;
    std::memcpy(r.R, c->R, sizeof r.R);
    std::memcpy(r.t, c->t, sizeof r.t);
    return r;
}

RasterSettings toSettings(const orc_settings *s)
{
    RasterSettings r;
    r.tileSize = s->tile_size;
    r.nearPlane = s->near_plane;
    r.eps2d = s->eps2d;
    std::memcpy(r.background, s->background, sizeof r.background);
    r.minAlpha = s->min_alpha;
    r.minTransmittance = s->min_transmittance;
    return r;
}

TrainOptions toOptions(const orc_train_options *o)
{
    TrainOptions t;
    t.iterations = o->iterations;
    t.maxGaussians = o->max_gaussians;
    t.ssimLambda = o->ssim_lambda;
    t.lr.means = o->lr.means;
    t.lr.scales = o->lr.scales;
    t.lr.quats = o->lr.quats;
    t.lr.opacities = o->lr.opacities;
    t.lr.colors = o->lr.colors;
    t.meansLrFinal = o->means_lr_final;
    t.densifyGradThreshold = o->densify_grad_threshold;
    t.percentDense = o->percent_dense;
    t.refineStart = o->refine_start;
    t.refineStop = o->refine_stop;
    t.refineInterval = o->refine_interval;
    t.opacityResetInterval = o->opacity_reset_interval;
    t.pruneOpacity = o->prune_opacity;
    t.seed = o->seed;
    t.raster = toSettings(&o->raster);
    return t;
}

struct RefRaster {
    std::unique_ptr<ThreadPool> pool;
    std::unique_ptr<Rasterizer> raster;
};

struct RefTrainer {
    std::unique_ptr<ThreadPool> pool;
    std::unique_ptr<Trainer> trainer;
};

} // namespace

extern "C" {

void *ref_raster_new(int threads)
{
    auto *r = new RefRaster();
    r->pool.reset(new ThreadPool(threads));
    r->raster.reset(new Rasterizer(*r->pool));
    return r;
}
void ref_raster_free(void *h) { delete static_cast<RefRaster *>(h); }

void ref_raster_forward(void *h, const orc_model *m, const orc_camera *cam, const orc_settings *st,
                        float *rgb)
{
    auto *r = static_cast<RefRaster *>(h);
    const GaussianModel g = toModel(m);
    std::vector<float> out;
    r->raster->forward(g, toCam(cam), toSettings(st), out);
    std::copy(out.begin(), out.end(), rgb);
}

long ref_raster_num_pairs(void *h) { return long(static_cast<RefRaster *>(h)->raster->m_tileList.size()); }

void ref_raster_get_state(void *h, int *radius, int *tile_offsets, int *tile_list, float *final_t,
                          int *n_contrib)
{
    const Rasterizer &r = *static_cast<RefRaster *>(h)->raster;
    std::copy(r.m_radius.begin(), r.m_radius.end(), radius);
    std::copy(r.m_tileOffsets.begin(), r.m_tileOffsets.end(), tile_offsets);
    std::copy(r.m_tileList.begin(), r.m_tileList.end(), tile_list);
    std::copy(r.m_finalT.begin(), r.m_finalT.end(), final_t);
    std::copy(r.m_nContrib.begin(), r.m_nContrib.end(), n_contrib);
}

void ref_raster_get_projection(void *h, float *mean2d, float *conic, float *depth, float *rgb,
                               float *alpha, int *rect, uint8_t *flags)
{
    const Rasterizer &r = *static_cast<RefRaster *>(h)->raster;
    std::copy(r.m_mean2d.begin(), r.m_mean2d.end(), mean2d);
    std::copy(r.m_conic.begin(), r.m_conic.end(), conic);
    std::copy(r.m_depth.begin(), r.m_depth.end(), depth);
    std::copy(r.m_rgb.begin(), r.m_rgb.end(), rgb);
    std::copy(r.m_alpha.begin(), r.m_alpha.end(), alpha);
    std::copy(r.m_rect.begin(), r.m_rect.end(), rect);
    std::copy(r.m_flags.begin(), r.m_flags.end(), flags);
}

void ref_raster_backward(void *h, const orc_model *m, const float *dl, orc_model *grads, float *ndc)
{
    auto *r = static_cast<RefRaster *>(h);
    const GaussianModel g = toModel(m);
    GaussianModel gr = toModel(grads);
    r->raster->backward(g, dl, gr, ndc);
    fromModel(gr, grads);
}

orc_loss_value ref_photometric_loss(const float *render, const float *target, int w, int h,
                                    float lambda, float *grad, int threads)
{
    ThreadPool pool(threads);
    const LossValue v = photometricLoss(render, target, w, h, lambda, grad, pool);
    return {v.total, v.l1, v.ssim};
}

float ref_psnr(const float *a, const float *b, long count) { return psnr(a, b, size_t(count)); }

// `steps` consecutive Adam steps with the same gradients, starting from zero moments.
void ref_adam_steps(orc_model *p, const orc_model *g, const orc_lrs *lr, int steps, orc_model *m_out,
                    orc_model *v_out)
{
    ThreadPool pool(1);
    GaussianModel params = toModel(p);
    const GaussianModel grads = toModel(g);
    Adam adam;
    adam.reset(params.size());
    LearningRates rates;
    rates.means = lr->means;
    rates.scales = lr->scales;
    rates.quats = lr->quats;
    rates.opacities = lr->opacities;
    rates.colors = lr->colors;
    for (int s = 0; s < steps; ++s)
        adam.step(params, grads, rates, pool);
    fromModel(params, p);
    if (m_out)
        fromModel(adam.m_m, m_out);
    if (v_out)
        fromModel(adam.m_v, v_out);
}

// SplatPly.h: the reference's own .ply writer and reader (for the exporter's cross-check).
int ref_write_splat_ply(const char *path, const orc_model *m)
{
    return writeSplatPly(path, toModel(m)) ? 1 : 0;
}

// Returns the vertex count (-1 on error); fills `out` when it is non-null.
long ref_read_splat_ply(const char *path, orc_model *out)
{
    GaussianModel g;
    bool splat = false;
    if (!readSplatPly(path, g, nullptr, &splat) || !splat)
        return -1;
    if (out)
        fromModel(g, out);
    return long(g.size());
}

// SplatPly.h: the reference's point-cloud writer and reader. ref_read_points_ply returns the
// point count (-1 on error) and fills xyz / rgb (3 per point) when they are non-null.
int ref_write_points_ply(const char *path, const float *xyz, const float *rgb, long rgb_count, long n)
{
    std::vector<float> x(xyz, xyz + size_t(n) * 3), c;
    if (rgb)
        c.assign(rgb, rgb + rgb_count);
    return writePointsPly(path, x, c) ? 1 : 0;
}

long ref_read_points_ply(const char *path, float *xyz, float *rgb)
{
    std::vector<float> x, c;
    if (!readPointsPly(path, x, c))
        return -1;
    if (xyz)
        std::copy(x.begin(), x.end(), xyz);
    if (rgb)
        std::copy(c.begin(), c.end(), rgb);
    return long(x.size() / 3);
}

// readSplatPly of any file: the model (splat or point-cloud fallback); returns the size or -1,
// *is_splat the kind.
long ref_read_splat_ply_any(const char *path, orc_model *out, int *is_splat)
{
    GaussianModel g;
    bool splat = false;
    if (!readSplatPly(path, g, nullptr, &splat))
        return -1;
    if (is_splat)
        *is_splat = splat ? 1 : 0;
    if (out)
        fromModel(g, out);
    return long(g.size());
}

void ref_init_from_points(int n, const float *xyz, const float *rgb, int rgb_count, orc_model *out)
{
    std::vector<float> x(xyz, xyz + size_t(n) * 3), c;
    if (rgb)
        c.assign(rgb, rgb + rgb_count);
    const GaussianModel g = initGaussiansFromPoints(x, c);
    fromModel(g, out);
}

void *ref_trainer_new(const orc_model *initial, const orc_camera *cams, const uint8_t *rgb8,
                      int n_images, float extent, const orc_train_options *opt, int threads)
{
    auto *t = new RefTrainer();
    t->pool.reset(new ThreadPool(threads));
    std::vector<TrainImage> images;
    size_t off = 0;
    for (int k = 0; k < n_images; ++k) {
        TrainImage img;
        img.camera = toCam(&cams[k]);
        const size_t len = size_t(cams[k].width) * size_t(cams[k].height) * 3;
        img.rgb.assign(rgb8 + off, rgb8 + off + len);
        off += len;
        images.push_back(std::move(img));
    }
    t->trainer.reset(new Trainer(toModel(initial), std::move(images), extent, toOptions(opt), *t->pool));
    return t;
}
void ref_trainer_free(void *h) { delete static_cast<RefTrainer *>(h); }
void ref_trainer_step(void *h) { static_cast<RefTrainer *>(h)->trainer->step(); }
int ref_trainer_done(void *h) { return static_cast<RefTrainer *>(h)->trainer->done(); }
int ref_trainer_size(void *h) { return int(static_cast<RefTrainer *>(h)->trainer->model().size()); }
void ref_trainer_get_model(void *h, orc_model *out) { fromModel(static_cast<RefTrainer *>(h)->trainer->model(), out); }
void ref_trainer_stats(void *h, int *iteration, float *loss, float *psnr_, long *gaussians)
{
    const TrainStats &s = static_cast<RefTrainer *>(h)->trainer->stats();
    *iteration = s.iteration;
    *loss = s.loss;
    *psnr_ = s.psnr;
    *gaussians = long(s.gaussians);
}
float ref_trainer_evaluate_psnr(void *h) { return static_cast<RefTrainer *>(h)->trainer->evaluatePsnr(); }

// Runs Trainer::densifyAndPrune once on the current model with the given statistics; the
// trainer's own mt19937 supplies the split normals.
void ref_trainer_densify(void *h, const float *grad_accum, const int *grad_count, int opacity_was_reset)
{
    Trainer &t = *static_cast<RefTrainer *>(h)->trainer;
    const size_t n = t.m_model.size();
    t.m_gradAccum.assign(grad_accum, grad_accum + n);
    t.m_gradCount.assign(grad_count, grad_count + n);
    t.m_opacityWasReset = opacity_was_reset != 0;
    t.densifyAndPrune();
}

// Draws `count` normals the way densifyAndPrune does: one fresh std::normal_distribution over
// the trainer's generator (seeded with `seed`), used to reproduce the split offsets.
void ref_normals(unsigned seed, int count, float *out)
{
    std::mt19937 rng(seed);
    std::normal_distribution<float> normal(0.0f, 1.0f);
    for (int k = 0; k < count; ++k)
        out[k] = normal(rng);
}

} // extern "C"
