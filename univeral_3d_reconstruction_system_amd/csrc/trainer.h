// Host-side state shared by the C ABI (api.cpp) and the batched engine (batch.cpp): the packed
// device model, the rasterizer / optimiser / trainer objects and the trainer's step helpers.
#pragma once
#include <random>
#include <vector>

#include "kernels.h"

namespace u3d {

// Parameter-shaped device storage: [means 3c | scales 3c | quats 4c | opacities c | colors 3c]
// for capacity c, so a whole model (or its gradients / moments) is one allocation.
struct DevModel {
    DevBuf<float> buf;
    int64_t n = 0;
    int64_t cap = 0;
    int reserve(int64_t c)
    {
        if (c <= cap)
            return U3D_OK;
        U3D_TRY(buf.ensure(size_t(c) * 14));
        cap = c;
        return U3D_OK;
    }
    u3d_gaussians view() const
    {
        float *b = buf.p;
        return {b, b + 3 * cap, b + 6 * cap, b + 10 * cap, b + 11 * cap, n};
    }
    int zero(hipStream_t s) const
    {
        if (cap > 0)
            U3D_HIP(hipMemsetAsync(buf.p, 0, size_t(cap) * 14 * sizeof(float), s));
        return U3D_OK;
    }
    void swap(DevModel &o)
    {
        std::swap(buf.p, o.buf.p);
        std::swap(buf.cap, o.buf.cap);
        std::swap(n, o.n);
        std::swap(cap, o.cap);
    }
};

inline int copyFields(hipStream_t s, const u3d_gaussians &src, const u3d_gaussians &dst, int64_t n,
               hipMemcpyKind kind)
{
    if (n <= 0)
        return U3D_OK;
    const size_t f = sizeof(float);
    U3D_HIP(hipMemcpyAsync(dst.means, src.means, size_t(n) * 3 * f, kind, s));
    U3D_HIP(hipMemcpyAsync(dst.scales, src.scales, size_t(n) * 3 * f, kind, s));
    U3D_HIP(hipMemcpyAsync(dst.quats, src.quats, size_t(n) * 4 * f, kind, s));
    U3D_HIP(hipMemcpyAsync(dst.opacities, src.opacities, size_t(n) * f, kind, s));
    U3D_HIP(hipMemcpyAsync(dst.colors, src.colors, size_t(n) * 3 * f, kind, s));
    return U3D_OK;
}

inline int checkModel(const u3d_gaussians *m, const char *what)
{
    U3D_CHECK(m != nullptr, U3D_ERR_INVALID, std::string("null ") + what);
    U3D_CHECK(m->n >= 0 && m->n < (int64_t(1) << 31), U3D_ERR_INVALID, std::string("bad size of ") + what);
    if (m->n > 0)
        U3D_CHECK(m->means && m->scales && m->quats && m->opacities && m->colors, U3D_ERR_INVALID,
                  std::string("null field in ") + what);
    return U3D_OK;
}

inline int bitsFor(int64_t values)
{
    int b = 1;
    while ((int64_t(1) << b) < values)
        ++b;
    return b;
}

class BatchEngine;
constexpr int kMaxViewsPerStep = 64;
int createBatchEngine(u3d_ctx *ctx, bool useGraph, BatchEngine **out);
void destroyBatchEngine(BatchEngine *e);

} // namespace u3d

struct u3d_raster {
    u3d_ctx *ctx = nullptr;
    int64_t n = 0;
    int64_t numPairs = 0;
    u3d_camera cam{};
    u3d_raster_settings st{};
    u3d::RasterGeometry geo{};
    bool haveForward = false;
    u3d::DevBuf<float4> rec0, rec1;
    u3d::DevBuf<float> rec2, depth;
    u3d::DevBuf<int> radius;
    u3d::DevBuf<int4> rect;
    u3d::DevBuf<uint8_t> flags;
    u3d::DevBuf<uint32_t> depthKey, nTiles, keyAlt, valA, valAlt, ntSorted, pairOffsets;
    u3d::DevBuf<uint32_t> pairTile, pairGauss, pairTileAlt, pairGaussAlt;
    u3d::DevBuf<uint32_t> hist, histScan, scanScratch, total;
    u3d::DevBuf<uint64_t> sortScratch; // small-sort state, lookback words and histograms
    uint32_t sortEpoch = 0;            // small-sort calls so far (tags its lookback words)
    u3d::DevBuf<int> tileOffsets;
    const int *tileList = nullptr;
    // tile binning (tile grids the binning supports): the final lists (sorted on depth),
    // per-(group, tile) counters, the per-tile sort's scratch, the lists in index order, the tile
    // orders
    // (heavy-first walks, longest lists first)
    u3d::DevBuf<uint32_t> tileListBuf, indexList, binCounts, tileTotal, tileSortScratch;
    u3d::DevBuf<int> heavyFirst, longFirst;
    // the pair count: the projection's per-block sums and the pinned host word
    // (call number << 32 | pair count) their sum is published to
    u3d::DevBuf<uint32_t> blockTiles;
    u3d::PinnedScratch pinnedTotal;
    uint32_t totalSeq = 0;
    u3d::DevBuf<float> finalT, grad2d;
    bool grad2dZero = false; // grad2d holds zeros in all cap entries
    // segmented backward (16-pixel tiles): forward checkpoints and the work-item scan
    bool segmented = false;
    u3d::DevBuf<float4> ckpt;
    u3d::DevBuf<uint64_t> hitBits; // the forward's cull per (tile step, quadrant)
    u3d::DevBuf<uint32_t> segCount, segStart, segTotal;
    u3d::DevBuf<unsigned> segCounter;
    u3d::DevBuf<int> itemTile;
    int occSeg = 0;
    u3d::DevBuf<int> nContrib;
    u3d::DevBuf<float> dbgMean2d, dbgConic, dbgColor, dbgOpacity;
    u3d::ScreenBuffers screen() const
    {
        return {rec0.p, rec1.p, rec2.p, depth.p, radius.p, rect.p, flags.p, depthKey.p, nTiles.p};
    }
};

struct u3d_adam {
    u3d_ctx *ctx = nullptr;
    u3d::DevModel m, v, spare;
    int step = 0;
};

struct u3d_trainer {
    u3d_ctx *ctx = nullptr;
    u3d::DevModel params, spare, grads, adamM, adamV;
    int adamStep = 0;
    std::vector<u3d_camera> cams;
    std::vector<size_t> imageOffset;
    u3d::DevBuf<uint8_t> targets;
    float extent = 1.0f;
    u3d_train_options opt{};
    u3d_raster raster;
    std::mt19937 rng;
    std::vector<int> order;
    size_t orderPos = 0;
    int iteration = 0;
    bool opacityWasReset = false;
    u3d::DevBuf<float> render, dloss, gradAccum, evalOut;
    u3d::DevBuf<float> ssimMaps; // the loss's SSIM partials, 9 floats per pixel
    // Per-step loss statistics {total, l1, ssim, psnr} per view, [iterations][B][4], in pinned
    // host memory: the loss's final reduction stores them there (psnr last, with a release),
    // over a sentinel in the psnr slot, so the host reads a step's values without a stream sync.
    u3d::PinnedScratch stepStats;
    float *statsSlot(size_t k) { return static_cast<float *>(stepStats.p) + k * 4; }
    u3d::DevBuf<int> gradCount;
    u3d::DevBuf<double> partials;
    int statsFolded = 0;
    u3d_train_stats stats{};
    // steps, Gaussians, pixels, pairs, views, steps whose pair list overflowed its capacity;
    // the evaluated pairs (last counter) accumulate on the device, in evalPairs
    int64_t counters[U3D_NUM_COUNTERS] = {};
    u3d::DevBuf<unsigned long long> evalPairs; // 64 partial counts
    // Per-view tile walk costs (largest contributor count per tile at the view's last forward)
    // that order the next forward of the same view; view k at walkCostOffset[k].
    u3d::DevBuf<uint32_t> walkCost;
    std::vector<size_t> walkCostOffset;
    // Spherical-harmonic coefficients beyond DC (extension): 45 planes of n floats (coefficient
    // c of Gaussian i at c * n + i), their Adam moments, the batched engine's gradient and a
    // spare for densification; capacity 45 * cap each.
    int shDegree = 0;
    u3d::DevBuf<float> shP, shG, shM, shV, shSpare;
    // Batched training (extension): views per step, whole-step graph, the engine that runs it
    // (owned by this trainer when it trains alone; null inside a scene set).
    int viewsPerStep = 1;
    bool useGraph = false;
    u3d::BatchEngine *engine = nullptr;
};

namespace u3d {

// evalCount (nullable, 64 device counters): the forward adds its evaluated (pixel, Gaussian)
// pairs, the sum of the contributor counts, when it composites 16-pixel tiles.
// sh (nullable): view-dependent colour coefficients.
// walkCost (nullable, tiles entries): this view's per-tile walk costs from its last forward,
// which order the tiles heavy-first, replaced by this forward's (16-pixel tiles).
int rasterForwardImpl(u3d_raster *r, const u3d_gaussians &m, const u3d_camera &cam,
                      const u3d_raster_settings &st, float *rgb, const ShCoefs *sh = nullptr,
                      unsigned long long *evalCount = nullptr, uint32_t *walkCost = nullptr);
void adamCoefficients(int step, const u3d_learning_rates &lr, float stepSize[5], float *invSqrtBc2);
// Next view of the trainer's shuffled epoch order (Trainer.cpp:57-64).
int trainerNextView(u3d_trainer *t);
// Position learning rate of the current iteration (Trainer.cpp:82-84).
float trainerMeansLr(const u3d_trainer *t);
// ++iteration, then the densification / opacity-reset schedule (Trainer.cpp:94-101).
int trainerFinishStep(u3d_trainer *t);
// The trainer's SH coefficients as kernels read them: one plane of n floats per coefficient.
inline ShCoefs trainerSh(const u3d_trainer *t)
{
    ShCoefs sh;
    if (t->shDegree > 0) {
        sh.coef = t->shP.p;
        sh.gs = 1;
        sh.cs = t->params.n;
        sh.degree = t->shDegree;
    }
    return sh;
}
// SH coefficients train at the colour rate times this scale (3DGS uses 1/20).
inline float shRateScale(const u3d_trainer *t)
{
    return t->opt.sh_lr_scale > 0.0f ? t->opt.sh_lr_scale : 1.0f / 20.0f;
}
// One optimiser step of every trainer in `scenes` (viewsPerStep views each) through `engine`:
// one projection, one sort and one persistent compositing launch per direction for all views.
int batchTrainStep(BatchEngine *engine, u3d_trainer *const *scenes, int numScenes);

} // namespace u3d
