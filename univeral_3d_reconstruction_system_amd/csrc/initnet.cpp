// Seed-point initialisation on the device: the kNN-distance initialisation of the reference
// (initGaussiansFromPoints, GaussianModel.cpp:68-144) and the learned initialiser that corrects
// it (no reference counterpart).
//
// The learned initialiser is an MLP 80 -> 64 -> 256 -> 256 -> 256 -> 16 (GELU after the three
// 256-wide layers; 68 of the 80 inputs and 11 of the 16 outputs are used). Weights and
// activations are BF16, accumulation FP32 on MFMA (kernels_mlp.hip); the optimiser keeps FP32
// master weights and refreshes their BF16 copies after each step.
#include <algorithm>
#include <cmath>
#include <memory>
#include <random>
#include <vector>

#include "trainer.h"

namespace u3d {

namespace {

constexpr int kLayers = 5;
constexpr int kIn[kLayers] = {kFeatureDim, 64, 256, 256, 256};
constexpr int kOut[kLayers] = {64, 256, 256, 256, 16};
constexpr int kGelu[kLayers] = {0, 1, 1, 1, 0};
constexpr int kOutputs = 11;

int64_t paramCount()
{
    int64_t c = 0;
    for (int l = 0; l < kLayers; ++l)
        c += int64_t(kIn[l]) * kOut[l] + kOut[l];
    return c;
}

// Seed-point hash grid (sorted 64-bit cell keys) and per-point scratch.
struct PointScratch {
    DevBuf<uint32_t> keyLo, keyHi, idx, keyLoAlt, keyHiAlt, idxAlt, hist, histScan;
    DevBuf<float> bounds, partial, meanD2, viewFrac;
    DevBuf<u3d_camera> cams;
    float lo[3] = {0, 0, 0};
    float cell = 1.0f;
    PointGrid grid{};
};

// Bounds, cell size (host cbrtf, as the reference) and the sorted key grid.
int buildGrid(u3d_ctx *ctx, PointScratch &ps, const float *xyz, int64_t n)
{
    const hipStream_t s = ctx->stream;
    const size_t nn = size_t(std::max<int64_t>(n, 1));
    U3D_TRY(ps.bounds.ensure(8));
    U3D_TRY(ps.partial.ensure(size_t(kBoundsBlocks) * 6));
    U3D_TRY(launchPointBounds(s, xyz, n, ps.partial.p, ps.bounds.p));
    float b[6];
    U3D_HIP(hipMemcpyAsync(b, ps.bounds.p, sizeof b, hipMemcpyDeviceToHost, s));
    U3D_HIP(hipStreamSynchronize(s));
    // GaussianModel.cpp:75-87
    const float extent = std::max({b[3] - b[0], b[4] - b[1], b[5] - b[2], 1e-6f});
    ps.cell = std::max(extent / std::cbrt(float(n)) * 2.0f, 1e-6f);
    for (int k = 0; k < 3; ++k)
        ps.lo[k] = b[k];
    for (DevBuf<uint32_t> *d : {&ps.keyLo, &ps.keyHi, &ps.idx, &ps.keyLoAlt, &ps.keyHiAlt, &ps.idxAlt})
        U3D_TRY(d->ensure(nn));
    U3D_TRY(ps.hist.ensure(radixHistSize(int64_t(nn))));
    U3D_TRY(ps.histScan.ensure(radixScanScratchSize(int64_t(nn))));
    U3D_TRY(launchCellKeys(s, xyz, n, ps.lo, ps.cell, ps.keyLo.p, ps.keyHi.p, ps.idx.p));
    // 64-bit key = two stable 32-bit passes: low half, then the (permuted) high half.
    RadixScratch rs{ps.hist.p, ps.hist.cap, ps.histScan.p, ps.histScan.cap};
    bool alt = false;
    U3D_TRY(radixSortPairs(s, ps.keyLo.p, ps.idx.p, ps.keyLoAlt.p, ps.idxAlt.p, n, 32, rs, &alt));
    uint32_t *order = alt ? ps.idxAlt.p : ps.idx.p;
    uint32_t *loSorted = alt ? ps.keyLoAlt.p : ps.keyLo.p;
    uint32_t *hiPerm = alt ? ps.keyLo.p : ps.keyLoAlt.p; // free buffer
    U3D_TRY(launchGatherU32(s, ps.keyHi.p, order, hiPerm, n, n));
    // (hi, position in the low-sorted order) -> stable sort by hi
    uint32_t *pos = alt ? ps.idx.p : ps.idxAlt.p; // free buffer
    U3D_TRY(launchIota(s, pos, n));
    uint32_t *hiOut = ps.keyHi.p, *posOut = ps.keyHiAlt.p;
    bool alt2 = false;
    U3D_TRY(radixSortPairs(s, hiPerm, pos, hiOut, posOut, n, 32, rs, &alt2));
    const uint32_t *hiSorted = alt2 ? hiOut : hiPerm;
    const uint32_t *posSorted = alt2 ? posOut : pos;
    // final arrays: lo[posSorted], order[posSorted]
    uint32_t *loFinal = alt2 ? pos : posOut;
    uint32_t *idxFinal = alt2 ? hiPerm : hiOut;
    U3D_TRY(launchGatherU32(s, loSorted, posSorted, loFinal, n, n));
    U3D_TRY(launchGatherU32(s, order, posSorted, idxFinal, n, n));
    ps.grid = PointGrid{hiSorted, loFinal, idxFinal};
    return U3D_OK;
}

int knnInit(u3d_ctx *ctx, PointScratch &ps, const float *xyz, const float *rgb, int rgbCount,
            int64_t n, u3d_gaussians &out)
{
    U3D_TRY(buildGrid(ctx, ps, xyz, n));
    U3D_TRY(ps.meanD2.ensure(size_t(std::max<int64_t>(n, 1))));
    volatile float p = 0.1f; // logit(0.1f) with the host logf (GaussianModel.h:12)
    const float pf = p;
    const float opacityLogit = std::log(pf / (1.0f - pf));
    const hipStream_t s = ctx->stream;
    U3D_TRY(launchKnnInit(s, xyz, rgb, rgbCount, n, ps.lo, ps.cell, ps.grid, opacityLogit, out,
                          ps.meanD2.p));
    // Log-scales with the host libm, as the reference (GaussianModel.cpp:131): the reference's
    // logf is not correctly rounded, so a device log differs from it in the last bit for about
    // one point in a thousand. n floats each way, once per initialisation.
    std::vector<float> d2(static_cast<size_t>(n)), sc(static_cast<size_t>(n) * 3);
    U3D_HIP(hipMemcpyAsync(d2.data(), ps.meanD2.p, d2.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    U3D_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < d2.size(); ++i) {
        const float logScale = std::log(std::sqrt(std::max(d2[i], 1e-14f)));
        sc[i * 3] = sc[i * 3 + 1] = sc[i * 3 + 2] = logScale;
    }
    U3D_HIP(hipMemcpyAsync(out.scales, sc.data(), sc.size() * sizeof(float), hipMemcpyHostToDevice, s));
    U3D_HIP(hipStreamSynchronize(s));
    return U3D_OK;
}

int features(u3d_ctx *ctx, PointScratch &ps, const float *xyz, const float *rgb, int rgbCount,
             int64_t n, const u3d_camera *camsHost, int numCams, float nearPlane, float *feat)
{
    const hipStream_t s = ctx->stream;
    const float *vf = nullptr;
    if (camsHost && numCams > 0) {
        U3D_TRY(ps.cams.ensure(size_t(numCams)));
        U3D_TRY(ps.viewFrac.ensure(size_t(std::max<int64_t>(n, 1))));
        U3D_HIP(hipMemcpyAsync(ps.cams.p, camsHost, sizeof(u3d_camera) * size_t(numCams),
                               hipMemcpyHostToDevice, s));
        U3D_TRY(launchViewFraction(s, xyz, n, ps.cams.p, numCams, nearPlane, ps.viewFrac.p));
        vf = ps.viewFrac.p;
        U3D_HIP(hipStreamSynchronize(s)); // the host camera array may go away after return
    }
    return launchKnnFeatures(s, xyz, rgb, rgbCount, n, ps.lo, ps.cell, ps.grid, ps.meanD2.p, vf, feat);
}

} // namespace
} // namespace u3d

using namespace u3d;

struct u3d_init_net {
    u3d_ctx *ctx = nullptr;
    DevBuf<float> W[kLayers], b[kLayers], gW[kLayers], gb[kLayers];
    DevBuf<float> mW[kLayers], vW[kLayers], mb[kLayers], vb[kLayers];
    DevBuf<uint16_t> Wb[kLayers];
    int adamStep = 0;
    // activations of the last forward (capacity rows)
    int64_t cap = 0;
    DevBuf<uint16_t> X, A[kLayers], dHb[2];
    DevBuf<float> H[kLayers], dH[2], dOut, loss, colW;
    PointScratch points;

    int reserve(int64_t rows)
    {
        if (rows <= cap)
            return U3D_OK;
        const size_t r = size_t(rows);
        U3D_TRY(X.ensure(r * kFeatureDim));
        for (int l = 0; l < kLayers; ++l) {
            U3D_TRY(A[l].ensure(r * size_t(kOut[l])));
            U3D_TRY(H[l].ensure(r * size_t(kOut[l])));
        }
        for (int k = 0; k < 2; ++k) {
            U3D_TRY(dHb[k].ensure(r * 256));
            U3D_TRY(dH[k].ensure(r * 256));
        }
        U3D_TRY(dOut.ensure(r * 16));
        cap = rows;
        return U3D_OK;
    }

    int forward(const float *feat, int64_t n)
    {
        const hipStream_t s = ctx->stream;
        U3D_TRY(reserve(n));
        U3D_TRY(launchToBf16(s, feat, X.p, n * kFeatureDim));
        const uint16_t *in = X.p;
        for (int l = 0; l < kLayers; ++l) {
            GemmArgs g{};
            g.A = in;
            g.sam = kIn[l];
            g.sak = 1;
            g.B = Wb[l].p; // B(k, n) = W[n][k]
            g.sbk = 1;
            g.sbn = kIn[l];
            g.M = int(n);
            g.N = kOut[l];
            g.K = kIn[l];
            g.C = H[l].p;
            g.ldc = kOut[l];
            g.bias = b[l].p;
            g.Y = l + 1 < kLayers ? A[l].p : nullptr;
            g.ldy = kOut[l];
            g.gelu = kGelu[l];
            g.splitK = 1;
            U3D_TRY(launchGemmBf16(s, g));
            in = A[l].p;
        }
        return U3D_OK;
    }

    // Gradients of the last forward given dOut (fp32 [n x 16]); accumulates into gW / gb.
    int backward(int64_t n)
    {
        const hipStream_t s = ctx->stream;
        for (int l = 0; l < kLayers; ++l) {
            U3D_HIP(hipMemsetAsync(gW[l].p, 0, size_t(kIn[l]) * kOut[l] * sizeof(float), s));
            U3D_HIP(hipMemsetAsync(gb[l].p, 0, size_t(kOut[l]) * sizeof(float), s));
        }
        // dH of the output layer = dOut; bf16 copy for the GEMMs
        const float *dHf = dOut.p;
        U3D_TRY(launchToBf16(s, dOut.p, dHb[0].p, n * 16));
        const uint16_t *dHbCur = dHb[0].p;
        int cur = 0;
        const int splits = int(std::max<int64_t>(1, std::min<int64_t>(128, n / 512)));
        for (int l = kLayers - 1; l >= 0; --l) {
            const int64_t ldH = kOut[l];
            // gW[l] += dH^T A_{l-1}  (A_{-1} = X)
            const uint16_t *prev = l > 0 ? A[l - 1].p : X.p;
            GemmArgs w{};
            w.A = dHbCur; // A(m = out, k = row) = dH[row][out]
            w.sam = 1;
            w.sak = ldH;
            w.B = prev; // B(k = row, n = in) = prev[row][in]
            w.sbk = kIn[l];
            w.sbn = 1;
            w.M = kOut[l];
            w.N = kIn[l];
            w.K = int(n);
            w.C = gW[l].p;
            w.ldc = kIn[l];
            w.splitK = splits;
            U3D_TRY(launchGemmBf16(s, w));
            U3D_TRY(launchColumnSums(s, dHf, ldH, int(n), kOut[l], gb[l].p));
            if (l == 0)
                break;
            // dH[l-1] = (dH[l] W[l]) * GELU'(H[l-1]) (layer 0 is linear: no gate)
            const int nxt = cur ^ 1;
            GemmArgs d{};
            d.A = dHbCur; // A(m = row, k = out)
            d.sam = ldH;
            d.sak = 1;
            d.B = Wb[l].p; // B(k = out, n = in) = W[out][in]
            d.sbk = kIn[l];
            d.sbn = 1;
            d.M = int(n);
            d.N = kIn[l];
            d.K = kOut[l];
            d.C = dH[nxt].p;
            d.ldc = kIn[l];
            d.Y = dHb[nxt].p;
            d.ldy = kIn[l];
            d.gate = kGelu[l - 1] ? H[l - 1].p : nullptr;
            d.ldg = kOut[l - 1];
            d.splitK = 1;
            U3D_TRY(launchGemmBf16(s, d));
            dHf = dH[nxt].p;
            dHbCur = dHb[nxt].p;
            cur = nxt;
        }
        return U3D_OK;
    }
};

extern "C" {

int u3d_points_knn_init(u3d_ctx *ctx, const float *xyz, const float *rgb, int rgb_count, int64_t n,
                        u3d_gaussians *out)
{
    U3D_TRY(useDevice(ctx));
    U3D_CHECK(out && n >= 0 && (n == 0 || xyz), U3D_ERR_INVALID, "bad argument");
    U3D_CHECK(rgb_count == 0 || rgb, U3D_ERR_INVALID, "null rgb");
    U3D_TRY(checkModel(out, "out"));
    U3D_CHECK(out->n >= n, U3D_ERR_CAPACITY, "output model smaller than the point count");
    if (n == 0)
        return U3D_OK;
    PointScratch ps;
    U3D_TRY(knnInit(ctx, ps, xyz, rgb, rgb_count, n, *out));
    U3D_HIP(hipStreamSynchronize(ctx->stream));
    return U3D_OK;
}

int u3d_points_features(u3d_ctx *ctx, const float *xyz, const float *rgb, int rgb_count, int64_t n,
                        const u3d_camera *cameras, int n_cameras, float near_plane, float *features_out)
{
    U3D_TRY(useDevice(ctx));
    U3D_CHECK(n >= 0 && (n == 0 || (xyz && features_out)), U3D_ERR_INVALID, "bad argument");
    U3D_CHECK(rgb_count == 0 || rgb, U3D_ERR_INVALID, "null rgb");
    if (n == 0)
        return U3D_OK;
    PointScratch ps;
    DevModel tmp;
    U3D_TRY(tmp.reserve(n));
    tmp.n = n;
    u3d_gaussians tv = tmp.view();
    U3D_TRY(knnInit(ctx, ps, xyz, rgb, rgb_count, n, tv));
    U3D_TRY(features(ctx, ps, xyz, rgb, rgb_count, n, cameras, n_cameras, near_plane, features_out));
    U3D_HIP(hipStreamSynchronize(ctx->stream));
    return U3D_OK;
}

int64_t u3d_init_net_param_count(void) { return paramCount(); }

int u3d_init_net_create(u3d_ctx *ctx, unsigned seed, u3d_init_net **out)
{
    U3D_TRY(useDevice(ctx));
    U3D_CHECK(out != nullptr, U3D_ERR_INVALID, "null out");
    std::unique_ptr<u3d_init_net> net(new u3d_init_net());
    net->ctx = ctx;
    const hipStream_t s = ctx->stream;
    std::mt19937 rng(seed);
    std::vector<float> host;
    for (int l = 0; l < kLayers; ++l) {
        const size_t nw = size_t(kIn[l]) * kOut[l], nb = size_t(kOut[l]);
        for (DevBuf<float> *d : {&net->W[l], &net->gW[l], &net->mW[l], &net->vW[l]})
            U3D_TRY(d->ensure(nw));
        for (DevBuf<float> *d : {&net->b[l], &net->gb[l], &net->mb[l], &net->vb[l]})
            U3D_TRY(d->ensure(nb));
        U3D_TRY(net->Wb[l].ensure(nw));
        // He-uniform for the GELU layers, Glorot-uniform otherwise; the output layer starts at
        // zero so that an untrained network reproduces the kNN initialisation.
        const float bound = l == kLayers - 1 ? 0.0f
                            : kGelu[l]        ? std::sqrt(6.0f / float(kIn[l]))
                                              : std::sqrt(6.0f / float(kIn[l] + kOut[l]));
        std::uniform_real_distribution<float> u(-bound, bound);
        host.assign(nw, 0.0f);
        for (float &x : host)
            x = bound > 0.0f ? u(rng) : 0.0f;
        U3D_HIP(hipMemcpyAsync(net->W[l].p, host.data(), nw * sizeof(float), hipMemcpyHostToDevice, s));
        U3D_HIP(hipStreamSynchronize(s));
        U3D_HIP(hipMemsetAsync(net->b[l].p, 0, nb * sizeof(float), s));
        U3D_HIP(hipMemsetAsync(net->mW[l].p, 0, nw * sizeof(float), s));
        U3D_HIP(hipMemsetAsync(net->vW[l].p, 0, nw * sizeof(float), s));
        U3D_HIP(hipMemsetAsync(net->mb[l].p, 0, nb * sizeof(float), s));
        U3D_HIP(hipMemsetAsync(net->vb[l].p, 0, nb * sizeof(float), s));
        U3D_TRY(launchToBf16(s, net->W[l].p, net->Wb[l].p, int64_t(nw)));
    }
    U3D_TRY(net->loss.ensure(4));
    U3D_TRY(net->colW.ensure(16));
    U3D_HIP(hipStreamSynchronize(s));
    *out = net.release();
    return U3D_OK;
}

void u3d_init_net_destroy(u3d_init_net *net)
{
    if (!net)
        return;
    (void)hipSetDevice(net->ctx->device);
    (void)hipStreamSynchronize(net->ctx->stream);
    delete net;
}

namespace {
int copyParams(u3d_init_net *net, float *host, bool toHost, bool grads)
{
    const hipStream_t s = net->ctx->stream;
    size_t o = 0;
    for (int l = 0; l < kLayers; ++l) {
        const size_t nw = size_t(kIn[l]) * kOut[l], nb = size_t(kOut[l]);
        float *w = grads ? net->gW[l].p : net->W[l].p, *b = grads ? net->gb[l].p : net->b[l].p;
        if (toHost) {
            U3D_HIP(hipMemcpyAsync(host + o, w, nw * sizeof(float), hipMemcpyDeviceToHost, s));
            U3D_HIP(hipMemcpyAsync(host + o + nw, b, nb * sizeof(float), hipMemcpyDeviceToHost, s));
        } else {
            U3D_HIP(hipMemcpyAsync(w, host + o, nw * sizeof(float), hipMemcpyHostToDevice, s));
            U3D_HIP(hipMemcpyAsync(b, host + o + nw, nb * sizeof(float), hipMemcpyHostToDevice, s));
            U3D_TRY(launchToBf16(s, net->W[l].p, net->Wb[l].p, int64_t(nw)));
        }
        o += nw + nb;
    }
    U3D_HIP(hipStreamSynchronize(s));
    return U3D_OK;
}
} // namespace

int u3d_init_net_get_params(u3d_init_net *net, float *params_host)
{
    U3D_CHECK(net && params_host, U3D_ERR_INVALID, "null argument");
    U3D_TRY(useDevice(net->ctx));
    return copyParams(net, params_host, true, false);
}

int u3d_init_net_set_params(u3d_init_net *net, const float *params_host)
{
    U3D_CHECK(net && params_host, U3D_ERR_INVALID, "null argument");
    U3D_TRY(useDevice(net->ctx));
    return copyParams(net, const_cast<float *>(params_host), false, false);
}

int u3d_init_net_get_grads(u3d_init_net *net, float *grads_host)
{
    U3D_CHECK(net && grads_host, U3D_ERR_INVALID, "null argument");
    U3D_TRY(useDevice(net->ctx));
    return copyParams(net, grads_host, true, true);
}

int u3d_init_net_forward(u3d_init_net *net, const float *features, int64_t n, float *out)
{
    U3D_CHECK(net && n >= 0 && (n == 0 || (features && out)), U3D_ERR_INVALID, "bad argument");
    U3D_TRY(useDevice(net->ctx));
    if (n == 0)
        return U3D_OK;
    U3D_TRY(net->forward(features, n));
    U3D_HIP(hipMemcpyAsync(out, net->H[kLayers - 1].p, size_t(n) * 16 * sizeof(float),
                           hipMemcpyDeviceToDevice, net->ctx->stream));
    return U3D_OK;
}

int u3d_init_net_train_step(u3d_init_net *net, const float *features, const float *targets,
                            const float *weights16_host, int64_t n, float lr, float *loss_host)
{
    U3D_CHECK(net && features && targets && weights16_host && n > 0, U3D_ERR_INVALID, "bad argument");
    U3D_TRY(useDevice(net->ctx));
    const hipStream_t s = net->ctx->stream;
    U3D_TRY(net->forward(features, n));
    U3D_HIP(hipMemcpyAsync(net->colW.p, weights16_host, 16 * sizeof(float), hipMemcpyHostToDevice, s));
    U3D_HIP(hipMemsetAsync(net->loss.p, 0, sizeof(float), s));
    U3D_TRY(launchMseLoss(s, net->H[kLayers - 1].p, targets, net->colW.p, int(n), kOutputs, 16,
                          net->dOut.p, net->loss.p));
    U3D_TRY(net->backward(n));
    ++net->adamStep;
    const float bc1 = 1.0f - std::pow(0.9f, float(net->adamStep));
    const float bc2 = 1.0f - std::pow(0.999f, float(net->adamStep));
    const float step = lr / bc1, inv = 1.0f / std::sqrt(bc2);
    for (int l = 0; l < kLayers; ++l) {
        U3D_TRY(launchAdamBf16(s, net->W[l].p, net->gW[l].p, net->mW[l].p, net->vW[l].p,
                               int64_t(kIn[l]) * kOut[l], step, inv, net->Wb[l].p));
        U3D_TRY(launchAdamBf16(s, net->b[l].p, net->gb[l].p, net->mb[l].p, net->vb[l].p, kOut[l],
                               step, inv, nullptr));
    }
    if (loss_host) {
        U3D_HIP(hipMemcpyAsync(loss_host, net->loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
        U3D_HIP(hipStreamSynchronize(s));
    }
    return U3D_OK;
}

int u3d_points_learned_init(u3d_ctx *ctx, u3d_init_net *net, const float *xyz, const float *rgb,
                            int rgb_count, int64_t n, const u3d_camera *cameras, int n_cameras,
                            float near_plane, u3d_gaussians *out)
{
    U3D_TRY(useDevice(ctx));
    U3D_CHECK(out && n >= 0 && (n == 0 || xyz), U3D_ERR_INVALID, "bad argument");
    U3D_CHECK(rgb_count == 0 || rgb, U3D_ERR_INVALID, "null rgb");
    U3D_TRY(checkModel(out, "out"));
    U3D_CHECK(out->n >= n, U3D_ERR_CAPACITY, "output model smaller than the point count");
    U3D_CHECK(!net || net->ctx == ctx, U3D_ERR_INVALID, "network belongs to another context");
    if (n == 0)
        return U3D_OK;
    std::unique_ptr<PointScratch> own;
    if (!net)
        own.reset(new PointScratch());
    PointScratch &ps = net ? net->points : *own;
    U3D_TRY(knnInit(ctx, ps, xyz, rgb, rgb_count, n, *out));
    if (net) { // no network: the kNN initialisation is the result (fallback)
        DevBuf<float> feat;
        U3D_TRY(feat.ensure(size_t(n) * kFeatureDim));
        U3D_TRY(features(ctx, ps, xyz, rgb, rgb_count, n, cameras, n_cameras, near_plane, feat.p));
        U3D_TRY(net->forward(feat.p, n));
        U3D_TRY(launchApplyCorrections(ctx->stream, net->H[kLayers - 1].p, 16, n, *out));
    }
    U3D_HIP(hipStreamSynchronize(ctx->stream));
    return U3D_OK;
}

} // extern "C"
