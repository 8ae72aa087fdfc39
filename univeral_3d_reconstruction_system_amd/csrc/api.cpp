// C ABI of libu3d_hip (include/u3d_hip.h): context, rasterizer, loss, Adam, densification and
// the device-resident trainer. Host orchestration only; every kernel lives in a kernels_*.hip
// translation unit.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <vector>

#include "trainer.h"

// Device scratch of u3d_photometric_loss_noalloc, for callers that cannot allocate: the loss
// kernels' per-block partial sums and the reduced value. Each GPU holds its own copy for the
// life of the process; host code reaches it only through the address hipGetSymbolAddress
// returns for that device (the array's name is a host-side stand-in).
constexpr size_t kStaticScratchBytes = size_t(1) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace u3d {

namespace {
thread_local std::string tError;
}

void setError(const std::string &msg) { tError = msg; }

int failHip(hipError_t e, const char *what, const char *file, int line)
{
    tError = std::string(what) + ": " + hipGetErrorString(e) + " (" + file + ":" + std::to_string(line) + ")";
    return U3D_ERR_HIP;
}

int waitPinnedWord(const uint64_t *p, uint32_t seq, hipStream_t s, uint32_t *value, bool *flag)
{
    // the word is (flag << 63 | seq << 32 | value) with a 31-bit seq
    auto seqOf = [](uint64_t w) { return uint32_t(w >> 32) & 0x7fffffffu; };
    if (flag)
        *flag = false;
    // Tight polling for the first millisecond (the usual wait is tens of microseconds); after
    // that the stream is also asked for an error about every 100 microseconds.
    using Clock = std::chrono::steady_clock;
    const Clock::time_point start = Clock::now();
    Clock::time_point nextQuery = start + std::chrono::milliseconds(1);
    for (uint32_t polls = 1;; ++polls) {
        const uint64_t w = __atomic_load_n(p, __ATOMIC_ACQUIRE);
        if (seqOf(w) == seq) {
            *value = uint32_t(w);
            if (flag)
                *flag = (w >> 63) != 0;
            return U3D_OK;
        }
        __builtin_ia32_pause();
        if ((polls & 255u) != 0 || Clock::now() < nextQuery)
            continue;
        nextQuery = Clock::now() + std::chrono::microseconds(100);
        const hipError_t e = hipStreamQuery(s);
        if (e == hipSuccess) {
            // the stream has drained, so the word was stored unless something went wrong
            const uint64_t w2 = __atomic_load_n(p, __ATOMIC_ACQUIRE);
            U3D_CHECK(seqOf(w2) == seq, U3D_ERR_HIP,
                      "the device-side pair count never arrived");
            *value = uint32_t(w2);
            if (flag)
                *flag = (w2 >> 63) != 0;
            return U3D_OK;
        }
        if (e != hipErrorNotReady)
            return failHip(e, "hipStreamQuery", __FILE__, __LINE__);
    }
}

// The device address of `scratch` on `device` (current), resolved once per device.
int staticScratchAddress(int device, unsigned char **out)
{
    static std::mutex mu;
    static std::vector<unsigned char *> perDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (size_t(device) >= perDevice.size())
        perDevice.resize(size_t(device) + 1, nullptr);
    if (!perDevice[size_t(device)]) {
        void *p = nullptr;
        U3D_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        U3D_CHECK(p != nullptr, U3D_ERR_HIP, "static scratch has no device address");
        perDevice[size_t(device)] = static_cast<unsigned char *>(p);
    }
    *out = perDevice[size_t(device)];
    return U3D_OK;
}

int waitPinnedU32(const uint32_t *p, uint32_t pending, hipStream_t s)
{
    // As waitPinnedWord: tight polling first, then the stream is asked for an error.
    using Clock = std::chrono::steady_clock;
    Clock::time_point nextQuery = Clock::now() + std::chrono::milliseconds(1);
    for (uint32_t polls = 1;; ++polls) {
        if (__atomic_load_n(p, __ATOMIC_ACQUIRE) != pending)
            return U3D_OK;
        __builtin_ia32_pause();
        if ((polls & 255u) != 0 || Clock::now() < nextQuery)
            continue;
        nextQuery = Clock::now() + std::chrono::microseconds(100);
        const hipError_t e = hipStreamQuery(s);
        if (e == hipSuccess) {
            U3D_CHECK(__atomic_load_n(p, __ATOMIC_ACQUIRE) != pending, U3D_ERR_HIP,
                      "a step's loss statistics never arrived");
            return U3D_OK;
        }
        if (e != hipErrorNotReady)
            return failHip(e, "hipStreamQuery", __FILE__, __LINE__);
    }
}

int useDevice(u3d_ctx *ctx)
{
    U3D_CHECK(ctx != nullptr, U3D_ERR_INVALID, "null context");
    U3D_HIP(hipSetDevice(ctx->device));
    return U3D_OK;
}

} // namespace u3d

using namespace u3d;

namespace u3d {

int rasterForwardImpl(u3d_raster *r, const u3d_gaussians &m, const u3d_camera &cam,
                      const u3d_raster_settings &st, float *rgb, const ShCoefs *shp,
                      unsigned long long *evalCount, uint32_t *walkCost)
{
    const ShCoefs sh = shp ? *shp : ShCoefs{};
    U3D_CHECK(sh.degree >= 0 && sh.degree <= 3, U3D_ERR_INVALID, "sh_degree must be 0..3");
    U3D_CHECK(st.tile_size > 0, U3D_ERR_INVALID, "tile_size must be positive");
    U3D_CHECK(cam.width >= 0 && cam.height >= 0, U3D_ERR_INVALID, "negative image size");
    const hipStream_t s = r->ctx->stream;
    const int64_t n = m.n;
    const int ts = st.tile_size;
    const int tilesX = (cam.width + ts - 1) / ts, tilesY = (cam.height + ts - 1) / ts;
    const int tiles = tilesX * tilesY;
    const size_t P = size_t(cam.width) * size_t(cam.height);
    r->n = n;
    r->cam = cam;
    r->st = st;
    r->geo = {cam.width, cam.height, ts, tilesX, tilesY};
    const size_t nn = size_t(n > 0 ? n : 1);
    U3D_TRY(r->rec0.ensure(nn));
    U3D_TRY(r->rec1.ensure(nn));
    U3D_TRY(r->rec2.ensure(nn));
    U3D_TRY(r->depth.ensure(nn));
    U3D_TRY(r->radius.ensure(nn));
    U3D_TRY(r->rect.ensure(nn));
    U3D_TRY(r->flags.ensure(nn));
    U3D_TRY(r->depthKey.ensure(nn));
    U3D_TRY(r->nTiles.ensure(nn));
    U3D_TRY(r->keyAlt.ensure(nn));
    U3D_TRY(r->valA.ensure(nn));
    U3D_TRY(r->valAlt.ensure(nn));
    U3D_TRY(r->ntSorted.ensure(nn));
    U3D_TRY(r->pairOffsets.ensure(nn));
    U3D_TRY(r->total.ensure(4));
    U3D_TRY(r->scanScratch.ensure(scanScratchSize(int64_t(nn))));
    U3D_TRY(r->tileOffsets.ensure(size_t(tiles) + 1));
    U3D_TRY(r->finalT.ensure(P > 0 ? P : 1));
    U3D_TRY(r->nContrib.ensure(P > 0 ? P : 1));
    const ScreenBuffers sb = r->screen();

    U3D_TRY(r->pinnedTotal.ensure(sizeof(uint64_t)));
    U3D_TRY(r->blockTiles.ensure(size_t(std::max(1, projectBlocks(n)))));
    // the device stores (seq << 32 | K) there; seq is never 0, the word's value until then
    uint64_t *pinnedWord = static_cast<uint64_t *>(r->pinnedTotal.p);
    __atomic_store_n(pinnedWord, uint64_t(0), __ATOMIC_RELEASE);
    r->totalSeq = (r->totalSeq + 1u) & 0x7fffffffu; // 31 bits, never 0
    if (r->totalSeq == 0u)
        r->totalSeq = 1u;
    const uint32_t seq = r->totalSeq;
    {
        StageScope scope(r->ctx, U3D_STAGE_PROJECT);
        U3D_TRY(launchProject(s, m, cam, st, tilesX, tilesY, sb, sh, r->blockTiles.p));
    }
    // The pair count K is the sum of the projection's block sums; it runs beside the next
    // launch and reaches the host through pinned memory while the device works on.
    SideSum pairSum;
    if (n > 0) {
        pairSum.in = r->blockTiles.p;
        pairSum.count = projectBlocks(n);
        pairSum.dev = r->total.p;
        pairSum.host = pinnedWord;
        pairSum.seq = seq;
    }
    std::unique_ptr<StageScope> sortScope(new StageScope(r->ctx, U3D_STAGE_SORT));
    uint32_t K = 0;
    const int *order = nullptr; // heavy-first tile order for the forward (binning path)
    if (tileBinningGroup(n, tilesX, tilesY) > 0) {
        // Tile lists in Gaussian index order (no pair count needed for the counts), then each
        // list sorted on the depth keys. The host reads the pair count K, summed by the first
        // workgroup of the count launch, while the counts run, and sizes the lists by it.
        U3D_TRY(r->binCounts.ensure(std::max<size_t>(1, tileBinningCountsSize(n, tilesX, tilesY))));
        U3D_TRY(r->tileTotal.ensure(tileBinningTotalsSize(tilesX, tilesY)));
        U3D_TRY(r->heavyFirst.ensure(size_t(tiles)));
        U3D_TRY(r->longFirst.ensure(size_t(tiles)));
        U3D_TRY(launchTileCounts(s, n, nullptr, r->nTiles.p, r->rect.p, nullptr, nullptr, tilesX,
                                 tilesY, r->binCounts.p, r->binCounts.cap, r->tileTotal.p,
                                 nullptr, n, &pairSum));
        if (n > 0)
            U3D_TRY(waitPinnedWord(pinnedWord, seq, s, &K, nullptr));
        const size_t kk = size_t(K > 0 ? K : 1);
        U3D_TRY(r->tileListBuf.ensure(kk));
        U3D_TRY(r->indexList.ensure(kk));
        U3D_TRY(r->tileSortScratch.ensure(tileSortScratchSize(r->indexList.cap)));
        U3D_TRY(launchTileEmit(s, n, nullptr, r->nTiles.p, r->rect.p, tilesX, tilesY,
                               r->binCounts.p, r->tileTotal.p, r->tileOffsets.p, r->indexList.p,
                               kk, r->heavyFirst.p, ts == 16 ? walkCost : nullptr, nullptr,
                               r->longFirst.p));
        if (n > 0)
            U3D_TRY(launchTileSort(s, tiles, r->tileOffsets.p, r->longFirst.p, r->indexList.p,
                                   r->tileListBuf.p, r->depthKey.p, r->tileSortScratch.p,
                                   r->tileSortScratch.cap, r->indexList.cap));
        order = n > 0 ? r->heavyFirst.p : nullptr;
        r->tileList = reinterpret_cast<const int *>(r->tileListBuf.p);
    } else {
        // Many tiles: depth order of the visible Gaussians (invisible keys sort last; ties keep
        // index order), pairs emitted in depth order, then a stable radix sort on the tile id.
        U3D_TRY(r->hist.ensure(radixHistSize(int64_t(nn))));
        U3D_TRY(r->histScan.ensure(radixScanScratchSize(int64_t(nn))));
        RadixScratch rs{r->hist.p, r->hist.cap, r->histScan.p, r->histScan.cap};
        rs.valsAreIndex = true;
        {
            const uint64_t *before = r->sortScratch.p;
            U3D_TRY(r->sortScratch.ensure(onesweepScratchWords()));
            if (r->sortScratch.p != before) // the small sort expects zeros in a new scratch
                U3D_TRY(launchFillI32(s, reinterpret_cast<int *>(r->sortScratch.p), 0,
                                      int64_t(r->sortScratch.cap) * 2));
            rs.onesweep = r->sortScratch.p;
            rs.onesweepCap = r->sortScratch.cap;
            rs.onesweepEpoch = &r->sortEpoch;
            rs.spinLimit = r->ctx->sortSpinLimit;
        }
        rs.sideIn = pairSum.in;
        rs.sideCount = pairSum.count;
        rs.sideDev = pairSum.dev;
        rs.sideHost = pairSum.host;
        rs.sideSeq = pairSum.seq;
        bool alt = false;
        U3D_TRY(radixSortPairs(s, r->depthKey.p, r->valA.p, r->keyAlt.p, r->valAlt.p, n, 32, rs,
                               &alt));
        const uint32_t *sorted = alt ? r->valAlt.p : r->valA.p;
        if (n > 0)
            U3D_TRY(waitPinnedWord(pinnedWord, seq, s, &K, nullptr));
        // The pair emission has no check of the depth sort, so its error word is read here (a
        // synchronisation, on this rare path only).
        if (n > 0) {
            uint32_t sortError = 0;
            U3D_TRY(onesweepErrorWord(r->sortScratch.p, s, &sortError));
            U3D_CHECK(sortError == 0, U3D_ERR_HIP,
                      "depth sort: a lookback wait was abandoned (the order is not valid)");
        }
        const size_t kk = size_t(K > 0 ? K : 1);
        U3D_TRY(launchGatherU32(s, r->nTiles.p, sorted, r->ntSorted.p, n, n));
        U3D_TRY(exclusiveScanU32(s, r->ntSorted.p, r->pairOffsets.p, n, r->total.p,
                                 r->scanScratch.p, r->scanScratch.cap));
        U3D_TRY(r->pairTile.ensure(kk));
        U3D_TRY(r->pairGauss.ensure(kk));
        U3D_TRY(r->pairTileAlt.ensure(kk));
        U3D_TRY(r->pairGaussAlt.ensure(kk));
        U3D_TRY(r->hist.ensure(radixHistSize(int64_t(kk))));
        U3D_TRY(r->histScan.ensure(radixScanScratchSize(int64_t(kk))));
        RadixScratch rs2{r->hist.p, r->hist.cap, r->histScan.p, r->histScan.cap};
        U3D_TRY(launchEmitPairs(s, n, sorted, r->pairOffsets.p, r->rect.p, tilesX, r->pairTile.p,
                                r->pairGauss.p, r->ntSorted.p, int64_t(K)));
        bool alt2 = false;
        U3D_TRY(radixSortPairs(s, r->pairTile.p, r->pairGauss.p, r->pairTileAlt.p,
                               r->pairGaussAlt.p, int64_t(K), bitsFor(tiles), rs2, &alt2));
        const uint32_t *tileSorted = alt2 ? r->pairTileAlt.p : r->pairTile.p;
        r->tileList = reinterpret_cast<const int *>(alt2 ? r->pairGaussAlt.p : r->pairGauss.p);
        U3D_TRY(launchTileOffsets(s, tileSorted, int64_t(K), tiles, r->tileOffsets.p));
    }
    r->numPairs = int64_t(K);
    sortScope.reset();
    // Checkpoints for the segmented backward (16-pixel tiles): state at every boundary of
    // kSegEntries records of each tile list.
    r->segmented = r->ctx->segmentedBackward && ts == 16;
    SegmentState seg{nullptr, nullptr};
    if (r->segmented) {
        const size_t slots = size_t(K) / kSegEntries + size_t(tiles) + 1;
        U3D_TRY(r->ckpt.ensure(slots * 256));
        U3D_TRY(r->segCount.ensure(size_t(tiles) * kSegCountsPerTile + 1));
        U3D_TRY(r->hitBits.ensure((size_t(K) / 64 + size_t(tiles) + 2) * 4));
        seg = SegmentState{r->ckpt.p, r->segCount.p, r->hitBits.p};
    }
    StageScope scope(r->ctx, U3D_STAGE_RASTER_FWD);
    U3D_TRY(launchRasterForward(s, r->geo, st, r->tileOffsets.p, r->tileList, sb, rgb,
                                r->finalT.p, r->nContrib.p, seg, order, evalCount,
                                order && ts == 16 ? walkCost : nullptr));
    r->haveForward = true;
    return U3D_OK;
}

namespace {

// Buffers and arguments of the segmented backward's work-item scan for the last forward.
int segmentPrepArgs(u3d_raster *r, SegmentPrep *p)
{
    const int tiles = r->geo.tilesX * r->geo.tilesY;
    U3D_TRY(r->segStart.ensure(size_t(tiles) + 1));
    U3D_TRY(r->segTotal.ensure(4));
    U3D_TRY(r->segCounter.ensure(1));
    // items <= segments of all lists (one per kSegEntries positions and tile) + 1
    U3D_TRY(r->itemTile.ensure(size_t(r->numPairs) / kSegEntries + size_t(tiles) + 1));
    p->segCount = r->segCount.p;
    p->perTile = kSegCountsPerTile;
    p->tiles = tiles;
    p->segStart = r->segStart.p;
    p->totalItems = r->segTotal.p;
    p->itemTile = r->itemTile.p;
    p->counter = r->segCounter.p;
    return U3D_OK;
}

// consumeGrad2d: the chain rule zeroes the 2D gradients it reads (the trainer's step), so the
// next backward needs no clearing pass; otherwise they stay readable (u3d_raster_get_grad2d).
// prepDone: the work-item scan already ran (with the loss, segmentPrepArgs' buffers).
int rasterBackwardImpl(u3d_raster *r, const u3d_gaussians &m, const float *dl,
                       u3d_gaussians &grads, float *ndc, int *gradCount,
                       const ShCoefs *shp = nullptr, bool consumeGrad2d = false,
                       const LossFinalize *lossFinalize = nullptr,
                       const AdamStepArgs *adam = nullptr, bool prepDone = false)
{
    const ShCoefs sh = shp ? *shp : ShCoefs{};
    U3D_CHECK(r->haveForward, U3D_ERR_INVALID, "backward without a forward");
    U3D_CHECK(m.n == r->n && grads.n == r->n, U3D_ERR_INVALID, "model size differs from the last forward");
    const hipStream_t s = r->ctx->stream;
    const size_t g2 = size_t(r->n) * kGrad2dStride;
    const float *before = r->grad2d.p;
    U3D_TRY(r->grad2d.ensure(g2 > 0 ? g2 : 1));
    if (r->grad2d.p != before)
        r->grad2dZero = false;
    if (g2 && !r->grad2dZero)
        U3D_HIP(hipMemsetAsync(r->grad2d.p, 0, r->grad2d.cap * sizeof(float), s));
    const ScreenBuffers sb = r->screen();
    if (r->segmented) {
        SegmentPrep prep;
        U3D_TRY(segmentPrepArgs(r, &prep));
        if (!r->occSeg)
            r->occSeg = rasterSegmentedOccupancy();
        StageScope scope(r->ctx, U3D_STAGE_RASTER_BWD); // item scan + compositing kernel
        U3D_TRY(launchRasterSegmented(s, r->geo, r->st, r->segCount.p, r->segStart.p, r->segTotal.p,
                                      r->segCounter.p, r->itemTile.p, r->ctx->num_cus * r->occSeg,
                                      r->tileOffsets.p, r->tileList, sb, dl, r->finalT.p,
                                      r->nContrib.p, r->ckpt.p, r->grad2d.p, lossFinalize,
                                      prepDone, r->hitBits.p));
    } else {
        if (lossFinalize)
            U3D_TRY(launchLossFinalize(s, *lossFinalize));
        StageScope scope(r->ctx, U3D_STAGE_RASTER_BWD);
        U3D_TRY(launchRasterBackward(s, r->geo, r->st, r->tileOffsets.p, r->tileList, sb, dl,
                                     r->finalT.p, r->nContrib.p, r->grad2d.p));
    }
    StageScope scope(r->ctx, U3D_STAGE_CHAIN);
    if (adam) {
        // the trainer's step: chain rule and Adam in one pass, grads untouched
        U3D_TRY(launchChainAdam(s, m, r->cam, r->st, sb, r->grad2d.p, ndc, gradCount, sh, *adam));
        r->grad2dZero = true;
        return U3D_OK;
    }
    U3D_TRY(launchChainRule(s, m, r->cam, r->st, sb, r->grad2d.p, grads, ndc, gradCount, sh,
                            consumeGrad2d));
    r->grad2dZero = consumeGrad2d;
    return U3D_OK;
}

// Scratch views carved out of the context's shared scratch buffer.
int ctxScratch(u3d_ctx *ctx, size_t bytes, unsigned char **out)
{
    U3D_TRY(ctx->scratch.ensure(bytes));
    *out = ctx->scratch.p;
    return U3D_OK;
}

} // namespace
} // namespace u3d

// ---------------------------------------------------------------------------------- trainer
// ----------------------------------------------------------------------------------- C ABI
extern "C" {

const char *u3d_last_error(void) { return tError.c_str(); }
const char *u3d_version(void) { return "u3d_hip 0.1 (gfx950)"; }

int u3d_ctx_create(int device, u3d_ctx **out)
{
    U3D_CHECK(out != nullptr, U3D_ERR_INVALID, "null out");
    *out = nullptr;
    int count = 0;
    U3D_HIP(hipGetDeviceCount(&count));
    U3D_CHECK(device >= 0 && device < count, U3D_ERR_INVALID, "no such HIP device");
    U3D_HIP(hipSetDevice(device));
    std::unique_ptr<u3d_ctx> c(new u3d_ctx());
    c->device = device;
    U3D_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    U3D_HIP(hipGetDeviceProperties(&prop, device));
    c->num_cus = prop.multiProcessorCount;
    U3D_TRY(staticScratchAddress(device, &c->staticScratch));
    *out = c.release();
    return U3D_OK;
}

void u3d_ctx_destroy(u3d_ctx *ctx)
{
    if (!ctx)
        return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int u3d_ctx_synchronize(u3d_ctx *ctx)
{
    U3D_TRY(useDevice(ctx));
    U3D_HIP(hipStreamSynchronize(ctx->stream));
    return U3D_OK;
}

void *u3d_ctx_stream(u3d_ctx *ctx) { return ctx ? reinterpret_cast<void *>(ctx->stream) : nullptr; }

int u3d_malloc(u3d_ctx *ctx, size_t bytes, void **out)
{
    U3D_TRY(useDevice(ctx));
    U3D_HIP(hipMalloc(out, bytes > 0 ? bytes : 1));
    return U3D_OK;
}

int u3d_free(u3d_ctx *ctx, void *ptr)
{
    U3D_TRY(useDevice(ctx));
    if (ptr) {
        U3D_HIP(hipStreamSynchronize(ctx->stream));
        U3D_HIP(hipFree(ptr));
    }
    return U3D_OK;
}

int u3d_memcpy_h2d(u3d_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    U3D_TRY(useDevice(ctx));
    if (bytes) {
        U3D_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        U3D_HIP(hipStreamSynchronize(ctx->stream));
    }
    return U3D_OK;
}

int u3d_memcpy_d2h(u3d_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    U3D_TRY(useDevice(ctx));
    if (bytes) {
        U3D_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        U3D_HIP(hipStreamSynchronize(ctx->stream));
    }
    return U3D_OK;
}

int u3d_memset(u3d_ctx *ctx, void *dst, int value, size_t bytes)
{
    U3D_TRY(useDevice(ctx));
    if (bytes)
        U3D_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
    return U3D_OK;
}

// ------------------------------------------------------------------------------ rasterizer
int u3d_raster_create(u3d_ctx *ctx, u3d_raster **out)
{
    U3D_TRY(useDevice(ctx));
    U3D_CHECK(out != nullptr, U3D_ERR_INVALID, "null out");
    *out = new u3d_raster();
    (*out)->ctx = ctx;
    return U3D_OK;
}

void u3d_raster_destroy(u3d_raster *r)
{
    if (!r)
        return;
    (void)hipSetDevice(r->ctx->device);
    (void)hipStreamSynchronize(r->ctx->stream);
    delete r;
}

int u3d_raster_forward(u3d_raster *r, const u3d_gaussians *model, const u3d_camera *camera,
                       const u3d_raster_settings *settings, float *rgb)
{
    U3D_CHECK(r && camera && settings, U3D_ERR_INVALID, "null argument");
    U3D_TRY(useDevice(r->ctx));
    U3D_TRY(checkModel(model, "model"));
    U3D_CHECK(rgb != nullptr || size_t(camera->width) * size_t(camera->height) == 0, U3D_ERR_INVALID, "null rgb");
    return rasterForwardImpl(r, *model, *camera, *settings, rgb);
}

int u3d_raster_backward(u3d_raster *r, const u3d_gaussians *model, const float *dl_drgb,
                        u3d_gaussians *grads, float *mean2d_grad_ndc)
{
    U3D_CHECK(r && dl_drgb, U3D_ERR_INVALID, "null argument");
    U3D_TRY(useDevice(r->ctx));
    U3D_TRY(checkModel(model, "model"));
    U3D_TRY(checkModel(grads, "grads"));
    return rasterBackwardImpl(r, *model, dl_drgb, *grads, mean2d_grad_ndc, nullptr);
}

int u3d_raster_forward_sh(u3d_raster *r, const u3d_gaussians *model, const float *sh_rest,
                          int sh_degree, const u3d_camera *camera,
                          const u3d_raster_settings *settings, float *rgb)
{
    U3D_CHECK(r && camera && settings, U3D_ERR_INVALID, "null argument");
    U3D_TRY(useDevice(r->ctx));
    U3D_TRY(checkModel(model, "model"));
    U3D_CHECK(rgb != nullptr || size_t(camera->width) * size_t(camera->height) == 0, U3D_ERR_INVALID, "null rgb");
    // the caller's coefficients are [n][15][3]
    ShCoefs sh;
    sh.coef = const_cast<float *>(sh_rest);
    sh.degree = sh_rest ? sh_degree : 0;
    return rasterForwardImpl(r, *model, *camera, *settings, rgb, &sh);
}

int u3d_raster_backward_sh(u3d_raster *r, const u3d_gaussians *model, const float *sh_rest,
                           int sh_degree, const float *dl_drgb, u3d_gaussians *grads,
                           float *grad_sh_rest, float *mean2d_grad_ndc)
{
    U3D_CHECK(r && dl_drgb, U3D_ERR_INVALID, "null argument");
    U3D_CHECK(sh_degree >= 0 && sh_degree <= 3, U3D_ERR_INVALID, "sh_degree must be 0..3");
    U3D_TRY(useDevice(r->ctx));
    U3D_TRY(checkModel(model, "model"));
    U3D_TRY(checkModel(grads, "grads"));
    ShCoefs sh;
    sh.coef = const_cast<float *>(sh_rest);
    sh.grad = grad_sh_rest;
    sh.degree = sh_rest ? sh_degree : 0;
    return rasterBackwardImpl(r, *model, dl_drgb, *grads, mean2d_grad_ndc, nullptr, &sh);
}

int u3d_raster_radii(u3d_raster *r, const int **radii, int64_t *n)
{
    U3D_CHECK(r && radii && n, U3D_ERR_INVALID, "null argument");
    *radii = r->radius.p;
    *n = r->haveForward ? r->n : 0;
    return U3D_OK;
}

int u3d_raster_get_state(u3d_raster *r, u3d_raster_state *out)
{
    U3D_CHECK(r && out, U3D_ERR_INVALID, "null argument");
    U3D_CHECK(r->haveForward, U3D_ERR_INVALID, "no forward pass yet");
    U3D_TRY(useDevice(r->ctx));
    const size_t nn = size_t(r->n > 0 ? r->n : 1);
    U3D_TRY(r->dbgMean2d.ensure(nn * 2));
    U3D_TRY(r->dbgConic.ensure(nn * 3));
    U3D_TRY(r->dbgColor.ensure(nn * 3));
    U3D_TRY(r->dbgOpacity.ensure(nn));
    U3D_TRY(launchUnpackScreen(r->ctx->stream, r->n, r->screen(), r->dbgMean2d.p, r->dbgConic.p,
                               r->dbgColor.p, r->dbgOpacity.p));
    U3D_HIP(hipStreamSynchronize(r->ctx->stream));
    if (r->sortScratch.p) {
        uint32_t sortError = 0;
        U3D_TRY(onesweepErrorWord(r->sortScratch.p, r->ctx->stream, &sortError));
        U3D_CHECK(sortError == 0, U3D_ERR_HIP,
                  "depth sort: a lookback wait was abandoned (the order is not valid)");
    }
    out->n = r->n;
    out->num_pairs = r->numPairs;
    out->tiles_x = r->geo.tilesX;
    out->tiles_y = r->geo.tilesY;
    out->radius = r->radius.p;
    out->rect = reinterpret_cast<const int *>(r->rect.p);
    out->depth = r->depth.p;
    out->mean2d = r->dbgMean2d.p;
    out->conic = r->dbgConic.p;
    out->color = r->dbgColor.p;
    out->opacity = r->dbgOpacity.p;
    out->flags = r->flags.p;
    out->tile_offsets = r->tileOffsets.p;
    out->tile_list = r->tileList;
    out->final_t = r->finalT.p;
    out->n_contrib = r->nContrib.p;
    return U3D_OK;
}

int u3d_raster_get_grad2d(u3d_raster *r, const float **grad2d)
{
    U3D_CHECK(r && grad2d, U3D_ERR_INVALID, "null argument");
    *grad2d = r->grad2d.p;
    return U3D_OK;
}

// ------------------------------------------------------------------------------------ loss
int u3d_photometric_loss(u3d_ctx *ctx, const float *render, const float *target, int width,
                         int height, float ssim_lambda, float *grad, u3d_loss_value *out_host)
{
    U3D_TRY(useDevice(ctx));
    U3D_CHECK(width >= 0 && height >= 0, U3D_ERR_INVALID, "negative image size");
    U3D_CHECK((render && target && grad) || size_t(width) * size_t(height) == 0, U3D_ERR_INVALID, "null image");
    // [partials: doubles][result: 4 floats][SSIM partial maps: 9 floats per pixel], 256-byte
    // aligned parts
    const size_t np = lossPartialsSize(width, height) + 1;
    const size_t resultOffset = (np * sizeof(double) + 255) / 256 * 256;
    const size_t mapsOffset = resultOffset + 256;
    unsigned char *sc = nullptr;
    U3D_TRY(ctxScratch(ctx, mapsOffset + lossMapsSize(width, height) * sizeof(float), &sc));
    double *partials = reinterpret_cast<double *>(sc);
    float *result = reinterpret_cast<float *>(sc + resultOffset);
    float *maps = reinterpret_cast<float *>(sc + mapsOffset);
    U3D_TRY(launchPhotometricLoss(ctx->stream, render, target, nullptr, width, height, ssim_lambda,
                                  grad, partials, np, result, nullptr, nullptr, maps));
    float host[4];
    U3D_HIP(hipMemcpyAsync(host, result, sizeof host, hipMemcpyDeviceToHost, ctx->stream));
    U3D_HIP(hipStreamSynchronize(ctx->stream));
    if (out_host)
        *out_host = {host[0], host[1], host[2]};
    return U3D_OK;
}

int u3d_photometric_loss_noalloc(u3d_ctx *ctx, const float *render, const float *target, int width,
                                 int height, float ssim_lambda, float *grad, float *value_device)
{
    U3D_TRY(useDevice(ctx));
    U3D_CHECK(width >= 0 && height >= 0, U3D_ERR_INVALID, "negative image size");
    U3D_CHECK((render && target && grad) || size_t(width) * size_t(height) == 0, U3D_ERR_INVALID, "null image");
    U3D_CHECK(ctx->staticScratch != nullptr, U3D_ERR_HIP, "static scratch unavailable");
    // [partials: doubles][result: 4 floats, 256-byte aligned]
    const size_t np = lossPartialsSize(width, height) + 1;
    const size_t resultOffset = (np * sizeof(double) + 255) / 256 * 256;
    U3D_CHECK(resultOffset + 4 * sizeof(float) <= kStaticScratchBytes, U3D_ERR_CAPACITY,
              "image too large for the static loss scratch (" + std::to_string(kStaticScratchBytes) +
                  " bytes: at most " + std::to_string((kStaticScratchBytes - 256) / sizeof(double) / 9 - 1) +
                  " blocks of 32x32 pixels)");
    double *partials = reinterpret_cast<double *>(ctx->staticScratch);
    float *result = value_device ? value_device
                                 : reinterpret_cast<float *>(ctx->staticScratch + resultOffset);
    return launchPhotometricLoss(ctx->stream, render, target, nullptr, width, height, ssim_lambda,
                                 grad, partials, np, result);
}

int u3d_psnr(u3d_ctx *ctx, const float *a, const float *b, int64_t count, float *out_host)
{
    U3D_TRY(useDevice(ctx));
    U3D_CHECK(out_host != nullptr, U3D_ERR_INVALID, "null out");
    const size_t np = psnrPartialsSize(count);
    unsigned char *sc = nullptr;
    U3D_TRY(ctxScratch(ctx, np * sizeof(double) + 64, &sc));
    float *result = reinterpret_cast<float *>(sc + np * sizeof(double));
    U3D_TRY(launchPsnr(ctx->stream, a, b, nullptr, count, reinterpret_cast<double *>(sc), np, result));
    U3D_HIP(hipMemcpyAsync(out_host, result, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    U3D_HIP(hipStreamSynchronize(ctx->stream));
    return U3D_OK;
}

// ------------------------------------------------------------------------------------ adam
int u3d_adam_create(u3d_ctx *ctx, u3d_adam **out)
{
    U3D_TRY(useDevice(ctx));
    U3D_CHECK(out != nullptr, U3D_ERR_INVALID, "null out");
    *out = new u3d_adam();
    (*out)->ctx = ctx;
    return U3D_OK;
}

void u3d_adam_destroy(u3d_adam *a)
{
    if (!a)
        return;
    (void)hipSetDevice(a->ctx->device);
    (void)hipStreamSynchronize(a->ctx->stream);
    delete a;
}

int u3d_adam_reset(u3d_adam *a, int64_t n)
{
    U3D_CHECK(a && n >= 0, U3D_ERR_INVALID, "bad argument");
    U3D_TRY(useDevice(a->ctx));
    U3D_TRY(a->m.reserve(n > 0 ? n : 1));
    U3D_TRY(a->v.reserve(n > 0 ? n : 1));
    a->m.n = a->v.n = n;
    U3D_TRY(a->m.zero(a->ctx->stream));
    U3D_TRY(a->v.zero(a->ctx->stream));
    a->step = 0;
    return U3D_OK;
}

namespace {

// Compacts `src` by keep (device, n entries) into `dst`, appending `extraZeros` zeroed entries.
int compactModel(u3d_ctx *ctx, const DevModel &src, DevModel &dst, const uint8_t *keep,
                 int64_t extraZeros, int64_t *newN)
{
    const hipStream_t s = ctx->stream;
    const int64_t n = src.n;
    DevBuf<uint32_t> flag, pos, counts, scratch;
    const size_t nn = size_t(n > 0 ? n : 1);
    U3D_TRY(flag.ensure(nn));
    U3D_TRY(pos.ensure(nn));
    U3D_TRY(counts.ensure(4));
    U3D_TRY(scratch.ensure(scanScratchSize(int64_t(nn))));
    DevBuf<uint8_t> noKind;
    U3D_TRY(noKind.ensure(nn));
    U3D_HIP(hipMemsetAsync(noKind.p, 0, nn, s));
    DevBuf<uint32_t> extra, split;
    U3D_TRY(extra.ensure(nn));
    U3D_TRY(split.ensure(nn));
    U3D_TRY(launchDensifyFlags(s, n, keep, noKind.p, flag.p, extra.p, split.p));
    U3D_TRY(exclusiveScanU32(s, flag.p, pos.p, n, counts.p + 2, scratch.p, scratch.cap));
    uint32_t host[4] = {uint32_t(extraZeros), 0, 0, 0};
    U3D_HIP(hipMemcpyAsync(&host[2], counts.p + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    U3D_HIP(hipStreamSynchronize(s));
    const int64_t total = int64_t(host[2]) + extraZeros;
    U3D_TRY(dst.reserve(std::max<int64_t>(std::max<int64_t>(total, src.cap), 1)));
    U3D_HIP(hipMemcpyAsync(counts.p, host, sizeof host, hipMemcpyHostToDevice, s));
    DevModel &d = dst;
    d.n = total;
    u3d_gaussians sv = src.view(), dv = d.view();
    // tail zeros: compactZeroTail zeroes [nKeep, nKeep + clones + 2*splits); extras given as clones
    const int64_t grid = std::max<int64_t>(n, extraZeros);
    U3D_TRY(launchCompactZeroTail(s, sv, dv, grid, keep, pos.p, counts.p));
    U3D_HIP(hipStreamSynchronize(s));
    *newN = total;
    return U3D_OK;
}

} // namespace

int u3d_adam_compact(u3d_adam *a, const uint8_t *keep)
{
    U3D_CHECK(a && (keep || a->m.n == 0), U3D_ERR_INVALID, "bad argument");
    U3D_TRY(useDevice(a->ctx));
    int64_t nn = 0;
    U3D_TRY(compactModel(a->ctx, a->m, a->spare, keep, 0, &nn));
    a->m.swap(a->spare);
    U3D_TRY(compactModel(a->ctx, a->v, a->spare, keep, 0, &nn));
    a->v.swap(a->spare);
    return U3D_OK;
}

int u3d_adam_grow(u3d_adam *a, int64_t count)
{
    U3D_CHECK(a && count >= 0, U3D_ERR_INVALID, "bad argument");
    U3D_TRY(useDevice(a->ctx));
    const hipStream_t s = a->ctx->stream;
    for (DevModel *mm : {&a->m, &a->v}) {
        const int64_t old = mm->n, nn = old + count;
        if (nn > mm->cap) {
            DevModel bigger;
            U3D_TRY(bigger.reserve(nn));
            U3D_TRY(bigger.zero(s));
            bigger.n = nn;
            U3D_TRY(copyFields(s, mm->view(), bigger.view(), old, hipMemcpyDeviceToDevice));
            U3D_HIP(hipStreamSynchronize(s));
            mm->swap(bigger);
        } else {
            mm->n = nn;
            const u3d_gaussians v = mm->view();
            const size_t f = sizeof(float);
            if (count) {
                U3D_HIP(hipMemsetAsync(v.means + old * 3, 0, size_t(count) * 3 * f, s));
                U3D_HIP(hipMemsetAsync(v.scales + old * 3, 0, size_t(count) * 3 * f, s));
                U3D_HIP(hipMemsetAsync(v.quats + old * 4, 0, size_t(count) * 4 * f, s));
                U3D_HIP(hipMemsetAsync(v.opacities + old, 0, size_t(count) * f, s));
                U3D_HIP(hipMemsetAsync(v.colors + old * 3, 0, size_t(count) * 3 * f, s));
            }
        }
    }
    return U3D_OK;
}

int u3d_adam_reset_opacity_moments(u3d_adam *a)
{
    U3D_CHECK(a != nullptr, U3D_ERR_INVALID, "null adam");
    U3D_TRY(useDevice(a->ctx));
    if (a->m.n > 0) {
        U3D_HIP(hipMemsetAsync(a->m.view().opacities, 0, size_t(a->m.n) * sizeof(float), a->ctx->stream));
        U3D_HIP(hipMemsetAsync(a->v.view().opacities, 0, size_t(a->v.n) * sizeof(float), a->ctx->stream));
    }
    return U3D_OK;
}

} // extern "C"

namespace u3d {
void adamCoefficients(int step, const u3d_learning_rates &lr, float stepSize[5], float *invSqrtBc2)
{
    // Adam.cpp:16-17, 70-72 (host powf, as the reference).
    const float bc1 = 1.0f - std::pow(0.9f, float(step));
    const float bc2 = 1.0f - std::pow(0.999f, float(step));
    const float rates[5] = {lr.means, lr.scales, lr.quats, lr.opacities, lr.colors};
    for (int k = 0; k < 5; ++k)
        stepSize[k] = rates[k] / bc1;
    *invSqrtBc2 = 1.0f / std::sqrt(bc2);
}
} // namespace u3d

extern "C" {

int u3d_adam_step(u3d_adam *a, u3d_gaussians *params, const u3d_gaussians *grads,
                  const u3d_learning_rates *lr)
{
    U3D_CHECK(a && lr, U3D_ERR_INVALID, "null argument");
    U3D_TRY(useDevice(a->ctx));
    U3D_TRY(checkModel(params, "params"));
    U3D_TRY(checkModel(grads, "grads"));
    U3D_CHECK(params->n == a->m.n && grads->n == a->m.n, U3D_ERR_INVALID, "size differs from the moments");
    ++a->step;
    float stepSize[5], inv;
    adamCoefficients(a->step, *lr, stepSize, &inv);
    u3d_gaussians m = a->m.view(), v = a->v.view();
    return launchAdamStep(a->ctx->stream, *params, *grads, m, v, stepSize, inv);
}

int u3d_adam_moments(u3d_adam *a, u3d_gaussians *m, u3d_gaussians *v, int *step)
{
    U3D_CHECK(a != nullptr, U3D_ERR_INVALID, "null adam");
    if (m)
        *m = a->m.view();
    if (v)
        *v = a->v.view();
    if (step)
        *step = a->step;
    return U3D_OK;
}

// ------------------------------------------------------------------------------ densification
namespace {

struct DensifyScratch {
    DevBuf<uint32_t> demand, scan, keepFlag, extraCount, splitFlag, keepPos, extraPos, splitPos;
    DevBuf<uint32_t> counts, misc, scanScratch;
    int ensure(int64_t n)
    {
        const size_t nn = size_t(n > 0 ? n : 1);
        for (DevBuf<uint32_t> *b : {&demand, &scan, &keepFlag, &extraCount, &splitFlag, &keepPos,
                                    &extraPos, &splitPos})
            U3D_TRY(b->ensure(nn));
        U3D_TRY(counts.ensure(8));
        U3D_TRY(misc.ensure(8));
        U3D_TRY(scanScratch.ensure(scanScratchSize(int64_t(nn))));
        return U3D_OK;
    }
};

int densifyDecideImpl(u3d_ctx *ctx, DensifyScratch &ds, const u3d_gaussians &m, const float *accum,
                      const int *count, float extent, int maxG, float thr, float pd, float po,
                      int wasReset, uint8_t *keep, uint8_t *kind, uint32_t hostCounts[4])
{
    const hipStream_t s = ctx->stream;
    const int64_t n = m.n;
    U3D_TRY(ds.ensure(n));
    U3D_TRY(launchDensifyDemand(s, m, accum, count, extent, thr, pd, po, wasReset, keep, ds.demand.p));
    U3D_TRY(exclusiveScanU32(s, ds.demand.p, ds.scan.p, n, ds.misc.p + 4, ds.scanScratch.p,
                             ds.scanScratch.cap));
    const uint64_t budget = int64_t(maxG) > n ? uint64_t(int64_t(maxG) - n) : 0; // Trainer.cpp:109
    U3D_TRY(launchDensifyAccept(s, n, ds.demand.p, ds.scan.p, budget, keep, kind, ds.misc.p, ds.counts.p));
    U3D_HIP(hipMemcpyAsync(hostCounts, ds.counts.p, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    U3D_HIP(hipStreamSynchronize(s));
    return U3D_OK;
}

// Scatter into dst (capacity checked by the caller); counts must be on the device.
int densifyApplyImpl(u3d_ctx *ctx, DensifyScratch &ds, const u3d_gaussians &src,
                     u3d_gaussians &dst, const uint8_t *keep, const uint8_t *kind,
                     const float *normals, float logShrink)
{
    const hipStream_t s = ctx->stream;
    const int64_t n = src.n;
    U3D_TRY(ds.ensure(n));
    U3D_TRY(launchDensifyFlags(s, n, keep, kind, ds.keepFlag.p, ds.extraCount.p, ds.splitFlag.p));
    U3D_TRY(exclusiveScanU32(s, ds.keepFlag.p, ds.keepPos.p, n, ds.misc.p + 4, ds.scanScratch.p, ds.scanScratch.cap));
    U3D_TRY(exclusiveScanU32(s, ds.extraCount.p, ds.extraPos.p, n, ds.misc.p + 5, ds.scanScratch.p, ds.scanScratch.cap));
    U3D_TRY(exclusiveScanU32(s, ds.splitFlag.p, ds.splitPos.p, n, ds.misc.p + 6, ds.scanScratch.p, ds.scanScratch.cap));
    U3D_TRY(launchDensifyScatter(s, src, dst, n, keep, kind, ds.keepPos.p, ds.extraPos.p,
                                 ds.splitPos.p, ds.counts.p, normals, logShrink));
    return U3D_OK;
}

} // namespace

int u3d_densify_decide(u3d_ctx *ctx, const u3d_gaussians *model, const float *grad_accum,
                       const int *grad_count, float extent, int max_gaussians,
                       float grad_threshold, float percent_dense, float prune_opacity,
                       int opacity_was_reset, uint8_t *keep, uint8_t *kind,
                       int64_t *n_clone_host, int64_t *n_split_host, int64_t *n_keep_host)
{
    U3D_TRY(useDevice(ctx));
    U3D_TRY(checkModel(model, "model"));
    U3D_CHECK(model->n == 0 || (grad_accum && grad_count && keep && kind), U3D_ERR_INVALID, "null argument");
    DensifyScratch ds;
    uint32_t c[4];
    U3D_TRY(densifyDecideImpl(ctx, ds, *model, grad_accum, grad_count, extent, max_gaussians,
                              grad_threshold, percent_dense, prune_opacity, opacity_was_reset, keep,
                              kind, c));
    if (n_clone_host) *n_clone_host = c[0];
    if (n_split_host) *n_split_host = c[1];
    if (n_keep_host) *n_keep_host = c[2];
    return U3D_OK;
}

int u3d_densify_apply(u3d_ctx *ctx, u3d_gaussians *model, const uint8_t *keep,
                      const uint8_t *kind, const float *normals, float log_shrink,
                      int64_t *new_n_host)
{
    U3D_TRY(useDevice(ctx));
    U3D_TRY(checkModel(model, "model"));
    const hipStream_t s = ctx->stream;
    const int64_t n = model->n;
    DensifyScratch ds;
    U3D_TRY(ds.ensure(n));
    // counts from keep / kind
    U3D_HIP(hipMemsetAsync(ds.counts.p, 0, 4 * sizeof(uint32_t), s));
    U3D_HIP(hipMemsetAsync(ds.misc.p, 0xff, 2 * sizeof(uint32_t), s));
    // densifyCount (through launchDensifyAccept's counting kernel) needs kind writable; use a copy
    DevBuf<uint8_t> kindCopy;
    U3D_TRY(kindCopy.ensure(size_t(n > 0 ? n : 1)));
    if (n > 0)
        U3D_HIP(hipMemcpyAsync(kindCopy.p, kind, size_t(n), hipMemcpyDeviceToDevice, s));
    U3D_TRY(launchDensifyFlags(s, n, keep, kindCopy.p, ds.keepFlag.p, ds.extraCount.p, ds.splitFlag.p));
    uint32_t tot[3] = {0, 0, 0};
    U3D_TRY(exclusiveScanU32(s, ds.keepFlag.p, ds.keepPos.p, n, ds.misc.p + 4, ds.scanScratch.p, ds.scanScratch.cap));
    U3D_TRY(exclusiveScanU32(s, ds.extraCount.p, ds.extraPos.p, n, ds.misc.p + 5, ds.scanScratch.p, ds.scanScratch.cap));
    U3D_TRY(exclusiveScanU32(s, ds.splitFlag.p, ds.splitPos.p, n, ds.misc.p + 6, ds.scanScratch.p, ds.scanScratch.cap));
    U3D_HIP(hipMemcpyAsync(tot, ds.misc.p + 4, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    U3D_HIP(hipStreamSynchronize(s));
    const uint32_t splits = tot[2], clones = tot[1] - 2 * tot[2], kept = tot[0];
    const uint32_t counts[4] = {clones, splits, kept, 0};
    U3D_HIP(hipMemcpyAsync(ds.counts.p, counts, sizeof counts, hipMemcpyHostToDevice, s));
    const int64_t total = int64_t(kept) + int64_t(tot[1]);
    DevModel tmp;
    U3D_TRY(tmp.reserve(std::max<int64_t>(total, 1)));
    tmp.n = total;
    u3d_gaussians dv = tmp.view();
    U3D_TRY(launchDensifyScatter(s, *model, dv, n, keep, kindCopy.p, ds.keepPos.p, ds.extraPos.p,
                                 ds.splitPos.p, ds.counts.p, normals, log_shrink));
    U3D_TRY(copyFields(s, dv, *model, total, hipMemcpyDeviceToDevice));
    U3D_HIP(hipStreamSynchronize(s));
    model->n = total;
    if (new_n_host)
        *new_n_host = total;
    return U3D_OK;
}

// -------------------------------------------------------------------------------- trainer
int u3d_trainer_create(u3d_ctx *ctx, const float *means, const float *scales, const float *quats,
                       const float *opacities, const float *colors, int64_t n,
                       const u3d_camera *cameras, const uint8_t *rgb8, int n_images,
                       float scene_extent, const u3d_train_options *options, u3d_trainer **out)
{
    U3D_TRY(useDevice(ctx));
    U3D_CHECK(out && options && n >= 0 && n_images >= 0, U3D_ERR_INVALID, "bad argument");
    U3D_CHECK(n == 0 || (means && scales && quats && opacities && colors), U3D_ERR_INVALID, "null model");
    U3D_CHECK(n_images == 0 || (cameras && rgb8), U3D_ERR_INVALID, "null images");
    U3D_CHECK(options->raster.tile_size > 0, U3D_ERR_INVALID, "tile_size must be positive");
    U3D_CHECK(options->raster.tile_size % 16 == 0 || (options->views_per_step <= 1 && !options->use_graph),
              U3D_ERR_UNSUPPORTED, "the batched engine (views_per_step > 1 or use_graph) needs a tile_size "
              "that is a multiple of 16");
    std::unique_ptr<u3d_trainer> t(new u3d_trainer());
    t->ctx = ctx;
    t->raster.ctx = ctx;
    t->opt = *options;
    t->extent = std::max(scene_extent, 1e-6f); // Trainer.cpp:41
    t->rng.seed(options->seed);
    const hipStream_t s = ctx->stream;
    const int64_t cap = std::max<int64_t>(std::max<int64_t>(n, options->max_gaussians), 1);
    for (DevModel *m : {&t->params, &t->spare, &t->grads, &t->adamM, &t->adamV}) {
        U3D_TRY(m->reserve(cap));
        m->n = n;
    }
    U3D_TRY(t->adamM.zero(s));
    U3D_TRY(t->adamV.zero(s));
    const u3d_gaussians host{const_cast<float *>(means), const_cast<float *>(scales),
                             const_cast<float *>(quats), const_cast<float *>(opacities),
                             const_cast<float *>(colors), n};
    U3D_TRY(copyFields(s, host, t->params.view(), n, hipMemcpyHostToDevice));
    size_t total = 0, maxPix = 1;
    for (int k = 0; k < n_images; ++k) {
        U3D_CHECK(cameras[k].width >= 0 && cameras[k].height >= 0, U3D_ERR_INVALID, "negative image size");
        t->cams.push_back(cameras[k]);
        t->imageOffset.push_back(total);
        const size_t p = size_t(cameras[k].width) * size_t(cameras[k].height);
        total += p * 3;
        maxPix = std::max(maxPix, p);
    }
    U3D_TRY(t->targets.ensure(total > 0 ? total : 1));
    if (total)
        U3D_HIP(hipMemcpyAsync(t->targets.p, rgb8, total, hipMemcpyHostToDevice, s));
    U3D_TRY(t->render.ensure(maxPix * 3));
    U3D_TRY(t->dloss.ensure(maxPix * 3));
    U3D_TRY(t->ssimMaps.ensure(maxPix * 9));
    U3D_TRY(t->gradAccum.ensure(size_t(cap)));
    U3D_TRY(t->gradCount.ensure(size_t(cap)));
    U3D_TRY(t->evalPairs.ensure(64));
    {
        const int ts = options->raster.tile_size;
        size_t costs = 0;
        for (const u3d_camera &c : t->cams) {
            t->walkCostOffset.push_back(costs);
            costs += size_t((c.width + ts - 1) / ts) * size_t((c.height + ts - 1) / ts);
        }
        U3D_TRY(t->walkCost.ensure(costs > 0 ? costs : 1));
        U3D_HIP(hipMemsetAsync(t->walkCost.p, 0, t->walkCost.cap * sizeof(uint32_t), s));
    }
    U3D_HIP(hipMemsetAsync(t->evalPairs.p, 0, 64 * sizeof(unsigned long long), s));
    U3D_HIP(hipMemsetAsync(t->gradAccum.p, 0, size_t(cap) * sizeof(float), s));
    U3D_HIP(hipMemsetAsync(t->gradCount.p, 0, size_t(cap) * sizeof(int), s));
    U3D_CHECK(options->sh_degree >= 0 && options->sh_degree <= 3, U3D_ERR_INVALID, "sh_degree must be 0..3");
    U3D_CHECK(options->views_per_step >= 0 && options->views_per_step <= kMaxViewsPerStep,
              U3D_ERR_INVALID, "views_per_step must be 0..64");
    t->viewsPerStep = std::max(1, options->views_per_step);
    t->useGraph = options->use_graph != 0;
    t->shDegree = options->sh_degree;
    if (t->shDegree > 0) {
        for (DevBuf<float> *b : {&t->shP, &t->shG, &t->shM, &t->shV, &t->shSpare}) {
            U3D_TRY(b->ensure(size_t(cap) * 45));
            U3D_HIP(hipMemsetAsync(b->p, 0, b->cap * sizeof(float), s));
        }
    }
    {
        const size_t slots = size_t(std::max(1, options->iterations)) * size_t(t->viewsPerStep);
        U3D_TRY(t->stepStats.ensure(slots * 4 * sizeof(float)));
        uint32_t *w = static_cast<uint32_t *>(t->stepStats.p);
        for (size_t k = 0; k < slots * 4; ++k)
            w[k] = kStatsPending;
    }
    U3D_TRY(t->evalOut.ensure(size_t(std::max(1, n_images)) + 4));
    size_t parts = 1024;
    for (const u3d_camera &c : t->cams)
        parts = std::max(parts, lossPartialsSize(c.width, c.height) + 1);
    U3D_TRY(t->partials.ensure(parts));
    U3D_HIP(hipStreamSynchronize(s));
    t->stats.gaussians = n;
    if (t->viewsPerStep > 1 || t->useGraph)
        U3D_TRY(createBatchEngine(ctx, t->useGraph, &t->engine));
    *out = t.release();
    return U3D_OK;
}

void u3d_trainer_destroy(u3d_trainer *t)
{
    if (!t)
        return;
    (void)hipSetDevice(t->ctx->device);
    (void)hipStreamSynchronize(t->ctx->stream);
    destroyBatchEngine(t->engine);
    delete t;
}

int u3d_trainer_done(u3d_trainer *t) { return t && t->iteration >= t->opt.iterations ? 1 : 0; }
int u3d_trainer_iteration(u3d_trainer *t) { return t ? t->iteration : 0; }
int64_t u3d_trainer_size(u3d_trainer *t) { return t ? t->params.n : 0; }

namespace {

int trainerDensify(u3d_trainer *t)
{
    // Trainer::densifyAndPrune, Trainer.cpp:105-169.
    u3d_ctx *ctx = t->ctx;
    const hipStream_t s = ctx->stream;
    const int64_t n = t->params.n;
    DensifyScratch ds;
    DevBuf<uint8_t> keep, kind;
    U3D_TRY(keep.ensure(size_t(n > 0 ? n : 1)));
    U3D_TRY(kind.ensure(size_t(n > 0 ? n : 1)));
    uint32_t c[4];
    const u3d_gaussians pv = t->params.view();
    U3D_TRY(densifyDecideImpl(ctx, ds, pv, t->gradAccum.p, t->gradCount.p, t->extent,
                              t->opt.max_gaussians, t->opt.densify_grad_threshold,
                              t->opt.percent_dense, t->opt.prune_opacity, t->opacityWasReset ? 1 : 0,
                              keep.p, kind.p, c));
    // Split offsets: 6 draws per accepted split from a fresh N(0, 1) over the trainer's
    // generator, in index order (Trainer.cpp:114, 150).
    std::normal_distribution<float> normal(0.0f, 1.0f);
    std::vector<float> normals(size_t(c[1]) * 6);
    for (float &x : normals)
        x = normal(t->rng);
    DevBuf<float> dn;
    U3D_TRY(dn.ensure(normals.size() > 0 ? normals.size() : 1));
    if (!normals.empty())
        U3D_HIP(hipMemcpyAsync(dn.p, normals.data(), normals.size() * sizeof(float), hipMemcpyHostToDevice, s));
    volatile float shrinkArg = 1.6f;
    const float logShrink = std::log(float(shrinkArg));
    const int64_t newN = int64_t(c[2]) + int64_t(c[0]) + 2 * int64_t(c[1]);
    U3D_CHECK(newN <= t->params.cap, U3D_ERR_CAPACITY, "densification exceeds the model capacity");
    t->spare.n = newN;
    u3d_gaussians sv = t->spare.view();
    U3D_TRY(densifyApplyImpl(ctx, ds, pv, sv, keep.p, kind.p, dn.p, logShrink));
    t->params.swap(t->spare);
    t->params.n = newN;
    if (t->shDegree > 0) {
        // coefficient planes of stride n -> planes of stride newN
        U3D_TRY(launchScatterPlanes(s, t->shP.p, t->shSpare.p, 45, n, newN, keep.p, kind.p,
                                    ds.keepPos.p, ds.extraPos.p, ds.counts.p, false));
        std::swap(t->shP.p, t->shSpare.p);
        std::swap(t->shP.cap, t->shSpare.cap);
        for (DevBuf<float> *mom : {&t->shM, &t->shV}) {
            U3D_TRY(launchScatterPlanes(s, mom->p, t->shSpare.p, 45, n, newN, keep.p, kind.p,
                                        ds.keepPos.p, ds.extraPos.p, ds.counts.p, true));
            std::swap(mom->p, t->shSpare.p);
            std::swap(mom->cap, t->shSpare.cap);
        }
    }
    // Adam moments: compacted, zeros appended (Adam.cpp:40-60).
    for (DevModel *mm : {&t->adamM, &t->adamV}) {
        t->spare.n = newN;
        u3d_gaussians src = mm->view(), dst = t->spare.view();
        src.n = n;
        U3D_TRY(launchCompactZeroTail(s, src, dst, n, keep.p, ds.keepPos.p, ds.counts.p));
        mm->swap(t->spare);
        mm->n = newN;
    }
    t->grads.n = newN;
    t->spare.n = 0;
    // Statistics restart (Trainer.cpp:167-168).
    if (newN > 0) {
        U3D_HIP(hipMemsetAsync(t->gradAccum.p, 0, size_t(newN) * sizeof(float), s));
        U3D_HIP(hipMemsetAsync(t->gradCount.p, 0, size_t(newN) * sizeof(int), s));
    }
    U3D_HIP(hipStreamSynchronize(s));
    return U3D_OK;
}

int trainerResetOpacity(u3d_trainer *t)
{
    // Trainer::resetOpacity, Trainer.cpp:171-178.
    const hipStream_t s = t->ctx->stream;
    volatile float p = 0.01f;
    const float pf = p;
    const float cap = std::log(pf / (1.0f - pf)); // logit(0.01f), GaussianModel.h:12
    const int64_t n = t->params.n;
    U3D_TRY(launchOpacityCap(s, t->params.view().opacities, n, cap));
    if (n > 0) {
        U3D_HIP(hipMemsetAsync(t->adamM.view().opacities, 0, size_t(n) * sizeof(float), s));
        U3D_HIP(hipMemsetAsync(t->adamV.view().opacities, 0, size_t(n) * sizeof(float), s));
    }
    t->opacityWasReset = true;
    return U3D_OK;
}

} // namespace

} // extern "C"

namespace u3d {

int trainerNextView(u3d_trainer *t)
{
    if (t->orderPos >= t->order.size()) {
        t->order.resize(t->cams.size());
        for (size_t k = 0; k < t->order.size(); ++k)
            t->order[k] = int(k);
        std::shuffle(t->order.begin(), t->order.end(), t->rng);
        t->orderPos = 0;
    }
    return t->order[t->orderPos++];
}

float trainerMeansLr(const u3d_trainer *t)
{
    const float tt = float(t->iteration) / float(std::max(1, t->opt.iterations - 1));
    return std::exp(std::log(t->opt.lr.means) * (1 - tt) + std::log(t->opt.means_lr_final) * tt) * t->extent;
}

int trainerFinishStep(u3d_trainer *t)
{
    ++t->iteration;
    const u3d_train_options &o = t->opt;
    const bool densify = t->iteration >= o.refine_start && t->iteration < o.refine_stop
                         && o.refine_interval > 0 && t->iteration % o.refine_interval == 0;
    const bool reset = o.opacity_reset_interval > 0 && t->iteration < o.refine_stop
                       && t->iteration % o.opacity_reset_interval == 0;
    if (densify || reset) {
        StageScope scope(t->ctx, U3D_STAGE_DENSIFY);
        if (densify)
            U3D_TRY(trainerDensify(t));
        if (reset)
            U3D_TRY(trainerResetOpacity(t));
    }
    return U3D_OK;
}

} // namespace u3d

extern "C" {

int u3d_trainer_step(u3d_trainer *t)
{
    // Trainer::step, Trainer.cpp:53-103.
    U3D_CHECK(t != nullptr, U3D_ERR_INVALID, "null trainer");
    U3D_TRY(useDevice(t->ctx));
    if (t->cams.empty() || u3d_trainer_done(t))
        return U3D_OK;
    if (t->engine) {
        u3d_trainer *one[1] = {t};
        return batchTrainStep(t->engine, one, 1);
    }
    const size_t view = size_t(trainerNextView(t));
    const u3d_camera &cam = t->cams[view];
    const hipStream_t s = t->ctx->stream;
    const u3d_gaussians pv = t->params.view();
    const ShCoefs shv = trainerSh(t);
    U3D_TRY(rasterForwardImpl(&t->raster, pv, cam, t->opt.raster, t->render.p, &shv, t->evalPairs.p,
                              t->walkCost.p + t->walkCostOffset[view]));
    t->counters[0] += 1;
    t->counters[1] += t->params.n;
    t->counters[2] += int64_t(cam.width) * cam.height;
    t->counters[3] += t->raster.numPairs;
    t->counters[4] += 1;
    float *stats = t->statsSlot(size_t(t->iteration));
    // The loss value's final reduction rides on the backward's first launch, and the
    // segmented backward's work-item scan on the loss launch.
    LossFinalize lossFinalize;
    SegmentPrep prep;
    const bool prepWithLoss = t->raster.segmented;
    if (prepWithLoss)
        U3D_TRY(segmentPrepArgs(&t->raster, &prep));
    {
        StageScope scope(t->ctx, U3D_STAGE_LOSS);
        U3D_TRY(launchPhotometricLoss(s, t->render.p, nullptr, t->targets.p + t->imageOffset[view],
                                      cam.width, cam.height, t->opt.ssim_lambda, t->dloss.p,
                                      t->partials.p, t->partials.cap, stats, &lossFinalize,
                                      prepWithLoss ? &prep : nullptr, t->ssimMaps.p));
    }
    t->grads.n = t->params.n;
    u3d_gaussians gv = t->grads.view();
    // Exponentially decaying position rate scaled by the scene size (Trainer.cpp:82-84).
    u3d_learning_rates lr = t->opt.lr;
    lr.means = trainerMeansLr(t);
    ++t->adamStep;
    AdamStepArgs adam;
    adamCoefficients(t->adamStep, lr, adam.stepSize, &adam.invSqrtBc2);
    adam.m = t->adamM.view();
    adam.v = t->adamV.view();
    // Backward, then the chain rule and Adam in one pass: the 14 parameter gradients of a
    // Gaussian stay in registers (the gradient buffer is not used by this step), and its SH
    // coefficients take their update in the same pass (rate lr.colors * scale, Adam.cpp:70-72).
    ShCoefs shAdam = shv;
    if (t->shDegree > 0) {
        const float bc1 = 1.0f - std::pow(0.9f, float(t->adamStep));
        shAdam.m = t->shM.p;
        shAdam.v = t->shV.p;
        shAdam.stepSize = (lr.colors * shRateScale(t)) / bc1;
        shAdam.invSqrtBc2 = adam.invSqrtBc2;
    }
    U3D_TRY(rasterBackwardImpl(&t->raster, pv, t->dloss.p, gv, t->gradAccum.p, t->gradCount.p,
                               &shAdam, true, &lossFinalize, &adam, prepWithLoss));
    return trainerFinishStep(t);
}

int u3d_trainer_get_model(u3d_trainer *t, float *means, float *scales, float *quats,
                          float *opacities, float *colors)
{
    U3D_CHECK(t != nullptr, U3D_ERR_INVALID, "null trainer");
    U3D_TRY(useDevice(t->ctx));
    const int64_t n = t->params.n;
    U3D_CHECK(n == 0 || (means && scales && quats && opacities && colors), U3D_ERR_INVALID, "null output");
    const u3d_gaussians host{means, scales, quats, opacities, colors, n};
    U3D_TRY(copyFields(t->ctx->stream, t->params.view(), host, n, hipMemcpyDeviceToHost));
    U3D_HIP(hipStreamSynchronize(t->ctx->stream));
    return U3D_OK;
}

int u3d_trainer_get_moments(u3d_trainer *t, float *m_host, float *v_host)
{
    U3D_CHECK(t != nullptr, U3D_ERR_INVALID, "null trainer");
    U3D_TRY(useDevice(t->ctx));
    const int64_t n = t->params.n;
    U3D_CHECK(n == 0 || (m_host && v_host), U3D_ERR_INVALID, "null output");
    const DevModel *src[2] = {&t->adamM, &t->adamV};
    float *dst[2] = {m_host, v_host};
    for (int k = 0; k < 2; ++k) {
        float *h = dst[k];
        const u3d_gaussians hv{h, h + 3 * n, h + 6 * n, h + 10 * n, h + 11 * n, n};
        u3d_gaussians dv = src[k]->view();
        dv.n = n;
        U3D_TRY(copyFields(t->ctx->stream, dv, hv, n, hipMemcpyDeviceToHost));
    }
    U3D_HIP(hipStreamSynchronize(t->ctx->stream));
    return U3D_OK;
}

} // extern "C"

namespace {
// The trainer keeps coefficient c of Gaussian i at plane[c * n + i]; the API's layout is
// [n][15][3] (row i, column c).
int shPlanesToRows(u3d_trainer *t, const float *planes, float *rows)
{
    const size_t n = size_t(t->params.n);
    std::vector<float> tmp(n * 45);
    U3D_HIP(hipMemcpyAsync(tmp.data(), planes, tmp.size() * sizeof(float), hipMemcpyDeviceToHost,
                           t->ctx->stream));
    U3D_HIP(hipStreamSynchronize(t->ctx->stream));
    for (size_t c = 0; c < 45; ++c)
        for (size_t i = 0; i < n; ++i)
            rows[i * 45 + c] = tmp[c * n + i];
    return U3D_OK;
}
} // namespace

extern "C" {

int u3d_trainer_get_sh_rest(u3d_trainer *t, float *sh_rest_host)
{
    U3D_CHECK(t != nullptr, U3D_ERR_INVALID, "null trainer");
    U3D_CHECK(t->shDegree > 0, U3D_ERR_INVALID, "trainer has no SH coefficients (sh_degree 0)");
    U3D_TRY(useDevice(t->ctx));
    if (t->params.n == 0)
        return U3D_OK;
    U3D_CHECK(sh_rest_host != nullptr, U3D_ERR_INVALID, "null output");
    return shPlanesToRows(t, t->shP.p, sh_rest_host);
}

int u3d_trainer_set_sh_rest(u3d_trainer *t, const float *sh_rest_host)
{
    U3D_CHECK(t != nullptr, U3D_ERR_INVALID, "null trainer");
    U3D_CHECK(t->shDegree > 0, U3D_ERR_INVALID, "trainer has no SH coefficients (sh_degree 0)");
    U3D_TRY(useDevice(t->ctx));
    const size_t n = size_t(t->params.n);
    if (n == 0)
        return U3D_OK;
    U3D_CHECK(sh_rest_host != nullptr, U3D_ERR_INVALID, "null input");
    std::vector<float> planes(n * 45);
    for (size_t c = 0; c < 45; ++c)
        for (size_t i = 0; i < n; ++i)
            planes[c * n + i] = sh_rest_host[i * 45 + c];
    U3D_HIP(hipMemcpyAsync(t->shP.p, planes.data(), planes.size() * sizeof(float),
                           hipMemcpyHostToDevice, t->ctx->stream));
    U3D_HIP(hipStreamSynchronize(t->ctx->stream));
    return U3D_OK;
}

int u3d_trainer_get_sh_moments(u3d_trainer *t, float *m_host, float *v_host)
{
    U3D_CHECK(t != nullptr, U3D_ERR_INVALID, "null trainer");
    U3D_CHECK(t->shDegree > 0, U3D_ERR_INVALID, "trainer has no SH coefficients (sh_degree 0)");
    U3D_TRY(useDevice(t->ctx));
    if (t->params.n == 0)
        return U3D_OK;
    U3D_CHECK(m_host && v_host, U3D_ERR_INVALID, "null output");
    U3D_TRY(shPlanesToRows(t, t->shM.p, m_host));
    return shPlanesToRows(t, t->shV.p, v_host);
}

int u3d_trainer_get_densify_stats(u3d_trainer *t, float *accum_host, int *count_host)
{
    U3D_CHECK(t != nullptr, U3D_ERR_INVALID, "null trainer");
    U3D_TRY(useDevice(t->ctx));
    const size_t n = size_t(t->params.n);
    if (n) {
        U3D_CHECK(accum_host && count_host, U3D_ERR_INVALID, "null output");
        U3D_HIP(hipMemcpyAsync(accum_host, t->gradAccum.p, n * sizeof(float), hipMemcpyDeviceToHost,
                               t->ctx->stream));
        U3D_HIP(hipMemcpyAsync(count_host, t->gradCount.p, n * sizeof(int), hipMemcpyDeviceToHost,
                               t->ctx->stream));
        U3D_HIP(hipStreamSynchronize(t->ctx->stream));
    }
    return U3D_OK;
}

/* writeSplatPly(trainer.model()) (Reconstructor.cpp:361): one copy from the device, then the
 * host writer of ply.cpp. */
int u3d_trainer_export_ply(u3d_trainer *t, const char *path)
{
    U3D_CHECK(t != nullptr && path != nullptr, U3D_ERR_INVALID, "null argument");
    const int64_t n = t->params.n;
    std::vector<float> h(size_t(n) * 14), sh(t->shDegree > 0 ? size_t(n) * 45 : 0);
    float *b = h.data();
    U3D_TRY(u3d_trainer_get_model(t, b, b + 3 * n, b + 6 * n, b + 10 * n, b + 11 * n));
    if (t->shDegree > 0)
        U3D_TRY(u3d_trainer_get_sh_rest(t, sh.data()));
    return u3d_write_splat_ply(path, b, b + 3 * n, b + 6 * n, b + 10 * n, b + 11 * n, n,
                               t->shDegree > 0 ? sh.data() : nullptr, t->shDegree);
}

int u3d_trainer_stats(u3d_trainer *t, u3d_train_stats *out)
{
    // TrainStats EMA (Trainer.cpp:88-92), folded lazily from the per-step device values.
    U3D_CHECK(t && out, U3D_ERR_INVALID, "null argument");
    U3D_TRY(useDevice(t->ctx));
    // With several views per step, a step's value is the mean over its views. The values are in
    // pinned host memory; waiting for the last step's psnr word (stored by the backward's first
    // launch) lets the device go on with the rest of the step: no stream synchronisation.
    const int from = t->statsFolded, to = t->iteration;
    const size_t B = size_t(t->viewsPerStep);
    if (to > from) {
        for (int it = from; it < to; ++it)
            for (size_t k = 0; k < B; ++k)
                U3D_TRY(waitPinnedU32(reinterpret_cast<const uint32_t *>(t->statsSlot(size_t(it) * B + k) + 3),
                                      kStatsPending, t->ctx->stream));
        for (int it = from; it < to; ++it) {
            float loss = 0.0f, psnr = 0.0f;
            for (size_t k = 0; k < B; ++k) {
                const float *x = t->statsSlot(size_t(it) * B + k);
                loss += x[0];
                psnr += x[3];
            }
            if (B > 1) {
                loss /= float(B);
                psnr /= float(B);
            }
            const float k = t->stats.iteration == 0 ? 1.0f : 0.05f;
            t->stats.loss += k * (loss - t->stats.loss);
            t->stats.psnr += k * (psnr - t->stats.psnr);
            t->stats.iteration = it + 1;
        }
        t->statsFolded = to;
    }
    t->stats.gaussians = t->params.n;
    *out = t->stats;
    return U3D_OK;
}

namespace {
int evaluateAll(u3d_trainer *t, int mode, float *out)
{
    U3D_TRY(useDevice(t->ctx));
    if (t->cams.empty()) {
        *out = 0.0f;
        return U3D_OK;
    }
    const hipStream_t s = t->ctx->stream;
    const u3d_gaussians pv = t->params.view();
    for (size_t k = 0; k < t->cams.size(); ++k) {
        const u3d_camera &cam = t->cams[k];
        const ShCoefs shv = trainerSh(t);
        U3D_TRY(rasterForwardImpl(&t->raster, pv, cam, t->opt.raster, t->render.p, &shv));
        const int64_t count = int64_t(cam.width) * cam.height * 3;
        if (mode == 0)
            U3D_TRY(launchPsnr(s, t->render.p, nullptr, t->targets.p + t->imageOffset[k], count,
                               t->partials.p, t->partials.cap, t->evalOut.p + k));
        else
            U3D_TRY(launchMeanAbs(s, t->render.p, t->targets.p + t->imageOffset[k], count,
                                  t->partials.p, t->partials.cap, t->evalOut.p + k));
    }
    std::vector<float> v(t->cams.size());
    U3D_HIP(hipMemcpyAsync(v.data(), t->evalOut.p, v.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    U3D_HIP(hipStreamSynchronize(s));
    double sum = 0;
    for (float x : v)
        sum += x;
    *out = float(sum / double(v.size()));
    return U3D_OK;
}
} // namespace

int u3d_trainer_evaluate_psnr(u3d_trainer *t, float *out)
{
    U3D_CHECK(t && out, U3D_ERR_INVALID, "null argument");
    return evaluateAll(t, 0, out);
}

int u3d_trainer_evaluate_l1(u3d_trainer *t, float *out)
{
    U3D_CHECK(t && out, U3D_ERR_INVALID, "null argument");
    return evaluateAll(t, 1, out);
}

int u3d_trainer_render(u3d_trainer *t, const u3d_camera *camera, float *rgb_host)
{
    U3D_CHECK(t && camera && rgb_host, U3D_ERR_INVALID, "null argument");
    U3D_TRY(useDevice(t->ctx));
    const size_t count = size_t(camera->width) * size_t(camera->height) * 3;
    DevBuf<float> out;
    U3D_TRY(out.ensure(count > 0 ? count : 1));
    const ShCoefs shv = trainerSh(t);
    U3D_TRY(rasterForwardImpl(&t->raster, t->params.view(), *camera, t->opt.raster, out.p, &shv));
    U3D_HIP(hipMemcpyAsync(rgb_host, out.p, count * sizeof(float), hipMemcpyDeviceToHost, t->ctx->stream));
    U3D_HIP(hipStreamSynchronize(t->ctx->stream));
    return U3D_OK;
}

int u3d_ctx_set_option(u3d_ctx *ctx, int option, int value)
{
    U3D_TRY(useDevice(ctx));
    U3D_HIP(hipStreamSynchronize(ctx->stream));
    switch (option) {
    case U3D_OPT_SEGMENTED_BACKWARD:
        ctx->segmentedBackward = value != 0;
        return U3D_OK;
    case U3D_OPT_SORT_SPIN_LIMIT:
        ctx->sortSpinLimit = value;
        return U3D_OK;
    default:
        U3D_CHECK(false, U3D_ERR_INVALID, "unknown option");
    }
    return U3D_OK;
}

int u3d_ctx_set_timing(u3d_ctx *ctx, int enabled)
{
    U3D_TRY(useDevice(ctx));
    U3D_HIP(hipStreamSynchronize(ctx->stream));
    ctx->timer.collect();
    for (int k = 0; k < U3D_NUM_STAGES; ++k) {
        ctx->timer.totalMs[k] = 0.0;
        ctx->timer.count[k] = 0;
    }
    ctx->timer.enabled = enabled != 0;
    return U3D_OK;
}

int u3d_ctx_stage_times(u3d_ctx *ctx, double total_ms[U3D_NUM_STAGES], int64_t counts[U3D_NUM_STAGES])
{
    U3D_TRY(useDevice(ctx));
    ctx->timer.collect();
    for (int k = 0; k < U3D_NUM_STAGES; ++k) {
        if (total_ms)
            total_ms[k] = ctx->timer.totalMs[k];
        if (counts)
            counts[k] = ctx->timer.count[k];
    }
    return U3D_OK;
}

int u3d_trainer_counters(u3d_trainer *t, int64_t values[U3D_NUM_COUNTERS])
{
    U3D_CHECK(t && values, U3D_ERR_INVALID, "null argument");
    U3D_TRY(useDevice(t->ctx));
    unsigned long long part[64] = {};
    if (t->evalPairs.p) {
        U3D_HIP(hipMemcpyAsync(part, t->evalPairs.p, sizeof part, hipMemcpyDeviceToHost, t->ctx->stream));
        U3D_HIP(hipStreamSynchronize(t->ctx->stream));
    }
    unsigned long long evals = 0;
    for (unsigned long long p : part)
        evals += p;
    t->counters[U3D_NUM_COUNTERS - 1] = int64_t(evals);
    for (int k = 0; k < U3D_NUM_COUNTERS; ++k)
        values[k] = t->counters[k];
    return U3D_OK;
}

void u3d_expf_host_array(const float *x, float *y, int64_t n)
{
    static const uint64_t table[32] = U3D_EXP2_TABLE;
    for (int64_t i = 0; i < n; ++i) {
        bool special;
        const float sp = u3dmath::expf_special(x[i], &special);
        y[i] = special ? sp : u3dmath::expf_core(x[i], table);
    }
}

} // extern "C"
