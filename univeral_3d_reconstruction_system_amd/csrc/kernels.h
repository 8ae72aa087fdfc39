// Host-side launchers of the device kernels (one .hip translation unit each). All launch on
// the given stream and return U3D_OK or an error code; none of them synchronises.
#pragma once
#include "common.h"

namespace u3d {

// Per-Gaussian screen-space state written by the projection kernel. rec0/rec1/rec2 pack what
// the compositing kernels read per (tile, Gaussian) pair: {mx, my, A, B}, {C, alpha, r, g}, b.
struct ScreenBuffers {
    float4 *rec0;
    float4 *rec1;
    float *rec2;
    float *depth;
    int *radius;
    int4 *rect;
    uint8_t *flags;
    uint32_t *depthKey; // depth bits, 0xffffffff when not visible
    uint32_t *nTiles;   // tiles covered, 0 when not visible
};

// One (scene, view) render of a batch: several views of one scene and/or views of several
// scenes share one projection pass, one sort and one persistent compositing launch. Screen
// entries, tile ids and raster blocks of all items are concatenated; each item records its
// offsets into them and the scene arrays it reads and writes.
struct RenderItem {
    u3d_camera cam;
    int64_t screenBase; // first screen entry (Gaussian i of the item is entry screenBase + i)
    int64_t n;          // Gaussians of the scene
    int64_t blockBase;  // first global 16x16 raster block
    int64_t pixBase;    // first pixel of the item in the concatenated per-pixel buffers
    int tileBase, tilesX, tilesY, blocksX;
    // the item's tile offsets are tileOffsets[tileBase + offsetShift + t], t in [0, tiles]
    // (offsetShift = item index when every item's offsets are stored with their own end)
    int offsetShift;
    int shDegree;
    float ndcScale;     // densification statistic per view: undoes the 1/B of the averaged loss
    const float *means, *scales, *quats, *opacities, *colors;
    float *shRest; // coefficient planes of the scene (stride n), or null
    float *rgb, *finalT;
    int *nContrib;
    const float *dl;
    float *gMeans, *gScales, *gQuats, *gOpac, *gColors, *gShRest, *ndc;
    int *gradCount;
};

// One Adam update of a scalar parameter (Adam.cpp:13-28): beta1 0.9, beta2 0.999, and 1e-15
// added to sqrt(v) * invSqrtBc2; stepSize is the learning rate over 1 - beta1^t.
__device__ __forceinline__ void adamScalar(float &p, float &m, float &v, float g, float stepSize,
                                           float invSqrtBc2)
{
    const float mk = 0.9f * m + (1.0f - 0.9f) * g;
    const float vk = 0.999f * v + (1.0f - 0.999f) * g * g;
    m = mk;
    v = vk;
    p = p - stepSize * mk / (sqrtf(vk) * invSqrtBc2 + 1e-15f);
}

// Final reduction of the photometric loss (Loss.cpp:180-185): the per-block double partials
// (l1, ssim, squared error) summed in a fixed order, result = {total, l1, ssim, psnr}.
struct LossFinalize {
    const double *partials = nullptr; // numBlocks x 3
    int numBlocks = 0;
    float lambda = 0.0f, invCount = 0.0f;
    double count = 0.0;
    float *result = nullptr;
};

// The reduction on the calling workgroup (a multiple of 64 threads, at least 256): the first 256
// threads accumulate, so the summation order does not depend on the workgroup size.
__device__ __forceinline__ void lossFinalizeBlock(const LossFinalize &f)
{
    __shared__ double red[3][16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double v[3] = {0.0, 0.0, 0.0};
    if (tid < 256) {
        for (int i = tid; i < f.numBlocks; i += 256) {
            v[0] += f.partials[size_t(i) * 3];
            v[1] += f.partials[size_t(i) * 3 + 1];
            v[2] += f.partials[size_t(i) * 3 + 2];
        }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            v[j] += __shfl_xor(v[j], off, 64);
    }
    if (lane == 0 && wave < 16) {
#pragma unroll
        for (int j = 0; j < 3; ++j)
            red[j][wave] = v[j];
    }
    __syncthreads();
    if (tid != 0)
        return;
    const int waves = int(blockDim.x) / 64 < 16 ? int(blockDim.x) / 64 : 16;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int w = 0; w < waves; ++w) {
        a += red[0][w];
        b += red[1][w];
        c += red[2][w];
    }
    const float l1 = float(a * double(f.invCount));
    float ssim = 0.0f, total = l1;
    if (f.lambda > 0.0f) {
        ssim = float(b * double(f.invCount));
        total = (1.0f - f.lambda) * l1 + f.lambda * (1.0f - ssim);
    }
    const double mse = c / (f.count > 1.0 ? f.count : 1.0);
    f.result[0] = total;
    f.result[1] = l1;
    f.result[2] = ssim;
    // psnr last, released at system scope: a host that reads it from pinned memory (the
    // trainer's per-step statistics) also sees the three values before it
    const float psnr = mse <= 1e-12 ? 99.0f : float(-10.0 * log10(mse));
    __hip_atomic_store(reinterpret_cast<unsigned *>(&f.result[3]), __float_as_uint(psnr),
                       __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Work-item scan of the segmented backward: segStart = exclusive scan of the per-tile segment
// counts (tiles + 1 entries), itemTile maps each item to its tile, and totalItems and the
// claim counter are set for the backward's workgroups.
struct SegmentPrep {
    const uint32_t *segCount = nullptr;
    int perTile = 1; // segment counts per tile (their maximum is the tile's)
    int tiles = 0;
    uint32_t *segStart = nullptr, *totalItems = nullptr;
    int *itemTile = nullptr;
    unsigned *counter = nullptr;
};

// The scan on the calling workgroup (a multiple of 64 threads, at most 1024).
__device__ __forceinline__ void segmentPrepBlock(const SegmentPrep &p)
{
    __shared__ uint32_t carry[16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nt = int(blockDim.x), nw = nt / 64;
    uint32_t base = 0;
    for (int start = 0; start < p.tiles; start += nt) {
        const int t = start + tid;
        uint32_t v = 0;
        if (t < p.tiles)
            for (int q = 0; q < p.perTile; ++q)
                v = max(v, p.segCount[size_t(t) * size_t(p.perTile) + size_t(q)]);
        uint32_t x = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= off)
                x += y;
        }
        if (lane == 63)
            carry[wave] = x;
        __syncthreads();
        uint32_t offs = base, chunk = 0;
        for (int w = 0; w < nw; ++w) {
            offs += w < wave ? carry[w] : 0u;
            chunk += carry[w];
        }
        offs += x - v;
        if (t < p.tiles) {
            p.segStart[t] = offs;
            for (uint32_t k = 0; k < v; ++k)
                p.itemTile[offs + k] = t;
        }
        base += chunk;
        __syncthreads();
    }
    if (tid == 0) {
        p.segStart[p.tiles] = base;
        *p.totalItems = base;
        *p.counter = 0u;
    }
}

// Loss of one item of a batched step (target = uint8 / 255, HWC).
struct LossItem {
    const float *render;
    const uint8_t *target;
    float *grad;       // dL/drgb of the item, scaled by gradScale
    float *result;     // {total, l1, ssim, psnr}
    double *partials;  // 9 doubles per 32x32 block
    float *maps;       // SSIM partials, 9 floats per pixel (the split loss; null: fused)
    int width, height, blocksX, numBlocks;
    float invCount, gradScale;
};

// One scene's optimiser step in a batched step. Parameter-shaped buffers use the packed
// layout [means 3c | scales 3c | quats 4c | opacities c | colours 3c] for capacity c.
struct SceneStep {
    float *params, *grads, *adamM, *adamV;
    int64_t n, cap;
    float stepSize[5];
    float invSqrtBc2;
    float *shP, *shG, *shM, *shV; // 45 coefficient planes of n floats when shDegree > 0
    float shStepSize;
    int shDegree;
    float beta1, beta2;           // Adam decay rates (0.9, 0.999 unless batch-scaled)
};

// ---- kernels_project.hip
// Spherical-harmonic colour beyond DC, degree 1..3 (an extension). Coefficient c (0..44,
// c = 3 k + colour channel, basis function k) of Gaussian i is coef[i * gs + c * cs]: the API
// layout [n][15][3] is gs = 45, cs = 1; the trainer keeps one plane per coefficient, gs = 1 and
// cs = its current size n, so the lanes of a wavefront read consecutive addresses.
struct ShCoefs {
    float *coef = nullptr;   // read (written only by the fused Adam)
    float *grad = nullptr;   // gradient in the same layout, added to (nullable)
    int64_t gs = 45, cs = 1;
    int degree = 0;          // 0: no view-dependent colour
    // The trainer's fused update: Adam moments in the same layout (nullable: none). Every
    // coefficient of every Gaussian is updated, with a zero gradient where it has none.
    float *m = nullptr, *v = nullptr;
    float stepSize = 0.0f, invSqrtBc2 = 0.0f;
};
// blockTiles (optional, projectBlocks(n) entries): per projection block, the tiles covered.
int launchProject(hipStream_t s, const u3d_gaussians &m, const u3d_camera &cam,
                  const u3d_raster_settings &st, int tilesX, int tilesY, const ScreenBuffers &out,
                  const ShCoefs &sh, uint32_t *blockTiles = nullptr);
int projectBlocks(int64_t n);
// Chain rule of the backward pass (Rasterizer.cpp:432-544): per visible Gaussian, ADDS into
// grads / ndc. gradCount (nullable) is incremented where radius > 0 (Trainer.cpp:76-79).
// clearGrad2d: zero the consumed 2D gradients (the buffer is then all zero again).
// Adam coefficients of one step for the five parameter groups (means, scales, quats,
// opacities, colours) and the model-shaped first and second moments.
struct AdamStepArgs {
    u3d_gaussians m, v;
    float stepSize[5];
    float invSqrtBc2;
};
// The trainer's update after the backward: the chain rule of each Gaussian into registers and
// its Adam update from them (Adam runs for every Gaussian, visible or not), so the parameter
// gradients never reach memory; the 2D gradients are cleared after use. Same arithmetic as
// launchChainRule into zeroed gradients followed by launchAdamStep.
// sh: with m / v set, the SH coefficients take their Adam update in the same pass.
int launchChainAdam(hipStream_t s, const u3d_gaussians &params, const u3d_camera &cam,
                    const u3d_raster_settings &st, const ScreenBuffers &sb, float *grad2d,
                    float *ndc, int *gradCount, const ShCoefs &sh, const AdamStepArgs &adam);
int launchChainRule(hipStream_t s, const u3d_gaussians &m, const u3d_camera &cam,
                    const u3d_raster_settings &st, const ScreenBuffers &sb, float *grad2d,
                    u3d_gaussians &grads, float *ndc, int *gradCount, const ShCoefs &sh,
                    bool clearGrad2d = false);
// Batched projection over items (grid y = item); entryItem[e] = owning item of screen entry e.
int launchProjectBatched(hipStream_t s, const RenderItem *items, int numItems, int64_t maxN,
                         const u3d_raster_settings &st, const ScreenBuffers &out, int *entryItem);
// Batched chain rule. groups[g] = {first item, item count} of scene g (its items are
// consecutive); each thread adds the contributions of all items of its scene in item order
// (deterministic, no atomics) and clears the 2D gradients it consumed.
int launchChainBatched(hipStream_t s, const RenderItem *items, const int2 *groups, int numGroups,
                       int64_t maxN, const u3d_raster_settings &st, const ScreenBuffers &sb,
                       float *grad2d);
// Batched chain rule and Adam in one pass, for groups with no view-dependent colour: the views of
// a scene run on separate wavefronts (up to four per block of 64 Gaussians, maxViews = the
// largest group), their gradient sums are added in wavefront order and the update is
// adamBatchedKernel's; the gradient arrays are not written.
int launchChainAdamBatched(hipStream_t s, const RenderItem *items, const int2 *groups,
                           const SceneStep *scenes, int numGroups, int64_t maxN, int maxViews,
                           const u3d_raster_settings &st, const ScreenBuffers &sb, float *grad2d);
// Unpacks the records into the separate views of u3d_raster_state.
int launchUnpackScreen(hipStream_t s, int64_t n, const ScreenBuffers &sb, float *mean2d,
                       float *conic, float *color, float *opacity);

// ---- kernels_sort.hip
// Stable LSD radix sort of (key, value) pairs on bits [0, bits) of the key, `count` elements
// (device pointer to the count when countDev != nullptr, else `count`). Ping-pongs between
// (keys, vals) and (keysAlt, valsAlt); returns in *resultInAlt where the output landed.
struct RadixScratch {
    uint32_t *hist;    // radixHistSize(count)
    size_t histCap;
    uint32_t *scanTmp; // radixScanScratchSize(count)
    size_t scanTmpCap;
    int items = 0;     // keys per lane per block: 4 or 8 (0: by size)
    int radixBits = 8; // digit bits per pass of the multi-launch sort (1..8)
    bool valsAreIndex = false; // the input values are 0..count-1 (vals is then not read)
    // Optional small-sort scratch (onesweepScratchWords() words, zero-filled once when
    // allocated and then left to the sort) and its host-side call counter: counts up to 2^20
    // with no device-side count sort in one launch per 8-bit pass plus one histogram launch.
    // The launches carry the counter's value, so they must not be replayed from a graph.
    uint64_t *onesweep = nullptr;
    size_t onesweepCap = 0;
    uint32_t *onesweepEpoch = nullptr;
    // Optional sum of sideIn[0, sideCount) run with the sort, as launchSumU32 would (to sideDev
    // and, if given, to the pinned word sideHost with sideSeq): an extra workgroup of the
    // histogram launch on the small-sort path, its own launch before the sort otherwise.
    const uint32_t *sideIn = nullptr;
    int64_t sideCount = 0;
    uint32_t *sideDev = nullptr;
    uint64_t *sideHost = nullptr;
    uint32_t sideSeq = 0;
    // Bound on a small-sort pass's lookback re-reads before it gives up and sets the error word
    // (0: the default, far beyond any wait that ends; negative: every wait is given up at once,
    // a test hook for the error path).
    int spinLimit = 0;
};
size_t onesweepScratchWords();
// The small sort's error word (non-zero once a lookback wait was abandoned): synchronous read,
// then cleared. A histogram launch with a side sum to a pinned word also reports and clears it
// (bit 63 of that word).
int onesweepErrorWord(uint64_t *scratch, hipStream_t s, uint32_t *value);
const uint32_t *onesweepErrorPtr(const uint64_t *scratch);
int radixSortPairs(hipStream_t s, uint32_t *keys, uint32_t *vals, uint32_t *keysAlt,
                   uint32_t *valsAlt, int64_t count, int bits, RadixScratch &scratch,
                   bool *resultInAlt, const uint32_t *countDev = nullptr);
size_t radixHistSize(int64_t count);
size_t radixScanScratchSize(int64_t count);
// values[i] = i
int launchIota(hipStream_t s, uint32_t *vals, int64_t count);
// out[i] = in[idx[i]] for i < count (0 where idx[i] >= inCount)
int launchGatherU32(hipStream_t s, const uint32_t *in, const uint32_t *idx, uint32_t *out,
                    int64_t count, int64_t inCount);
// Sum of count u32 values (one workgroup) to *devOut and, if given, to *hostOut (pinned host
// memory mapped for the device) as the word (seq << 32 | sum), published by one system-scope
// release store; waitPinnedWord() reads it without an event on the stream.
int launchSumU32(hipStream_t s, const uint32_t *in, int64_t count, uint32_t *devOut,
                 uint64_t *hostOut = nullptr, uint32_t seq = 0);
// Exclusive scan of count u32 values (out may alias in); total written to *total (device).
int exclusiveScanU32(hipStream_t s, const uint32_t *in, uint32_t *out, int64_t count,
                     uint32_t *total, uint32_t *blockScratch, size_t blockScratchCap);
size_t scanScratchSize(int64_t count);
// Tile binning of depth-ordered Gaussians (sorted[i] = Gaussian, invisible ones last): writes
// each tile's list of Gaussian ids in depth order to tileList (tileOffsets[tiles] entries,
// tileOffsets[0, tiles] exclusive) without materialising or sorting (tile, Gaussian) pairs.
// Supports tile grids with (tilesX + 1) (tilesY + 1) <= 12288 (one LDS grid cell per tile).
// counts: tileBinningCountsSize entries; tileTotal: tileBinningTotalsSize entries (the tile
// totals, then the totals of blocks of 32 tiles).
int tileBinningGroup(int64_t n, int tilesX, int tilesY); // Gaussians per group, 0: unsupported
size_t tileBinningCountsSize(int64_t n, int tilesX, int tilesY);
size_t tileBinningTotalsSize(int tilesX, int tilesY);
// ntS, rectS (n entries each): scratch for the tile counts and rectangles in depth order.
// heavyFirst (optional, tiles entries): the tiles by decreasing list length, written unless
// n <= 0. Ids sorted[i] >= n count as invisible; list slots and offsets stay below listCapacity.
// sortError (nullable): the small sort's error word (onesweepErrorPtr); when set, every
// Gaussian counts as invisible. walkCost (nullable, tiles entries): per-tile walk costs of the
// last forward of this view (0: unknown), which order heavyFirst; cleared for the next forward.
// Placement of one image's lists in a list shared by several images (the batched engine):
// ids in sorted are below idLimit; the lists start at listBase of tileList (the offsets written
// include it) and take at most listCapacity slots; listSum / listMax (atomically) receive the
// image's pair count and overflow is set when it exceeds listCapacity.
struct BinningList {
    int64_t idLimit;
    uint32_t listBase;
    uint32_t *listSum, *listMax, *overflow;
};
int launchTileBinning(hipStream_t s, int64_t n, const uint32_t *sorted, const uint32_t *nTiles,
                      const int4 *rect, uint32_t *ntS, int4 *rectS, int tilesX, int tilesY,
                      uint32_t *counts, size_t countsCap, uint32_t *tileTotal, int *tileOffsets,
                      uint32_t *tileList, size_t listCapacity, int *heavyFirst = nullptr,
                      const uint32_t *sortError = nullptr, uint32_t *walkCost = nullptr,
                      const BinningList *list = nullptr);
// The two halves of launchTileBinning: the per-(group, tile) counts and their scan, and the
// emission (ntS / rectS are the arrays the counts were taken from: the depth-ordered copies, or
// nTiles / rect themselves when sorted is null). sorted == nullptr bins the Gaussians in index
// order: each tile's list comes out in index order (for launchTileSort). longFirst (nullable,
// tiles entries): the tiles by decreasing list length.
// side (nullable): a sum of side->in[0, side->count) as launchSumU32 does, run by an extra first
// workgroup of the count launch (the pair count K from the projection's block sums, published
// to the host while the counts run).
struct SideSum {
    const uint32_t *in = nullptr;
    int64_t count = 0;
    uint32_t *dev = nullptr;
    uint64_t *host = nullptr;
    uint32_t seq = 0;
};
int launchTileCounts(hipStream_t s, int64_t n, const uint32_t *sorted, const uint32_t *nTiles,
                     const int4 *rect, uint32_t *ntS, int4 *rectS, int tilesX, int tilesY,
                     uint32_t *counts, size_t countsCap, uint32_t *tileTotal,
                     const uint32_t *sortError, int64_t idLimit, const SideSum *side = nullptr);
int launchTileEmit(hipStream_t s, int64_t n, const uint32_t *sorted, const uint32_t *ntS,
                   const int4 *rectS, int tilesX, int tilesY, const uint32_t *counts,
                   const uint32_t *tileTotal, int *tileOffsets, uint32_t *tileList,
                   size_t listCapacity, int *heavyFirst, uint32_t *walkCost,
                   const BinningList *list, int *longFirst);
// Sorts each tile's list (listIn at the tile's offsets, Gaussian ids in index order) stably on
// depthKey[id] into listOut: the (depth bits, index) order. order (nullable): a permutation of
// the tiles, sorted in that order (the longest first). scratch: tileSortScratchSize(capacity)
// words for lists longer than one workgroup holds.
size_t tileSortScratchSize(size_t listCapacity);
int launchTileSort(hipStream_t s, int tiles, const int *tileOffsets, const int *order,
                   const uint32_t *listIn, uint32_t *listOut, const uint32_t *depthKey,
                   uint32_t *scratch, size_t scratchCap, size_t listCapacity);
// Index-order binning and per-tile sort of several images in one launch per phase (the batched
// engine). Image v bins entries [idBase, idBase + n) (its Gaussians; ids written are global);
// its (group, tile) counters, tile and block totals, offsets (tiles + 1 entries, including
// listBase), list slice and longest-first order start at the given offsets of shared buffers.
// binItemsLayout fills the derived fields and returns the buffer sizes and grids.
struct BinItem {
    int64_t n;
    uint32_t idBase;
    int tilesX, tilesY, group, numGroups;
    uint32_t countsBase, totalsBase, offsetsBase, listBase, orderBase;
    int countBlock0, scanBlock0, emitBlock0, sortBlock0; // first workgroup in each launch
};
struct BinItemsGrid {
    size_t counts = 0, totals = 0, orders = 0; // buffer entries
    int countBlocks = 0, scanBlocks = 0, emitBlocks = 0, sortBlocks = 0;
    int maxCells = 0, emitWaves = 0;
};
// items[v].n, idBase, tilesX, tilesY, offsetsBase and listBase are set by the caller. sortMap
// (sortBlocks entries, item << 24 | rank) deals the images' tiles to the sort's workgroups in
// rounds: round k takes the k-th longest list of every image that has one, so the longest lists
// of all images start first.
int binItemsLayout(std::vector<BinItem> &items, BinItemsGrid *grid, std::vector<uint32_t> *sortMap);
// sorted (nullable): every image's entries in depth order (global ids below idLimit, image v's
// at sorted + idBase); the lists then come out in depth order (no per-tile sort needed) and
// ntS / rectS receive the gathered tile counts and rectangles. longFirst is nullable.
int launchTileBinningItems(hipStream_t s, const BinItem *items, int numItems,
                           const BinItemsGrid &grid, const uint32_t *nTiles, const int4 *rect,
                           uint32_t *counts, uint32_t *tileTotal, int *tileOffsets,
                           uint32_t *tileList, size_t capacity, uint32_t *listSum,
                           uint32_t *listMax, uint32_t *overflow, int *longFirst,
                           const uint32_t *sorted = nullptr, uint32_t *ntS = nullptr,
                           int4 *rectS = nullptr, int64_t idLimit = 0);
int launchTileSortItems(hipStream_t s, const BinItem *items, const uint32_t *sortMap,
                        const BinItemsGrid &grid, const int *tileOffsets, const int *longFirst,
                        const uint32_t *listIn, uint32_t *listOut, const uint32_t *depthKey,
                        uint32_t *scratch, size_t scratchCap, size_t listCapacity);
// Writes (tile id, Gaussian) pairs for Gaussians in depth order: sorted[i] is the Gaussian,
// offsets[i] its first pair slot; slots at or past capacity are not written.
int launchEmitPairs(hipStream_t s, int64_t n, const uint32_t *sorted, const uint32_t *offsets,
                    const int4 *rect, int tilesX, uint32_t *pairTile, uint32_t *pairGauss,
                    const uint32_t *nTilesSorted, int64_t capacity);
// Batched emission: tile ids are global (item tileBase + local tile), values are screen entries.
// Pairs at or past `capacity` are dropped (the farthest in depth order) and *overflow is set.
int launchEmitPairsBatched(hipStream_t s, int64_t n, const uint32_t *sorted,
                           const uint32_t *offsets, const int4 *rect, const int *entryItem,
                           const RenderItem *items, uint32_t *pairTile, uint32_t *pairGauss,
                           const uint32_t *nTilesSorted, int64_t capacity, unsigned *overflow);
// Work order for the persistent compositing launch: keys[b] = 8-bit inverted log-scale length
// of block b's tile list, order[b] = b; a one-pass (8-bit) radix sort then lists heavy blocks
// first.
int launchBlockCosts(hipStream_t s, const RenderItem *items, int numItems, int64_t totalBlocks,
                     int tileSize, const int *tileOffsets, uint32_t *keys, uint32_t *order);
int launchBlockCostsPlain(hipStream_t s, int blocksX, int totalBlocks, int tileSize, int tilesX,
                          const int *tileOffsets, uint32_t *keys, uint32_t *order);
// tileOffsets[t] = first pair index of tile t, tileOffsets[tiles] = K (tile ids sorted).
// With countDev, the pair count is min(*countDev, numPairs) read on the device.
int launchTileOffsets(hipStream_t s, const uint32_t *pairTile, int64_t numPairs, int tiles,
                      int *tileOffsets, const uint32_t *countDev = nullptr);

// ---- kernels_raster.hip
struct RasterGeometry {
    int width, height, tileSize, tilesX, tilesY;
};
// Forward state the segmented backward starts from (16x16 tiles; ckpt == nullptr: not kept).
// For a tile t whose list starts at pair `begin`, checkpoint s >= 1 of its pixel p,
// ckpt[((begin / B) + t + s) * 256 + p] (B = segment length: kSegEntries for the single-image
// forward, the LDS batch in the batched engine), holds the transmittance before list position
// s * B and the colour composited from that position on (the single-image forward writes it
// only for pixels still walking there, the only ones whose backward reads it); the segment
// count is ceil(max contributor count / B), per tile quadrant in the single-image forward
// (segCount[4 t + q]) and per tile in the batched engine. Capacity: (K / B + tiles + 1) * 256.
struct SegmentState {
    float4 *ckpt;
    uint32_t *segCount;
    // Optional (single-image forward): the cull of every 64-entry step of a tile list by
    // quadrant, hitBits[(begin / 64 + tile + step) * 4 + quadrant], bit i for entry
    // begin + 64 step + i; the segmented backward reads it instead of recomputing the cull.
    uint64_t *hitBits = nullptr;
};
// The single-image forward (16-pixel tiles) writes one segment count per tile quadrant.
constexpr int kSegCountsPerTile = 4;
// List entries per segment of the single-image segmented backward (one work item each; the
// forward checkpoints every kSegEntries entries).
constexpr int kSegEntries = 64;
// blockOrder (optional, 16-pixel tiles only): workgroup i renders tile blockOrder[i].
int launchRasterForward(hipStream_t s, const RasterGeometry &g, const u3d_raster_settings &st,
                        const int *tileOffsets, const int *tileList, const ScreenBuffers &sb,
                        float *rgb, float *finalT, int *nContrib, const SegmentState &seg,
                        const int *blockOrder = nullptr, unsigned long long *evalCount = nullptr,
                        uint32_t *walkCost = nullptr);
// Segmented backward: one work item per (tile, segment of kSegEntries list positions) from the
// forward's per-tile segment counts. Scratch written here: segStart (tiles + 1 entries, the
// exclusive scan of segCount), *totalItems, counter, itemTile (at least Σ segCount entries:
// the owning tile of each item).
int launchRasterSegmented(hipStream_t s, const RasterGeometry &g, const u3d_raster_settings &st,
                          const uint32_t *segCount, uint32_t *segStart, uint32_t *totalItems,
                          unsigned *counter, int *itemTile, int numWorkgroups,
                          const int *tileOffsets, const int *tileList, const ScreenBuffers &sb,
                          const float *dl, const float *finalT, const int *nContrib,
                          const float4 *ckpt, float *grad2d, const LossFinalize *fin = nullptr,
                          bool prepDone = false, const uint64_t *hitBits = nullptr);
int rasterSegmentedOccupancy();
int launchRasterBackward(hipStream_t s, const RasterGeometry &g, const u3d_raster_settings &st,
                         const int *tileOffsets, const int *tileList, const ScreenBuffers &sb,
                         const float *dlDrgb, const float *finalT, const int *nContrib,
                         float *grad2d);

// Persistent compositing over the blocks of all items in `order` (nullptr = natural order);
// `counter` must be zero on entry (one unsigned). ldsBatch: tile-list records staged in LDS per
// round (64, 128 or 256; the compositing order and results do not depend on it). fewLanes
// (backward; 0, 4 or 16): Gaussians with at most that many contributing lanes in a wave skip
// the wave reduction and add per lane.
int launchRasterBatched(hipStream_t s, bool backward, const RenderItem *items, int numItems,
                        const uint32_t *order, int64_t totalBlocks, unsigned *counter,
                        int numWorkgroups, const int *tileOffsets, const int *tileList,
                        const ScreenBuffers &sb, const u3d_raster_settings &st, float *grad2d,
                        int ldsBatch, int fewLanes, const SegmentState &seg);

// Resident workgroups per CU of the persistent compositing kernel (occupancy limit).
int rasterBatchedOccupancy(bool backward, int ldsBatch);

// ---- kernels_points.hip
// Hash grid of seed points: point indices sorted by their 64-bit cell key (hi, lo halves).
struct PointGrid {
    const uint32_t *keyHi;
    const uint32_t *keyLo;
    const uint32_t *idx;
};
constexpr int kFeatureNeighbours = 16;
constexpr int kFeatureDim = 80; // 16 x (offset, distance) + colour + view fraction, zero padded
constexpr int kBoundsBlocks = 1024;
// out6 = {min x, min y, min z, max x, max y, max z}; partial holds 6 * kBoundsBlocks floats.
int launchPointBounds(hipStream_t s, const float *xyz, int64_t n, float *partial, float *out6);
int launchCellKeys(hipStream_t s, const float *xyz, int64_t n, const float lo[3], float cell,
                   uint32_t *keyLo, uint32_t *keyHi, uint32_t *idx);
// initGaussiansFromPoints per point (GaussianModel.cpp:103-141) except the log-scales: writes
// means, quaternions, opacities, colours and each point's mean squared distance to its 3
// nearest neighbours (the caller turns those into log-scales with the host libm).
int launchKnnInit(hipStream_t s, const float *xyz, const float *rgb, int rgbCount, int64_t n,
                  const float lo[3], float cell, const PointGrid &g, float opacityLogit,
                  u3d_gaussians &out, float *meanD2);
int launchKnnFeatures(hipStream_t s, const float *xyz, const float *rgb, int rgbCount, int64_t n,
                      const float lo[3], float cell, const PointGrid &g, const float *meanD2,
                      const float *viewFrac, float *features);
int launchViewFraction(hipStream_t s, const float *xyz, int64_t n, const u3d_camera *cams,
                       int numCams, float nearPlane, float *out);

// Adds the initialiser network's outputs (row stride ld: 3 log-scale, 4 quaternion, 1 opacity
// logit, 3 DC colour corrections) to a kNN-initialised model.
int launchApplyCorrections(hipStream_t s, const float *net, int ld, int64_t n, u3d_gaussians &m);

// ---- kernels_mlp.hip
// C[m][n] = sum_k A(m, k) B(k, n) on BF16 MFMA with FP32 accumulation; A(m, k) = A[m*sam + k*sak],
// B(k, n) = B[k*sbk + n*sbn]. Epilogue (no split): + bias[n]; C (fp32, ldc) receives the value;
// when gate is set the value is first multiplied by GELU'(gate[m][n]); Y (bf16, ldy) receives
// act(value). With splitK > 1 every split atomically adds its partial sum into C (zeroed by
// the caller) and no epilogue runs.
struct GemmArgs {
    const uint16_t *A;
    int64_t sam, sak;
    const uint16_t *B;
    int64_t sbk, sbn;
    int M, N, K;
    float *C;
    int64_t ldc;
    const float *bias;
    uint16_t *Y;
    int64_t ldy;
    int gelu;
    const float *gate;
    int64_t ldg;
    int splitK;
};
int launchGemmBf16(hipStream_t s, const GemmArgs &a);
// bf16 copy of an fp32 array (round to nearest even).
int launchToBf16(hipStream_t s, const float *in, uint16_t *out, int64_t count);
// Column sums: out[n] += sum_m in[m * ld + n], m < M, n < N (fp32 atomics).
int launchColumnSums(hipStream_t s, const float *in, int64_t ld, int M, int N, float *out);
// Masked mean squared error of the network output: y [M x ld] vs target [M x ld] over the
// first `width` columns weighted by w[col]; writes dL/dy (fp32) and adds the loss to *loss.
int launchMseLoss(hipStream_t s, const float *y, const float *target, const float *w, int M,
                  int width, int64_t ld, float *dy, float *loss);
// Adam over a flat fp32 array, then its bf16 copy.
int launchAdamBf16(hipStream_t s, float *p, const float *g, float *m, float *v, int64_t n,
                   float stepSize, float invSqrtBc2, uint16_t *pBf16);

// ---- kernels_loss.hip
// Fused L1 + D-SSIM value and gradient. target is float (targetU8 == nullptr) or uint8/255.
// Writes per-block double partials (l1, ssim, squared error) into partials and the final
// {total, l1, ssim, psnr} floats into result (device) via a one-block finalize kernel; with
// `deferred`, the finalize is not launched but described there, for a later launch to run it.
// With ssimMaps (9 * width * height floats of scratch) the loss runs as two launches that pass
// the SSIM partials through it (the same gradient and values, bit for bit, with a third less
// arithmetic); without, as one fused launch.
int launchPhotometricLoss(hipStream_t s, const float *render, const float *target,
                          const uint8_t *targetU8, int width, int height, float lambda,
                          float *grad, double *partials, size_t partialsCap, float *result,
                          LossFinalize *deferred = nullptr, const SegmentPrep *prep = nullptr,
                          float *ssimMaps = nullptr);
inline size_t lossMapsSize(int width, int height) { return size_t(9) * size_t(width) * size_t(height); }
int launchLossFinalize(hipStream_t s, const LossFinalize &f);
size_t lossPartialsSize(int width, int height);
void lossItemGeometry(int width, int height, int *blocksX, int *numBlocks);
// Where the batched step leaves its pair-count status: the loss finalisation's first workgroup
// copies total[0..3] to host[4 * slot .. 4 * slot + 3] (pinned host memory, system-scope stores)
// with slot = *ring, then advances *ring to (slot + 1) % ringSize (ring null: no status).
struct StepStatus {
    const uint32_t *total = nullptr;
    uint32_t *host = nullptr;
    uint32_t *ring = nullptr;
    int ringSize = 1;
};
// splitMaps: every item has maps (the two-launch loss).
int launchPhotometricLossBatched(hipStream_t s, const LossItem *items, int numItems,
                                 int maxBlocks, float lambda, const StepStatus &status = {},
                                 bool splitMaps = false);
// Squared-error sum of two float arrays -> result[0] = psnr (double partials scratch).
int launchPsnr(hipStream_t s, const float *a, const float *b, const uint8_t *bU8, int64_t count,
               double *partials, size_t partialsCap, float *result);
size_t psnrPartialsSize(int64_t count);
// Mean |a - b/255| -> result[0] (double partials).
int launchMeanAbs(hipStream_t s, const float *a, const uint8_t *bU8, int64_t count,
                  double *partials, size_t partialsCap, float *result);

// ---- kernels_adam.hip
// clearGrads: each gradient is zeroed once applied.
int launchAdamStep(hipStream_t s, u3d_gaussians &p, const u3d_gaussians &g, u3d_gaussians &m,
                   u3d_gaussians &v, const float stepSize[5], float invSqrtBc2,
                   bool clearGrads = false);
// Adam over one extra parameter array (SH coefficients).
int launchAdamArray(hipStream_t s, float *p, const float *g, float *m, float *v, int64_t len,
                    float stepSize, float invSqrtBc2);
// Densification of a row array of `width` floats per Gaussian: kept rows compacted, extras
// copied from their parent (or zeroed when zeroRows, for optimiser moments). Needs the
// scans and counts of the same densification.
int launchScatterRows(hipStream_t s, const float *src, float *dst, int width, int64_t n,
                      const uint8_t *keep, const uint8_t *kind, const uint32_t *keepPos,
                      const uint32_t *extraPos, const uint32_t *counts, bool zeroRows);
// The same for `planes` planes of n floats (stride n) into planes of stride newN.
int launchScatterPlanes(hipStream_t s, const float *src, float *dst, int planes, int64_t n,
                        int64_t newN, const uint8_t *keep, const uint8_t *kind,
                        const uint32_t *keepPos, const uint32_t *extraPos,
                        const uint32_t *counts, bool zeroRows);
// Adam over every scene of a batched step (blockIdx.y = scene); clears the gradients it read.
int launchAdamBatched(hipStream_t s, const SceneStep *scenes, int numScenes, int64_t maxN);
int launchFill(hipStream_t s, float *p, float value, int64_t count);
int launchFillI32(hipStream_t s, int *p, int value, int64_t count);
// First launch of a batched step: copies the step's table (tableBytes, a multiple of 16) from
// pinned ring slot *ring (slot k at pinned + k * slotBytes) to the device table, and zeroes
// zeroA[0..countA) and zeroB[0..countB) (counters of the step). One workgroup.
int launchStepPrologue(hipStream_t s, const unsigned char *pinned, size_t slotBytes,
                       size_t tableBytes, unsigned char *table, const uint32_t *ring,
                       uint32_t *zeroA, int countA, uint32_t *zeroB, int countB);
// o = min(o, cap) (Trainer.cpp:171-178)
int launchOpacityCap(hipStream_t s, float *opacities, int64_t n, float cap);
// Densification: per-Gaussian demand (Trainer.cpp:116-135 before the budget test).
int launchDensifyDemand(hipStream_t s, const u3d_gaussians &m, const float *accum,
                        const int *count, float extent, float gradThreshold, float percentDense,
                        float pruneOpacity, int opacityWasReset, uint8_t *keep, uint32_t *demand);
// Greedy budget (Trainer.cpp:130-158): given the exclusive scan of demand, marks the accepted
// candidates in kind and clears keep for accepted splits; writes {clones, splits, kept,
// extras} to counts (device, 4 x u32). firstReject is 1-element device scratch.
int launchDensifyAccept(hipStream_t s, int64_t n, const uint32_t *demand, const uint32_t *scan,
                        uint64_t budget, uint8_t *keep, uint8_t *kind, uint32_t *firstReject,
                        uint32_t *counts);
// Builds keepIdx (exclusive scan input) etc.: writes per-Gaussian u32 keep flags and extra
// counts for the scatter.
int launchDensifyFlags(hipStream_t s, int64_t n, const uint8_t *keep, const uint8_t *kind,
                       uint32_t *keepFlag, uint32_t *extraCount, uint32_t *splitFlag);
// Scatter into dst (fresh arrays): kept entries at keepPos, extras at nKeep + extraPos.
int launchDensifyScatter(hipStream_t s, const u3d_gaussians &src, u3d_gaussians &dst,
                         int64_t n, const uint8_t *keep, const uint8_t *kind,
                         const uint32_t *keepPos, const uint32_t *extraPos,
                         const uint32_t *splitPos, const uint32_t *counts, const float *normals,
                         float logShrink);
// Moment compaction: kept entries to keepPos, appended zeros for the extras.
int launchCompactZeroTail(hipStream_t s, const u3d_gaussians &src, u3d_gaussians &dst,
                          int64_t n, const uint8_t *keep, const uint32_t *keepPos,
                          const uint32_t *counts);

} // namespace u3d
