// Batched training engine (no reference counterpart; the reference trains one view of one scene
// per step, Trainer.cpp:53-103).
//
// Every (scene, view) render of a step is an item. All items share one projection launch, one
// depth sort over their concatenated screen entries (then a stable pass on the item, so each
// item's entries are contiguous and in depth order), the tile lists of each item built from
// difference grids straight into its own slice of one list (no (tile, Gaussian) pairs are
// materialised or sorted; images with more tiles than the binning's LDS grid holds fall back to
// emitting pairs and sorting them by global tile id), one persistent compositing launch per
// direction (workgroups pull 16x16 blocks of any item from an atomic counter, heaviest tile
// lists first), one loss launch, one chain-rule launch and one Adam launch. What changes from step to step (cameras, target and output
// pointers, optimiser coefficients, the stats slot) travels in one small table uploaded before
// the step, so the launches themselves only depend on the layout (Gaussian counts, image
// sizes, pair capacity): they are captured once as a hipGraph and replayed until the layout
// changes (densification, a grown pair capacity).
//
// The pair count K is never read back inside a step: the lists have a fixed capacity per item
// (or in all, on the fallback path). Pairs past the capacity are dropped (the farthest in depth
// order) and flagged; the host reads each step's counts and flag a few steps later and grows
// the capacity before it is reached again.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "trainer.h"

namespace u3d {

namespace {
constexpr int kRing = 4; // table uploads / status read-backs in flight
size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
int ceilDiv(int a, int b) { return (a + b - 1) / b; }
} // namespace

class BatchEngine
{
public:
    explicit BatchEngine(u3d_ctx *c, bool graph) : ctx(c), useGraph(graph) {}
    ~BatchEngine()
    {
        dropGraph();
        for (int k = 0; k < kRing; ++k) {
            if (slotDone[k])
                (void)hipEventDestroy(slotDone[k]);
            if (statusDone[k])
                (void)hipEventDestroy(statusDone[k]);
        }
        if (pinned)
            (void)hipHostFree(pinned);
        if (statusHost)
            (void)hipHostFree(statusHost);
    }

    int init()
    {
        for (int k = 0; k < kRing; ++k) {
            U3D_HIP(hipEventCreateWithFlags(&slotDone[k], hipEventDisableTiming));
            U3D_HIP(hipEventCreateWithFlags(&statusDone[k], hipEventDisableTiming));
            slotUsed[k] = false;
            statusStep[k] = -1;
        }
        U3D_HIP(hipHostMalloc(reinterpret_cast<void **>(&statusHost), sizeof(uint32_t) * 4 * kRing));
        U3D_TRY(ringPos.ensure(1));
        for (int k = 0; k < 3; ++k) {
            occFwd[k] = rasterBatchedOccupancy(false, 64 << k);
            occBwd[k] = rasterBatchedOccupancy(true, 64 << k);
        }
        return U3D_OK;
    }

    // New launch parameters: the captured graph (if any) is rebuilt on the next step.
    int setTuning(const u3d_batch_tuning &t)
    {
        U3D_CHECK(t.workgroups_per_cu >= 0 && t.workgroups_per_cu <= 64, U3D_ERR_INVALID,
                  "workgroups_per_cu must be 0..64");
        U3D_CHECK(t.lds_batch == 0 || t.lds_batch == 64 || t.lds_batch == 128 || t.lds_batch == 256,
                  U3D_ERR_INVALID, "lds_batch must be 0, 64, 128 or 256");
        U3D_CHECK(t.sort_items == 0 || t.sort_items == 4 || t.sort_items == 8, U3D_ERR_INVALID,
                  "sort_items must be 0, 4 or 8");
        U3D_CHECK(t.atomic_lanes == 0 || t.atomic_lanes == 4 || t.atomic_lanes == 16,
                  U3D_ERR_INVALID, "atomic_lanes must be 0, 4 or 16");
        U3D_CHECK(t.sort_radix == 0 || (t.sort_radix >= 4 && t.sort_radix <= 8), U3D_ERR_INVALID,
                  "sort_radix must be 0 or 4..8");
        U3D_CHECK(t.pair_sort >= 0 && t.pair_sort <= 2, U3D_ERR_INVALID, "pair_sort must be 0, 1 or 2");
        U3D_HIP(hipStreamSynchronize(ctx->stream));
        tuning.heavy_first = t.heavy_first != 0;
        tuning.workgroups_per_cu = t.workgroups_per_cu;
        tuning.lds_batch = t.lds_batch ? t.lds_batch : 256;
        tuning.sort_items = t.sort_items;
        tuning.tile_size = t.tile_size;
        tuning.atomic_lanes = t.atomic_lanes;
        tuning.sort_radix = t.sort_radix;
        if (t.pair_sort != tuning.pair_sort)
            needRelayout = true; // other buffers and list layout
        tuning.pair_sort = t.pair_sort;
        dropGraph();
        return U3D_OK;
    }
    u3d_batch_tuning tuning{1, 0, 256, 0, 0, 0, 0, 0};

    int step(u3d_trainer *const *scenes, int numScenes);
    int autotune(u3d_trainer *const *scenes, int numScenes, int stepsPerCandidate, bool tuneTile,
                 u3d_batch_tuning *chosen);
    size_t workingBytes() const { return allocated; }
    // Working-set estimate for one more member with n Gaussians, B views of P pixels each.
    static size_t estimateBytes(int64_t n, int views, size_t pixels)
    {
        const size_t entries = size_t(n) * size_t(views);
        const size_t perEntry = 2 * 16 + 4 * 4 + 16 + 1 + 8 * 4 + 9 * 4;
        const size_t pairs = entries * 24; // capacity for ~12 tiles per entry, 2 copies
        // per pixel: render, loss gradient, transmittance, contributor count, loss SSIM maps
        return entries * perEntry + pairs * 8 + size_t(views) * pixels * (40 + 36);
    }

    u3d_ctx *ctx;
    bool useGraph;

private:
    int occFwd[3] = {1, 1, 1}, occBwd[3] = {1, 1, 1}; // per LDS batch 64, 128, 256
    struct Member {
        const u3d_trainer *t;
        int64_t n;
        int width, height, views, shDegree;
        bool operator==(const Member &o) const
        {
            return t == o.t && n == o.n && width == o.width && height == o.height
                   && views == o.views && shDegree == o.shDegree;
        }
    };

    int relayout(const std::vector<u3d_trainer *> &act);
    int fillTable(const std::vector<u3d_trainer *> &act, unsigned char *dst);
    int probePairs();
    int enqueue();
    int capture();
    void dropGraph()
    {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    int poll(bool wait);
    ScreenBuffers screen() const
    {
        return {rec0.p, rec1.p, rec2.p, depth.p, radius.p, rect.p, flags.p, depthKey.p, nTiles.p};
    }

    // layout
    std::vector<Member> key;
    struct ItemGeometry {
        int64_t n, screenBase;
        int tileBase, tilesX, tilesY;
    };
    std::vector<ItemGeometry> geo; // per item, fixed for the layout
    // binned: per-image difference-grid binning, in index order followed by per-tile depth sorts
    // (pair_sort 0) or in the depth order of a sort of all entries (pair_sort 2); else a depth
    // sort of all entries, pairs and their sort on the tile id
    bool binned = false;
    bool depthBinned = false;
    int64_t itemCap = 0; // list slots per item (binned)
    int itemBits = 1;
    u3d_raster_settings st{};
    int numItems = 0, numGroups = 0, totalTiles = 0, maxLossBlocks = 0, tileBits = 1;
    int64_t totalEntries = 0, maxN = 0, totalBlocks = 0, pairCap = 0;
    int maxViews = 1;    // largest group
    bool splitLoss = true; // every image small enough for the split loss's 32-bit offsets
    bool anySh = false;  // a group with view-dependent colour (chain and Adam run apart)
    size_t totalPix = 0, partialsLen = 0;
    size_t offItems = 0, offLoss = 0, offSteps = 0, offGroups = 0, tableBytes = 0;
    bool needRelayout = true;
    int64_t grownCap = 0; // capacity requested by poll()
    size_t allocated = 0;

    // per-step table: pinned ring -> device
    unsigned char *pinned = nullptr;
    size_t pinnedSlot = 0;
    hipEvent_t slotDone[kRing] = {};
    bool slotUsed[kRing] = {};
    uint32_t *statusHost = nullptr; // pinned: {K, overflow, longest item K, -} per slot
    hipEvent_t statusDone[kRing] = {};
    int64_t statusStep[kRing] = {};
    std::vector<u3d_trainer *> statusScenes[kRing];
    int ring = 0;
    // The device's copy of ring: the step's first launch reads its table from pinned slot
    // *ringPos, the loss finalisation writes the status to that slot and advances it (both
    // inside the captured graph). ringSynced: *ringPos == ring.
    DevBuf<uint32_t> ringPos;
    bool ringSynced = false;
    int64_t stepNo = 0;

    // device buffers
    DevBuf<unsigned char> table;
    DevBuf<float4> rec0, rec1;
    DevBuf<float> rec2, depth, grad2d, rgb, dl, finalT;
    DevBuf<int> radius, entryItem, tileOffsets, nContrib;
    DevBuf<int4> rect;
    DevBuf<uint8_t> flags;
    DevBuf<uint32_t> depthKey, nTiles, keyAlt, valA, valAlt, ntSorted, pairOffsets, total;
    DevBuf<uint32_t> pairTile, pairGauss, pairTileAlt, pairGaussAlt, hist, histScan, scanScratch;
    DevBuf<uint32_t> blockKey, blockOrder, blockKeyAlt, blockOrderAlt;
    // binned path: the images' binning table (host copy and device copy, whose list slices
    // follow the per-item capacity), their (group, tile) counters, tile totals and longest-first
    // orders, the lists in index order and the per-tile sort's scratch
    std::vector<BinItem> binItems;
    BinItemsGrid binGrid;
    DevBuf<BinItem> binItemsDev;
    DevBuf<uint32_t> binCounts, binTotals, indexList, sortScratch, sortMap;
    PinnedScratch binStage; // the tables' upload staging (the stream is synchronised after)
    DevBuf<int4> rectSorted; // depth-order binning: the rectangles in depth order
    DevBuf<int> longFirst;
    DevBuf<unsigned> counters; // persistent work counters (forward, backward)
    DevBuf<double> partials;
    DevBuf<float> ssimMaps; // the loss's SSIM partials, 9 floats per pixel of every item

    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int BatchEngine::relayout(const std::vector<u3d_trainer *> &act)
{
    const hipStream_t s = ctx->stream;
    U3D_HIP(hipStreamSynchronize(s));
    dropGraph();
    ringSynced = false;
    st = act[0]->opt.raster;
    key.clear();
    numItems = 0;
    numGroups = int(act.size());
    totalTiles = 0;
    maxLossBlocks = 0;
    totalEntries = 0;
    maxN = 0;
    maxViews = 1;
    anySh = false;
    splitLoss = true;
    totalBlocks = 0;
    totalPix = 0;
    partialsLen = 0;
    geo.clear();
    binned = true;
    depthBinned = false;
    const int ts = st.tile_size;
    for (u3d_trainer *t : act) {
        U3D_CHECK(std::memcmp(&t->opt.raster, &st, sizeof st) == 0, U3D_ERR_UNSUPPORTED,
                  "batched scenes must share the raster settings");
        const int W = t->cams[0].width, H = t->cams[0].height;
        for (const u3d_camera &c : t->cams)
            U3D_CHECK(c.width == W && c.height == H, U3D_ERR_UNSUPPORTED,
                      "batched training needs one image size per scene");
        U3D_CHECK(t->grads.cap == t->params.cap && t->adamM.cap == t->params.cap
                      && t->adamV.cap == t->params.cap,
                  U3D_ERR_INVALID, "parameter, gradient and moment capacities differ");
        const int B = t->viewsPerStep;
        const int64_t n = t->params.n;
        key.push_back({t, n, W, H, B, t->shDegree});
        int lbx = 0, lnb = 0;
        lossItemGeometry(W, H, &lbx, &lnb);
        for (int k = 0; k < B; ++k)
            geo.push_back({n, totalEntries + n * k, totalTiles + ceilDiv(W, ts) * ceilDiv(H, ts) * k,
                           ceilDiv(W, ts), ceilDiv(H, ts)});
        binned = binned && tuning.pair_sort != 1 && tileBinningGroup(n, ceilDiv(W, ts), ceilDiv(H, ts)) > 0;
        numItems += B;
        totalEntries += n * B;
        maxN = std::max(maxN, n);
        maxViews = std::max(maxViews, B);
        splitLoss = splitLoss && size_t(W) * size_t(H) * 3 < (size_t(1) << 32);
        anySh = anySh || t->shDegree > 0;
        totalTiles += ceilDiv(W, ts) * ceilDiv(H, ts) * B;
        totalBlocks += int64_t(ceilDiv(W, 16)) * ceilDiv(H, 16) * B;
        totalPix += size_t(W) * size_t(H) * size_t(B);
        partialsLen += size_t(9) * size_t(lnb) * size_t(B);
        maxLossBlocks = std::max(maxLossBlocks, lnb);
    }
    U3D_CHECK(totalEntries < (int64_t(1) << 32) && totalTiles < (1 << 30), U3D_ERR_CAPACITY,
              "batch too large for 32-bit pair values");
    tileBits = bitsFor(totalTiles);
    itemBits = bitsFor(numItems);
    depthBinned = binned && tuning.pair_sort == 2;
    offItems = 0;
    offLoss = align256(size_t(numItems) * sizeof(RenderItem));
    offSteps = offLoss + align256(size_t(numItems) * sizeof(LossItem));
    offGroups = offSteps + align256(size_t(numGroups) * sizeof(SceneStep));
    tableBytes = offGroups + align256(size_t(numGroups) * sizeof(int2));

    const size_t ne = size_t(std::max<int64_t>(totalEntries, 1));
    U3D_TRY(table.ensure(tableBytes));
    if (pinnedSlot < tableBytes) {
        U3D_HIP(hipStreamSynchronize(s));
        if (pinned)
            U3D_HIP(hipHostFree(pinned));
        pinned = nullptr;
        pinnedSlot = align256(tableBytes * 2);
        U3D_HIP(hipHostMalloc(reinterpret_cast<void **>(&pinned), pinnedSlot * kRing));
        for (int k = 0; k < kRing; ++k)
            slotUsed[k] = false;
    }
    U3D_TRY(rec0.ensure(ne));
    U3D_TRY(rec1.ensure(ne));
    U3D_TRY(rec2.ensure(ne));
    U3D_TRY(depth.ensure(ne));
    U3D_TRY(radius.ensure(ne));
    U3D_TRY(rect.ensure(ne));
    U3D_TRY(flags.ensure(ne));
    U3D_TRY(depthKey.ensure(ne));
    U3D_TRY(nTiles.ensure(ne));
    U3D_TRY(keyAlt.ensure(ne));
    U3D_TRY(valA.ensure(ne));
    U3D_TRY(valAlt.ensure(ne));
    U3D_TRY(ntSorted.ensure(ne));
    U3D_TRY(pairOffsets.ensure(ne));
    U3D_TRY(entryItem.ensure(ne));
    U3D_TRY(grad2d.ensure(ne * kGrad2dStride));
    U3D_TRY(total.ensure(4));
    U3D_TRY(counters.ensure(4));
    // binned: every item's offsets with their own end (totalTiles + numItems entries)
    U3D_TRY(tileOffsets.ensure(size_t(totalTiles) + size_t(numItems)));
    if (binned) {
        binItems.assign(size_t(numItems), BinItem{});
        for (int v = 0; v < numItems; ++v) {
            const ItemGeometry &g = geo[size_t(v)];
            BinItem &it = binItems[size_t(v)];
            it.n = g.n;
            it.idBase = uint32_t(g.screenBase);
            it.tilesX = g.tilesX;
            it.tilesY = g.tilesY;
            it.offsetsBase = uint32_t(g.tileBase + v); // each image's offsets with their own end
        }
        std::vector<uint32_t> map;
        U3D_TRY(binItemsLayout(binItems, &binGrid, &map));
        U3D_TRY(binItemsDev.ensure(size_t(numItems)));
        U3D_TRY(sortMap.ensure(std::max<size_t>(map.size(), 1)));
        if (!map.empty()) {
            const size_t bytes = map.size() * sizeof(uint32_t);
            U3D_TRY(binStage.ensure(bytes));
            std::memcpy(binStage.p, map.data(), bytes);
            U3D_HIP(hipMemcpyAsync(sortMap.p, binStage.p, bytes, hipMemcpyHostToDevice, s));
            U3D_HIP(hipStreamSynchronize(s));
        }
        U3D_TRY(binCounts.ensure(std::max<size_t>(binGrid.counts, 1)));
        U3D_TRY(binTotals.ensure(std::max<size_t>(binGrid.totals, 1)));
        U3D_TRY(longFirst.ensure(std::max<size_t>(binGrid.orders, 1)));
        if (depthBinned)
            U3D_TRY(rectSorted.ensure(ne));
    }
    const size_t nb = size_t(std::max<int64_t>(totalBlocks, 1));
    U3D_TRY(blockKey.ensure(nb));
    U3D_TRY(blockOrder.ensure(nb));
    U3D_TRY(blockKeyAlt.ensure(nb));
    U3D_TRY(blockOrderAlt.ensure(nb));
    const size_t np = std::max<size_t>(totalPix, 1);
    U3D_TRY(rgb.ensure(np * 3));
    U3D_TRY(dl.ensure(np * 3));
    U3D_TRY(finalT.ensure(np));
    U3D_TRY(nContrib.ensure(np));
    U3D_TRY(partials.ensure(std::max<size_t>(partialsLen, 1)));
    U3D_TRY(ssimMaps.ensure(np * 9));
    // Gradients start at zero; afterwards the chain rule clears the 2D gradients it consumed and
    // Adam the parameter gradients it applied.
    U3D_HIP(hipMemsetAsync(grad2d.p, 0, grad2d.cap * sizeof(float), s));
    for (u3d_trainer *t : act) {
        t->grads.n = t->params.n;
        U3D_TRY(t->grads.zero(s));
        if (t->shDegree > 0)
            U3D_HIP(hipMemsetAsync(t->shG.p, 0, t->shG.cap * sizeof(float), s));
    }
    needRelayout = false;
    pairCap = 0; // sized by the probe of the first step of this layout (or grownCap)
    itemCap = 0;
    return U3D_OK;
}

int BatchEngine::fillTable(const std::vector<u3d_trainer *> &act, unsigned char *dst)
{
    RenderItem *items = reinterpret_cast<RenderItem *>(dst + offItems);
    LossItem *loss = reinterpret_cast<LossItem *>(dst + offLoss);
    SceneStep *steps = reinterpret_cast<SceneStep *>(dst + offSteps);
    int2 *groups = reinterpret_cast<int2 *>(dst + offGroups);
    const int ts = st.tile_size;
    int v = 0, tile = 0;
    int64_t screen = 0, block = 0;
    size_t pix = 0, part = 0;
    for (int g = 0; g < int(act.size()); ++g) {
        u3d_trainer *t = act[size_t(g)];
        const int B = t->viewsPerStep;
        const int W = t->cams[0].width, H = t->cams[0].height;
        const int64_t n = t->params.n;
        const int tilesX = ceilDiv(W, ts), tilesY = ceilDiv(H, ts);
        const int blocksX = ceilDiv(W, 16), blocksY = ceilDiv(H, 16);
        int lbx = 0, lnb = 0;
        lossItemGeometry(W, H, &lbx, &lnb);
        groups[g] = make_int2(v, B);
        const u3d_gaussians pv = t->params.view(), gv = t->grads.view();
        for (int k = 0; k < B; ++k, ++v) {
            const size_t view = size_t(trainerNextView(t));
            RenderItem it{};
            it.cam = t->cams[view];
            it.screenBase = screen;
            it.n = n;
            it.blockBase = block;
            it.pixBase = int64_t(pix);
            it.tileBase = tile;
            it.tilesX = tilesX;
            it.tilesY = tilesY;
            it.blocksX = blocksX;
            it.offsetShift = binned ? v : 0;
            it.shDegree = t->shDegree;
            it.ndcScale = float(B);
            it.means = pv.means;
            it.scales = pv.scales;
            it.quats = pv.quats;
            it.opacities = pv.opacities;
            it.colors = pv.colors;
            it.shRest = t->shDegree > 0 ? t->shP.p : nullptr;
            it.rgb = rgb.p + pix * 3;
            it.finalT = finalT.p + pix;
            it.nContrib = nContrib.p + pix;
            it.dl = dl.p + pix * 3;
            it.gMeans = gv.means;
            it.gScales = gv.scales;
            it.gQuats = gv.quats;
            it.gOpac = gv.opacities;
            it.gColors = gv.colors;
            it.gShRest = t->shDegree > 0 ? t->shG.p : nullptr;
            it.ndc = t->gradAccum.p;
            it.gradCount = t->gradCount.p;
            items[v] = it;
            LossItem li{};
            li.render = it.rgb;
            li.target = t->targets.p + t->imageOffset[view];
            li.grad = dl.p + pix * 3;
            li.result = t->statsSlot(size_t(t->iteration) * size_t(B) + size_t(k));
            li.partials = partials.p + part;
            li.maps = ssimMaps.p + pix * 9;
            li.width = W;
            li.height = H;
            li.blocksX = lbx;
            li.numBlocks = lnb;
            li.invCount = 1.0f / float(size_t(W) * size_t(H) * 3);
            li.gradScale = 1.0f / float(B); // the step's loss is the mean over its views
            loss[v] = li;
            screen += n;
            block += int64_t(blocksX) * blocksY;
            tile += tilesX * tilesY;
            pix += size_t(W) * size_t(H);
            part += size_t(9) * size_t(lnb);
        }
        // Optimiser coefficients of this step (Trainer.cpp:82-84, Adam.cpp:70-72); with batch
        // scaling, rates x sqrt(B) and decay rates beta^B.
        u3d_learning_rates lr = t->opt.lr;
        lr.means = trainerMeansLr(t);
        float b1 = 0.9f, b2 = 0.999f;
        if (t->opt.batch_scaling && B > 1) {
            const float r = std::sqrt(float(B));
            lr.means *= r;
            lr.scales *= r;
            lr.quats *= r;
            lr.opacities *= r;
            lr.colors *= r;
            b1 = std::pow(0.9f, float(B));
            b2 = std::pow(0.999f, float(B));
        }
        ++t->adamStep;
        SceneStep sc{};
        const float bc1 = 1.0f - std::pow(b1, float(t->adamStep));
        const float bc2 = 1.0f - std::pow(b2, float(t->adamStep));
        const float rates[5] = {lr.means, lr.scales, lr.quats, lr.opacities, lr.colors};
        for (int q = 0; q < 5; ++q)
            sc.stepSize[q] = rates[q] / bc1;
        sc.invSqrtBc2 = 1.0f / std::sqrt(bc2);
        sc.beta1 = b1;
        sc.beta2 = b2;
        sc.params = t->params.buf.p;
        sc.grads = t->grads.buf.p;
        sc.adamM = t->adamM.buf.p;
        sc.adamV = t->adamV.buf.p;
        sc.n = n;
        sc.cap = t->params.cap;
        sc.shDegree = t->shDegree;
        if (t->shDegree > 0) {
            sc.shP = t->shP.p;
            sc.shG = t->shG.p;
            sc.shM = t->shM.p;
            sc.shV = t->shV.p;
            sc.shStepSize = (lr.colors * shRateScale(t)) / bc1;
        }
        steps[g] = sc;
    }
    return U3D_OK;
}

// Sizes the pair capacity from this step's real pair count (projection, depth sort and scan
// run once more inside the step itself).
int BatchEngine::probePairs()
{
    const hipStream_t s = ctx->stream;
    const RenderItem *dItems = reinterpret_cast<const RenderItem *>(table.p + offItems);
    const ScreenBuffers sb = screen();
    U3D_TRY(launchStepPrologue(s, pinned, pinnedSlot, tableBytes, table.p, ringPos.p, nullptr, 0,
                               nullptr, 0));
    U3D_TRY(launchProjectBatched(s, dItems, numItems, maxN, st, sb, entryItem.p));
    U3D_TRY(scanScratch.ensure(scanScratchSize(totalEntries)));
    U3D_TRY(exclusiveScanU32(s, nTiles.p, pairOffsets.p, totalEntries, total.p, scanScratch.p,
                             scanScratch.cap));
    uint32_t K = 0;
    U3D_HIP(hipMemcpyAsync(&K, total.p, sizeof K, hipMemcpyDeviceToHost, s));
    // the pairs of each item: the scan at the item's first entry and at the next item's
    std::vector<uint32_t> starts(size_t(numItems), 0u);
    if (binned)
        for (int v = 0; v < numItems; ++v)
            if (geo[size_t(v)].screenBase < totalEntries)
                U3D_HIP(hipMemcpyAsync(&starts[size_t(v)], pairOffsets.p + geo[size_t(v)].screenBase,
                                       sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    U3D_HIP(hipStreamSynchronize(s));
    // Views differ in coverage: leave room for the spread between steps.
    const double margin = numItems >= 4 ? 1.3 : 1.6;
    if (binned) {
        uint32_t most = 0;
        for (int v = 0; v < numItems; ++v) {
            const uint32_t end = v + 1 < numItems ? starts[size_t(v) + 1] : K;
            most = std::max(most, end - starts[size_t(v)]);
        }
        // one capacity for every item: the views of a scene change from step to step
        int64_t cap = int64_t(double(most) * margin) + 65536;
        cap = std::max(cap, grownCap);
        const int64_t limit = (int64_t(1) << 31) / std::max(numItems, 1) - 1;
        U3D_CHECK(cap <= limit, U3D_ERR_CAPACITY, "tile lists of the batch exceed 2^31 entries");
        itemCap = cap;
        pairCap = cap * numItems;
        U3D_TRY(pairGauss.ensure(size_t(pairCap)));
        U3D_TRY(indexList.ensure(size_t(pairCap)));
        U3D_TRY(sortScratch.ensure(tileSortScratchSize(indexList.cap)));
        // the images' list slices follow the capacity
        for (int v = 0; v < numItems; ++v)
            binItems[size_t(v)].listBase = uint32_t(int64_t(v) * itemCap);
        const size_t bytes = binItems.size() * sizeof(BinItem);
        U3D_TRY(binStage.ensure(bytes));
        std::memcpy(binStage.p, binItems.data(), bytes);
        U3D_HIP(hipMemcpyAsync(binItemsDev.p, binStage.p, bytes, hipMemcpyHostToDevice, s));
        U3D_HIP(hipStreamSynchronize(s));
    } else {
        int64_t cap = int64_t(double(K) * margin) + 65536;
        cap = std::max(cap, grownCap);
        cap = std::min<int64_t>(cap, (int64_t(1) << 32) - 1);
        pairCap = cap;
        const size_t pc = size_t(pairCap);
        U3D_TRY(pairTile.ensure(pc));
        U3D_TRY(pairGauss.ensure(pc));
        U3D_TRY(pairTileAlt.ensure(pc));
        U3D_TRY(pairGaussAlt.ensure(pc));
    }
    const int64_t sortMax = std::max(std::max(totalEntries, binned ? 0 : pairCap), totalBlocks);
    U3D_TRY(hist.ensure(radixHistSize(sortMax)));
    U3D_TRY(histScan.ensure(radixScanScratchSize(sortMax)));
    U3D_TRY(scanScratch.ensure(scanScratchSize(sortMax)));
    allocated = 0;
    for (size_t b : {rec0.cap * 16, rec1.cap * 16, rec2.cap * 4, depth.cap * 4, grad2d.cap * 4,
                     radius.cap * 4, rect.cap * 16, flags.cap, depthKey.cap * 4, nTiles.cap * 4,
                     keyAlt.cap * 4, valA.cap * 4, valAlt.cap * 4, ntSorted.cap * 4,
                     pairOffsets.cap * 4, entryItem.cap * 4, pairTile.cap * 12, pairGauss.cap * 4,
                     binCounts.cap * 4, binTotals.cap * 4, indexList.cap * 4,
                     sortScratch.cap * 4, longFirst.cap * 4, sortMap.cap * 4, rectSorted.cap * 16,
                     hist.cap * 4,
                     histScan.cap * 4, scanScratch.cap * 4, rgb.cap * 4, dl.cap * 4, finalT.cap * 4,
                     nContrib.cap * 4, partials.cap * 8, blockKey.cap * 16, tileOffsets.cap * 4,
                     ssimMaps.cap * 4})
        allocated += b;
    return U3D_OK;
}

// Every launch of one step, in stream order. Nothing here allocates, synchronises or writes
// with a memset, so the sequence can be captured as a graph.
int BatchEngine::enqueue()
{
    const hipStream_t s = ctx->stream;
    const RenderItem *dItems = reinterpret_cast<const RenderItem *>(table.p + offItems);
    const LossItem *dLoss = reinterpret_cast<const LossItem *>(table.p + offLoss);
    const SceneStep *dSteps = reinterpret_cast<const SceneStep *>(table.p + offSteps);
    const int2 *dGroups = reinterpret_cast<const int2 *>(table.p + offGroups);
    const ScreenBuffers sb = screen();
    {
        StageScope scope(ctx, U3D_STAGE_PROJECT);
        // the step's table from its pinned slot; K, overflow flag, longest item list set and
        // the compositing work counters start at zero
        U3D_TRY(launchStepPrologue(s, pinned, pinnedSlot, tableBytes, table.p, ringPos.p, total.p,
                                   4, counters.p, 2));
        U3D_TRY(launchProjectBatched(s, dItems, numItems, maxN, st, sb, entryItem.p));
    }
    const uint32_t *order = nullptr;
    const int *tileList = nullptr;
    {
        StageScope scope(ctx, U3D_STAGE_SORT);
        RadixScratch rs{hist.p, hist.cap, histScan.p, histScan.cap, tuning.sort_items};
        rs.radixBits = tuning.sort_radix ? tuning.sort_radix : 8;
        if (depthBinned) {
            // all entries in depth order, then a stable pass on the item (each item's entries
            // contiguous at its screen base, in depth order, invisible ones last); every image's
            // tiles binned in that order straight into its slice of the lists
            U3D_TRY(launchIota(s, valA.p, totalEntries));
            bool alt = false;
            U3D_TRY(radixSortPairs(s, depthKey.p, valA.p, keyAlt.p, valAlt.p, totalEntries, 32, rs,
                                   &alt));
            uint32_t *k0 = alt ? depthKey.p : keyAlt.p, *k1 = alt ? keyAlt.p : depthKey.p;
            uint32_t *v1 = alt ? valAlt.p : valA.p, *v0 = alt ? valA.p : valAlt.p;
            U3D_TRY(launchGatherU32(s, reinterpret_cast<const uint32_t *>(entryItem.p), v1, k0,
                                    totalEntries, totalEntries));
            bool alt4 = false;
            U3D_TRY(radixSortPairs(s, k0, v1, k1, v0, totalEntries, itemBits, rs, &alt4));
            const uint32_t *byItem = alt4 ? v0 : v1;
            U3D_TRY(launchTileBinningItems(s, binItemsDev.p, numItems, binGrid, nTiles.p, rect.p,
                                           binCounts.p, binTotals.p, tileOffsets.p, pairGauss.p,
                                           size_t(itemCap), total.p, total.p + 2, total.p + 1,
                                           nullptr, byItem, ntSorted.p, rectSorted.p,
                                           totalEntries));
            tileList = reinterpret_cast<const int *>(pairGauss.p);
        } else if (binned) {
            // every image's tiles binned in index order into its slice of the lists, then each
            // list sorted on the depth keys: one launch per phase for all images
            U3D_TRY(launchTileBinningItems(s, binItemsDev.p, numItems, binGrid, nTiles.p, rect.p,
                                           binCounts.p, binTotals.p, tileOffsets.p, indexList.p,
                                           size_t(itemCap), total.p, total.p + 2, total.p + 1,
                                           longFirst.p));
            U3D_TRY(launchTileSortItems(s, binItemsDev.p, sortMap.p, binGrid, tileOffsets.p,
                                        longFirst.p, indexList.p, pairGauss.p, depthKey.p,
                                        sortScratch.p, sortScratch.cap, indexList.cap));
            tileList = reinterpret_cast<const int *>(pairGauss.p);
        } else {
            // all entries in depth order (a radix sort of the depth keys), pairs emitted in that
            // order and sorted on the global tile id
            U3D_TRY(launchIota(s, valA.p, totalEntries));
            bool alt = false;
            U3D_TRY(radixSortPairs(s, depthKey.p, valA.p, keyAlt.p, valAlt.p, totalEntries, 32, rs,
                                   &alt));
            const uint32_t *sorted = alt ? valAlt.p : valA.p;
            U3D_TRY(launchGatherU32(s, nTiles.p, sorted, ntSorted.p, totalEntries, totalEntries));
            U3D_TRY(exclusiveScanU32(s, ntSorted.p, pairOffsets.p, totalEntries, total.p,
                                     scanScratch.p, scanScratch.cap));
            U3D_TRY(launchEmitPairsBatched(s, totalEntries, sorted, pairOffsets.p, rect.p,
                                           entryItem.p, dItems, pairTile.p, pairGauss.p,
                                           ntSorted.p, pairCap, total.p + 1));
            bool alt2 = false;
            U3D_TRY(radixSortPairs(s, pairTile.p, pairGauss.p, pairTileAlt.p, pairGaussAlt.p,
                                   pairCap, tileBits, rs, &alt2, total.p));
            const uint32_t *tileSorted = alt2 ? pairTileAlt.p : pairTile.p;
            tileList = reinterpret_cast<const int *>(alt2 ? pairGaussAlt.p : pairGauss.p);
            U3D_TRY(launchTileOffsets(s, tileSorted, pairCap, totalTiles, tileOffsets.p, total.p));
        }
        if (tuning.heavy_first) {
            U3D_TRY(launchBlockCosts(s, dItems, numItems, totalBlocks, st.tile_size, tileOffsets.p,
                                     blockKey.p, blockOrder.p));
            bool alt3 = false;
            U3D_TRY(radixSortPairs(s, blockKey.p, blockOrder.p, blockKeyAlt.p, blockOrderAlt.p,
                                   totalBlocks, 8, rs, &alt3));
            order = alt3 ? blockOrderAlt.p : blockOrder.p;
        }
    }
    const int perCu = tuning.workgroups_per_cu;
    const int lds = tuning.lds_batch, li = lds == 64 ? 0 : (lds == 128 ? 1 : 2);
    {
        StageScope scope(ctx, U3D_STAGE_RASTER_FWD);
        const int occ = perCu > 0 ? perCu : occFwd[li];
        const int64_t wgs = std::min<int64_t>(int64_t(ctx->num_cus) * occ, totalBlocks);
        U3D_TRY(launchRasterBatched(s, false, dItems, numItems, order, totalBlocks, counters.p,
                                    int(std::max<int64_t>(wgs, 1)), tileOffsets.p, tileList, sb,
                                    st, nullptr, lds, 0, SegmentState{nullptr, nullptr}));
    }
    {
        StageScope scope(ctx, U3D_STAGE_LOSS);
        StepStatus status;
        status.total = total.p;
        status.host = statusHost;
        status.ring = ringPos.p;
        status.ringSize = kRing;
        U3D_TRY(launchPhotometricLossBatched(s, dLoss, numItems, maxLossBlocks,
                                             key.empty() ? 0.2f : key[0].t->opt.ssim_lambda,
                                             status, splitLoss));
    }
    {
        StageScope scope(ctx, U3D_STAGE_RASTER_BWD);
        const int occ = perCu > 0 ? perCu : occBwd[li];
        const int64_t wgs = std::min<int64_t>(int64_t(ctx->num_cus) * occ, totalBlocks);
        U3D_TRY(launchRasterBatched(s, true, dItems, numItems, order, totalBlocks, counters.p + 1,
                                    int(std::max<int64_t>(wgs, 1)), tileOffsets.p, tileList, sb,
                                    st, grad2d.p, lds, tuning.atomic_lanes,
                                    SegmentState{nullptr, nullptr}));
    }
    if (!anySh) {
        StageScope scope(ctx, U3D_STAGE_CHAIN);
        U3D_TRY(launchChainAdamBatched(s, dItems, dGroups, dSteps, numGroups, maxN, maxViews, st,
                                       sb, grad2d.p));
        return U3D_OK;
    }
    {
        StageScope scope(ctx, U3D_STAGE_CHAIN);
        U3D_TRY(launchChainBatched(s, dItems, dGroups, numGroups, maxN, st, sb, grad2d.p));
    }
    {
        StageScope scope(ctx, U3D_STAGE_ADAM);
        U3D_TRY(launchAdamBatched(s, dSteps, numGroups, maxN));
    }
    return U3D_OK;
}

int BatchEngine::capture()
{
    dropGraph();
    const hipStream_t s = ctx->stream;
    const bool timing = ctx->timer.enabled;
    ctx->timer.enabled = false; // no event records inside the graph
    U3D_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    ctx->timer.enabled = timing;
    if (rc != U3D_OK) {
        if (g)
            (void)hipGraphDestroy(g);
        return rc;
    }
    U3D_HIP(e);
    graph = g;
    U3D_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    return U3D_OK;
}

// Reads the pair counts and overflow flags of finished steps; asks for a larger capacity when a
// step came within 10% of it.
int BatchEngine::poll(bool wait)
{
    for (int k = 0; k < kRing; ++k) {
        if (statusStep[k] < 0)
            continue;
        if (wait)
            U3D_HIP(hipEventSynchronize(statusDone[k]));
        else if (hipEventQuery(statusDone[k]) != hipSuccess)
            continue;
        const int64_t K = int64_t(statusHost[4 * k]);
        const bool overflow = statusHost[4 * k + 1] != 0;
        const int64_t most = int64_t(statusHost[4 * k + 2]); // binned: the longest item list set
        for (u3d_trainer *t : statusScenes[k]) {
            if (statusScenes[k].size() == 1)
                t->counters[3] += K;
            if (overflow)
                t->counters[5] += 1;
        }
        statusStep[k] = -1;
        // binned: capacities are per item; otherwise for the whole batch
        const int64_t used = binned ? most : K, cap = binned ? itemCap : pairCap;
        if (double(used) > 0.9 * double(cap)) {
            grownCap = std::max(grownCap, int64_t(double(used) * 1.3) + 65536);
            needRelayout = true;
        }
    }
    return U3D_OK;
}

int BatchEngine::step(u3d_trainer *const *scenes, int numScenes)
{
    const hipStream_t s = ctx->stream;
    std::vector<u3d_trainer *> act;
    for (int k = 0; k < numScenes; ++k) {
        u3d_trainer *t = scenes[k];
        if (t && !t->cams.empty() && t->iteration < t->opt.iterations)
            act.push_back(t);
    }
    if (act.empty())
        return U3D_OK;
    U3D_TRY(poll(false));
    std::vector<Member> now;
    for (u3d_trainer *t : act)
        now.push_back({t, t->params.n, t->cams[0].width, t->cams[0].height, t->viewsPerStep,
                       t->shDegree});
    if (needRelayout || !(now == key)) {
        U3D_TRY(poll(true));
        U3D_TRY(relayout(act));
    }
    // This step's table: wait until the step that last used the slot has finished (its first
    // launch copies the table to the device, its loss finalisation writes the slot's status).
    const int slot = ring;
    if (slotUsed[slot])
        U3D_HIP(hipEventSynchronize(slotDone[slot]));
    unsigned char *host = pinned + size_t(slot) * pinnedSlot;
    U3D_TRY(fillTable(act, host));
    if (totalEntries > 0) {
        if (statusStep[slot] >= 0)
            U3D_TRY(poll(true));
        if (!ringSynced) {
            U3D_TRY(launchFillI32(s, reinterpret_cast<int *>(ringPos.p), slot, 1));
            ringSynced = true;
        }
        // a failed launch below leaves the device ring unknown
        ringSynced = false;
        if (pairCap == 0)
            U3D_TRY(probePairs());
        if (useGraph) {
            if (!exec)
                U3D_TRY(capture());
            StageScope scope(ctx, U3D_STAGE_GRAPH);
            U3D_HIP(hipGraphLaunch(exec, s));
        } else {
            U3D_TRY(enqueue());
        }
        ringSynced = true;
        U3D_HIP(hipEventRecord(slotDone[slot], s));
        slotUsed[slot] = true;
        U3D_HIP(hipEventRecord(statusDone[slot], s));
        statusStep[slot] = stepNo;
        statusScenes[slot] = act;
        ring = (ring + 1) % kRing;
    }
    ++stepNo;
    for (u3d_trainer *t : act) {
        const int64_t B = t->viewsPerStep;
        t->counters[0] += 1;
        t->counters[1] += t->params.n * B;
        t->counters[2] += int64_t(t->cams[0].width) * t->cams[0].height * B;
        t->counters[4] += B;
        U3D_TRY(trainerFinishStep(t));
    }
    return U3D_OK;
}

// Times candidate launch configurations on the scenes' next training steps (real steps: the
// choices below change no result beyond float summation order) and keeps the fastest:
// persistent workgroups per CU, heavy-first block order, LDS batch, radix items per lane and,
// when tuneTile, the tile size (which, as RasterSettings::tileSize, changes which Gaussians a
// tile keeps near the 3-sigma cut).
int BatchEngine::autotune(u3d_trainer *const *scenes, int numScenes, int stepsPerCandidate,
                          bool tuneTile, u3d_batch_tuning *chosen)
{
    const hipStream_t s = ctx->stream;
    std::vector<u3d_batch_tuning> cands;
    const int baseTile = numScenes > 0 ? scenes[0]->opt.raster.tile_size : 16;
    // stage 1: the compositing launches and the tile size (sorts at their defaults)
    for (int tile : tuneTile ? std::vector<int>{16, 32} : std::vector<int>{baseTile})
        for (int lds : {64, 128, 256})
            for (int heavy : {1, 0})
                for (int perCu : {0, 2})
                    for (int items : {0, 4})
                        cands.push_back({heavy, perCu, lds, items, tile, 0, 0, 0});
    const size_t stage1 = cands.size();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    U3D_HIP(hipEventCreate(&e0));
    U3D_HIP(hipEventCreate(&e1));
    float bestMs = 3.4e38f;
    u3d_batch_tuning best = tuning;
    best.tile_size = baseTile;
    int rc = U3D_OK;
    for (size_t ci = 0; ci < cands.size(); ++ci) {
        const u3d_batch_tuning c = cands[ci];
        bool live = false;
        for (int k = 0; k < numScenes; ++k)
            live |= scenes[k]->iteration < scenes[k]->opt.iterations;
        if (!live)
            break;
        if ((rc = setTuning(c)) != U3D_OK)
            break;
        for (int k = 0; k < numScenes; ++k)
            scenes[k]->opt.raster.tile_size = c.tile_size;
        needRelayout = needRelayout || c.tile_size != st.tile_size;
        if ((rc = step(scenes, numScenes)) != U3D_OK) // warm-up: relayout, probe, capture
            break;
        if ((rc = hipEventRecord(e0, s) == hipSuccess ? U3D_OK : U3D_ERR_HIP) != U3D_OK)
            break;
        int done = 0;
        for (; done < stepsPerCandidate; ++done)
            if ((rc = step(scenes, numScenes)) != U3D_OK)
                break;
        if (rc != U3D_OK || hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
            break;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (done > 0 && ms / float(done) < bestMs) {
            bestMs = ms / float(done);
            best = c;
        }
        if (ci + 1 == stage1) {
            // stage 2, from the best of stage 1: the sort radix and how the tile lists are built
            for (int radix : {8, 6, 4})
                for (int pairs : {0, 1, 2}) {
                    u3d_batch_tuning next = best;
                    next.sort_radix = radix;
                    next.pair_sort = pairs;
                    cands.push_back(next);
                }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    U3D_TRY(rc);
    U3D_TRY(setTuning(best));
    for (int k = 0; k < numScenes; ++k)
        scenes[k]->opt.raster.tile_size = best.tile_size;
    needRelayout = true;
    if (chosen)
        *chosen = best;
    return U3D_OK;
}

int createBatchEngine(u3d_ctx *ctx, bool useGraph, BatchEngine **out)
{
    std::unique_ptr<BatchEngine> e(new BatchEngine(ctx, useGraph));
    U3D_TRY(e->init());
    *out = e.release();
    return U3D_OK;
}

void destroyBatchEngine(BatchEngine *e) { delete e; }

int batchTrainStep(BatchEngine *engine, u3d_trainer *const *scenes, int numScenes)
{
    return engine->step(scenes, numScenes);
}

namespace {
size_t trainerBytes(const u3d_trainer *t)
{
    size_t b = (t->params.buf.cap + t->spare.buf.cap + t->grads.buf.cap + t->adamM.buf.cap
                + t->adamV.buf.cap) * sizeof(float);
    b += t->targets.cap + (t->render.cap + t->dloss.cap + t->ssimMaps.cap + t->gradAccum.cap) * sizeof(float);
    b += t->gradCount.cap * sizeof(int);
    b += (t->shP.cap + t->shG.cap + t->shM.cap + t->shV.cap + t->shSpare.cap) * sizeof(float);
    return b;
}
} // namespace

} // namespace u3d

using namespace u3d;

struct u3d_scene_set {
    u3d_ctx *ctx = nullptr;
    std::unique_ptr<BatchEngine> engine;
    std::vector<u3d_trainer *> members;
    size_t budget = 0;
    size_t memberBytes = 0;
    size_t engineEstimate = 0;
};

extern "C" {

int u3d_scene_set_create(u3d_ctx *ctx, int use_graph, size_t hbm_budget_bytes, u3d_scene_set **out)
{
    U3D_TRY(useDevice(ctx));
    U3D_CHECK(out != nullptr, U3D_ERR_INVALID, "null out");
    std::unique_ptr<u3d_scene_set> set(new u3d_scene_set());
    set->ctx = ctx;
    BatchEngine *e = nullptr;
    U3D_TRY(createBatchEngine(ctx, use_graph != 0, &e));
    set->engine.reset(e);
    if (hbm_budget_bytes == 0) {
        size_t freeB = 0, totalB = 0;
        U3D_HIP(hipMemGetInfo(&freeB, &totalB));
        hbm_budget_bytes = size_t(double(freeB) * 0.9);
    }
    set->budget = hbm_budget_bytes;
    *out = set.release();
    return U3D_OK;
}

void u3d_scene_set_destroy(u3d_scene_set *set)
{
    if (!set)
        return;
    (void)hipSetDevice(set->ctx->device);
    (void)hipStreamSynchronize(set->ctx->stream);
    for (u3d_trainer *t : set->members)
        u3d_trainer_destroy(t);
    delete set;
}

int u3d_scene_set_add(u3d_scene_set *set, u3d_trainer *scene)
{
    U3D_CHECK(set && scene, U3D_ERR_INVALID, "null argument");
    U3D_CHECK(scene->ctx == set->ctx, U3D_ERR_INVALID, "scene belongs to another context");
    for (u3d_trainer *t : set->members)
        U3D_CHECK(t != scene, U3D_ERR_INVALID, "scene already in the set");
    U3D_CHECK(scene->opt.raster.tile_size % 16 == 0, U3D_ERR_UNSUPPORTED,
              "scene sets need a tile_size that is a multiple of 16");
    const size_t own = trainerBytes(scene);
    size_t pix = 0;
    for (const u3d_camera &c : scene->cams)
        pix = std::max(pix, size_t(c.width) * size_t(c.height));
    const size_t work = BatchEngine::estimateBytes(std::max<int64_t>(scene->params.cap, 1),
                                                   scene->viewsPerStep, pix);
    U3D_CHECK(set->memberBytes + set->engineEstimate + own + work <= set->budget, U3D_ERR_CAPACITY,
              "scene does not fit the scene set's HBM budget");
    destroyBatchEngine(scene->engine); // the set's engine trains it from now on
    scene->engine = nullptr;
    set->members.push_back(scene);
    set->memberBytes += own;
    set->engineEstimate += work;
    return U3D_OK;
}

int u3d_scene_set_size(u3d_scene_set *set) { return set ? int(set->members.size()) : 0; }

u3d_trainer *u3d_scene_set_member(u3d_scene_set *set, int k)
{
    if (!set || k < 0 || k >= int(set->members.size()))
        return nullptr;
    return set->members[size_t(k)];
}

int u3d_scene_set_step(u3d_scene_set *set)
{
    U3D_CHECK(set != nullptr, U3D_ERR_INVALID, "null scene set");
    U3D_TRY(useDevice(set->ctx));
    return set->engine->step(set->members.data(), int(set->members.size()));
}

int u3d_scene_set_done(u3d_scene_set *set)
{
    if (!set)
        return 1;
    for (u3d_trainer *t : set->members)
        if (!u3d_trainer_done(t))
            return 0;
    return 1;
}

size_t u3d_scene_set_bytes(u3d_scene_set *set)
{
    if (!set)
        return 0;
    return set->memberBytes + std::max(set->engineEstimate, set->engine->workingBytes());
}

int u3d_trainer_get_tuning(u3d_trainer *t, u3d_batch_tuning *out)
{
    U3D_CHECK(t && out, U3D_ERR_INVALID, "null argument");
    U3D_CHECK(t->engine != nullptr, U3D_ERR_INVALID, "trainer does not use the batched engine");
    *out = t->engine->tuning;
    out->tile_size = t->opt.raster.tile_size;
    return U3D_OK;
}

int u3d_trainer_set_tuning(u3d_trainer *t, const u3d_batch_tuning *tuning)
{
    U3D_CHECK(t && tuning, U3D_ERR_INVALID, "null argument");
    U3D_CHECK(t->engine != nullptr, U3D_ERR_INVALID, "trainer does not use the batched engine");
    U3D_CHECK(tuning->tile_size == 0 || (tuning->tile_size > 0 && tuning->tile_size % 16 == 0),
              U3D_ERR_UNSUPPORTED, "tile_size must be a positive multiple of 16");
    if (tuning->tile_size)
        t->opt.raster.tile_size = tuning->tile_size;
    return t->engine->setTuning(*tuning);
}

int u3d_scene_set_set_tuning(u3d_scene_set *set, const u3d_batch_tuning *tuning)
{
    U3D_CHECK(set && tuning, U3D_ERR_INVALID, "null argument");
    U3D_CHECK(tuning->tile_size == 0 || (tuning->tile_size > 0 && tuning->tile_size % 16 == 0),
              U3D_ERR_UNSUPPORTED, "tile_size must be a positive multiple of 16");
    if (tuning->tile_size)
        for (u3d_trainer *t : set->members)
            t->opt.raster.tile_size = tuning->tile_size;
    return set->engine->setTuning(*tuning);
}

int u3d_trainer_autotune(u3d_trainer *t, int steps_per_candidate, int tune_tile_size,
                         u3d_batch_tuning *chosen)
{
    U3D_CHECK(t != nullptr && steps_per_candidate > 0, U3D_ERR_INVALID, "bad argument");
    U3D_CHECK(t->engine != nullptr, U3D_ERR_INVALID, "trainer does not use the batched engine");
    U3D_TRY(useDevice(t->ctx));
    u3d_trainer *one[1] = {t};
    return t->engine->autotune(one, 1, steps_per_candidate, tune_tile_size != 0, chosen);
}

int u3d_scene_set_autotune(u3d_scene_set *set, int steps_per_candidate, int tune_tile_size,
                           u3d_batch_tuning *chosen)
{
    U3D_CHECK(set != nullptr && steps_per_candidate > 0, U3D_ERR_INVALID, "bad argument");
    U3D_TRY(useDevice(set->ctx));
    return set->engine->autotune(set->members.data(), int(set->members.size()), steps_per_candidate,
                                 tune_tile_size != 0, chosen);
}

} // extern "C"
