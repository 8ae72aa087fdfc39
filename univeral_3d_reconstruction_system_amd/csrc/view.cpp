// C ABI of the viewer support (app/src/SplatView.cpp:117-180, 349-397): display attributes of a
// loaded model and the front-to-back order of a view, on the device.
#include "trainer.h"

namespace u3d {
int launchViewAttributes(hipStream_t s, const u3d_gaussians &m, uint8_t *rgba, float *cov6);
int viewRangeBlocks(int64_t n);
int launchDepthKeys(hipStream_t s, const float *centers, int64_t n, const float view_row2[4],
                    float *range, uint32_t *keys);
int launchPackView(hipStream_t s, const uint32_t *order, int64_t n, const float *centers,
                   const uint8_t *rgba, const float *cov6, void *packed);
} // namespace u3d

using namespace u3d;

extern "C" {

int u3d_view_splat_attributes(u3d_ctx *ctx, const u3d_gaussians *model, uint8_t *rgba, float *cov6)
{
    U3D_TRY(useDevice(ctx));
    U3D_TRY(checkModel(model, "model"));
    U3D_CHECK(model->n == 0 || (rgba && cov6), U3D_ERR_INVALID, "null output");
    U3D_TRY(launchViewAttributes(ctx->stream, *model, rgba, cov6));
    U3D_HIP(hipStreamSynchronize(ctx->stream));
    return U3D_OK;
}

int u3d_view_depth_order(u3d_ctx *ctx, const float *centers, int64_t n, const float *view_row2,
                         uint32_t *order, const uint8_t *rgba, const float *cov6, void *packed)
{
    U3D_TRY(useDevice(ctx));
    U3D_CHECK(n >= 0 && view_row2 != nullptr, U3D_ERR_INVALID, "bad argument");
    U3D_CHECK(n == 0 || (centers && order), U3D_ERR_INVALID, "null array");
    U3D_CHECK(!packed || n == 0 || (rgba && cov6), U3D_ERR_INVALID, "packing needs rgba and cov6");
    if (n == 0)
        return U3D_OK;
    const hipStream_t s = ctx->stream;
    const size_t nn = size_t(n);
    DevBuf<float> range;
    DevBuf<uint32_t> keys, keysAlt, orderAlt, hist, histScan;
    U3D_TRY(range.ensure(2 * size_t(viewRangeBlocks(n))));
    U3D_TRY(keys.ensure(nn));
    U3D_TRY(keysAlt.ensure(nn));
    U3D_TRY(orderAlt.ensure(nn));
    U3D_TRY(hist.ensure(radixHistSize(n)));
    U3D_TRY(histScan.ensure(radixScanScratchSize(n)));
    U3D_TRY(launchDepthKeys(s, centers, n, view_row2, range.p, keys.p));
    // the stable sort of the 16-bit keys with the indices as values: the counting sort's order
    RadixScratch rs{hist.p, hist.cap, histScan.p, histScan.cap};
    rs.valsAreIndex = true;
    bool alt = false;
    U3D_TRY(radixSortPairs(s, keys.p, order, keysAlt.p, orderAlt.p, n, 16, rs, &alt));
    if (alt)
        U3D_HIP(hipMemcpyAsync(order, orderAlt.p, nn * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    if (packed)
        U3D_TRY(launchPackView(s, order, n, centers, rgba, cov6, packed));
    U3D_HIP(hipStreamSynchronize(s));
    return U3D_OK;
}

} // extern "C"
