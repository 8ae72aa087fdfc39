// photometricLoss of the drop-in API over libu3d_hip; psnr on the host (Loss.cpp:187-196).
#include "u3d/splat/Loss.h"

#include "device.h"

#include <algorithm>
#include <cmath>

namespace u3d {

LossValue photometricLoss(const float *render, const float *target, int width, int height,
                          float ssimLambda, float *grad, ThreadPool &)
{
    static dev::Buffer bufR, bufT, bufG;
    u3d_ctx *ctx = dev::context();
    const std::size_t bytes = std::size_t(width) * std::size_t(height) * 3 * sizeof(float);
    float *r = static_cast<float *>(bufR.ensure(bytes + 4));
    float *t = static_cast<float *>(bufT.ensure(bytes + 4));
    float *g = static_cast<float *>(bufG.ensure(bytes + 4));
    dev::check(u3d_memcpy_h2d(ctx, r, render, bytes), "upload render");
    dev::check(u3d_memcpy_h2d(ctx, t, target, bytes), "upload target");
    u3d_loss_value v{};
    dev::check(u3d_photometric_loss(ctx, r, t, width, height, ssimLambda, g, &v), "u3d_photometric_loss");
    dev::check(u3d_memcpy_d2h(ctx, grad, g, bytes), "download grad");
    LossValue out;
    out.total = v.total;
    out.l1 = v.l1;
    out.ssim = v.ssim;
    return out;
}

float psnr(const float *a, const float *b, std::size_t count)
{
    double mse = 0;
    for (std::size_t k = 0; k < count; ++k) {
        const double d = double(a[k]) - double(b[k]);
        mse += d * d;
    }
    mse /= double(std::max<std::size_t>(count, 1));
    return mse <= 1e-12 ? 99.0f : float(-10.0 * std::log10(mse));
}

} // namespace u3d
