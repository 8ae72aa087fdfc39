// Projection (Rasterizer.cpp:137-204) and backward chain rule (Rasterizer.cpp:432-544): one
// thread per Gaussian over structure-of-arrays inputs. Both are HBM-bound streaming passes.
// Compiled with -ffp-contract=off: projection outputs are bit-identical to the reference.
#include "kernels.h"
#include "projection.h"

namespace u3d {

namespace {

constexpr int kThreads = 256;

// Basis functions of SH degree kDeg (1..3), or of the runtime degree (kDeg < 0; sh.coef
// non-null); 0: none. A compile-time degree keeps the basis arrays in registers.
template <int kDeg>
__device__ __forceinline__ int shTerms(const ShCoefs &sh)
{
    return kDeg > 0 ? shCount(kDeg) : (kDeg < 0 && sh.coef ? shCount(sh.degree) : 0);
}

// Projects Gaussian i of a model and writes its screen state at entry o; returns the number of
// tiles it covers. kDeg: the SH degree of the colour (0: the reference's DC colour; < 0: the
// runtime degree of sh).
template <int kDeg>
__device__ __forceinline__ uint32_t projectBody(
    int64_t i, int64_t o, const float *__restrict__ means, const float *__restrict__ scales,
    const float *__restrict__ quats, const float *__restrict__ opac,
    const float *__restrict__ colors, u3d_camera cam, float nearPlane, float eps2d, int ts,
    int tilesX, int tilesY, ScreenBuffers out, const ShCoefs &sh, float3 campos)
{
    int radius = 0;
    uint8_t flags = 0;
    uint32_t key = 0xffffffffu, nt = 0;
    do {
        const float mu[3] = {means[i * 3], means[i * 3 + 1], means[i * 3 + 2]};
        const float qr[4] = {quats[i * 4], quats[i * 4 + 1], quats[i * 4 + 2], quats[i * 4 + 3]};
        const float ls[3] = {scales[i * 3], scales[i * 3 + 1], scales[i * 3 + 2]};
        Projection p;
        if (!projectGaussian(mu, qr, ls, cam, nearPlane, eps2d, p))
            break;
        const float det = p.a * p.c - p.b * p.b;
        if (!(det > 0.0f))
            break;
        const float invDet = 1.0f / det;
        const float z = p.pc[2];
        const float u = cam.fx * p.pc[0] / z + cam.cx;
        const float v = cam.fy * p.pc[1] / z + cam.cy;
        const float mid = 0.5f * (p.a + p.c);
        const float disc = mid * mid - det;
        const float lambda = mid + sqrtf((0.1f < disc) ? disc : 0.1f); // std::max(0.1f, disc)
        const int r = cvtTruncI32(ceilf(3.0f * sqrtf(lambda)));
        const float rf = float(r);
        if (u + rf < 0.0f || v + rf < 0.0f || u - rf >= float(cam.width) || v - rf >= float(cam.height))
            break;
        const float tsf = float(ts);
        int tx0 = cvtTruncI32(floorf((u - rf) / tsf));
        int ty0 = cvtTruncI32(floorf((v - rf) / tsf));
        int tx1 = cvtTruncI32(floorf((u + rf) / tsf)) + 1;
        int ty1 = cvtTruncI32(floorf((v + rf) / tsf)) + 1;
        tx0 = tx0 > 0 ? tx0 : 0;
        ty0 = ty0 > 0 ? ty0 : 0;
        tx1 = tx1 < tilesX ? tx1 : tilesX;
        ty1 = ty1 < tilesY ? ty1 : tilesY;
        if (tx0 >= tx1 || ty0 >= ty1)
            break;
        flags = uint8_t((p.clampX ? 1 : 0) | (p.clampY ? 2 : 0));
        float rgb[3];
        const int nb = shTerms<kDeg>(sh);
        float basis[15];
        if (nb) {
            float d[3], len;
            viewDir(mu, campos, d, &len);
            shBasis(d[0], d[1], d[2], kDeg > 0 ? kDeg : sh.degree, basis, nullptr);
        }
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            float col;
            if (nb == 0) {
                col = 0.5f + kShC0 * colors[i * 3 + ch];
            } else {
                float sum = kShC0 * colors[i * 3 + ch];
                for (int k = 0; k < nb; ++k)
                    sum += basis[k] * sh.coef[i * sh.gs + int64_t(k * 3 + ch) * sh.cs];
                col = 0.5f + sum;
            }
            if (col < 0.0f)
                flags |= uint8_t(4 << ch);
            rgb[ch] = (col < 0.0f) ? 0.0f : col; // std::max(col, 0.0f)
        }
        const float alpha = 1.0f / (1.0f + expfExact(-opac[i]));
        out.rec0[o] = make_float4(u, v, p.c * invDet, -p.b * invDet);
        out.rec1[o] = make_float4(p.a * invDet, alpha, rgb[0], rgb[1]);
        out.rec2[o] = rgb[2];
        out.depth[o] = z;
        out.rect[o] = make_int4(tx0, ty0, tx1, ty1);
        radius = r;
        key = floatBits(z);
        nt = uint32_t(tx1 - tx0) * uint32_t(ty1 - ty0);
    } while (false);
    out.radius[o] = radius;
    out.flags[o] = flags;
    out.depthKey[o] = key;
    out.nTiles[o] = nt;
    return nt;
}

template <int kDeg>
__global__ __launch_bounds__(kThreads) void projectKernel(
    int64_t n, const float *__restrict__ means, const float *__restrict__ scales,
    const float *__restrict__ quats, const float *__restrict__ opac,
    const float *__restrict__ colors, u3d_camera cam, float nearPlane, float eps2d, int ts,
    int tilesX, int tilesY, ScreenBuffers out, ShCoefs sh, float3 campos, uint32_t *blockTiles)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    uint32_t nt = 0;
    if (i < n)
        nt = projectBody<kDeg>(i, i, means, scales, quats, opac, colors, cam, nearPlane, eps2d, ts,
                               tilesX, tilesY, out, sh, campos);
    if (!blockTiles)
        return;
    // this block's sum of the covered tiles
    __shared__ uint32_t waveSum[kThreads / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        nt += __shfl_xor(nt, off, 64);
    if ((threadIdx.x & 63) == 0)
        waveSum[threadIdx.x >> 6] = nt;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t sum = 0;
#pragma unroll
        for (int w = 0; w < kThreads / 64; ++w)
            sum += waveSum[w];
        blockTiles[blockIdx.x] = sum;
    }
}

// Batched projection: item blockIdx.y, Gaussian i of its scene -> entry screenBase + i; also
// records the owning item of every entry for the pair emission.
__global__ __launch_bounds__(kThreads) void projectBatchedKernel(const RenderItem *__restrict__ items,
                                                                 float nearPlane, float eps2d, int ts,
                                                                 ScreenBuffers out, int *entryItem)
{
    const RenderItem &it = items[blockIdx.y];
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= it.n)
        return;
    const int64_t o = it.screenBase + i;
    entryItem[o] = int(blockIdx.y);
    ShCoefs sh;
    sh.coef = it.shRest;
    sh.gs = 1;
    sh.cs = it.n;
    sh.degree = it.shRest ? it.shDegree : 0;
    projectBody<-1>(i, o, it.means, it.scales, it.quats, it.opacities, it.colors, it.cam, nearPlane,
                    eps2d, ts, it.tilesX, it.tilesY, out, sh, cameraCentre(it.cam));
}

// Where chainBody adds the parameter gradients of Gaussian i: the model-shaped gradient arrays
// (added to in place) or registers that start at zero (the same sums, term by term).
struct GradArrays {
    float *means, *scales, *quats, *opac, *colors;
    int64_t i;
    __device__ void mean(int k, float v) const { means[i * 3 + k] += v; }
    __device__ void scale(int k, float v) const { scales[i * 3 + k] += v; }
    __device__ void quat(int k, float v) const { quats[i * 4 + k] += v; }
    __device__ void opacity(float v) const { opac[i] += v; }
    __device__ void color(int k, float v) const { colors[i * 3 + k] += v; }
};
struct GradRegs {
    float means[3] = {0.0f, 0.0f, 0.0f}, scales[3] = {0.0f, 0.0f, 0.0f};
    float quats[4] = {0.0f, 0.0f, 0.0f, 0.0f}, opac = 0.0f, colors[3] = {0.0f, 0.0f, 0.0f};
    __device__ void mean(int k, float v) { means[k] += v; }
    __device__ void scale(int k, float v) { scales[k] += v; }
    __device__ void quat(int k, float v) { quats[k] += v; }
    __device__ void opacity(float v) { opac += v; }
    __device__ void color(int k, float v) { colors[k] += v; }
};

// Backward of the projection for one Gaussian (the quantities of Rasterizer.cpp:432-544, derived
// here in matrix form). With U = T M (the Gaussian's axes after the camera and the local affine
// map T = J W; M = R(q) diag(exp s)), the screen covariance is S = U U^T + eps2d I and the conic
// is Q = S^-1. Backward:
//   dS = -Q dQ Q                 (dQ symmetric; the off-diagonal conic entry B counts twice)
//   dU = 2 dS U,  dT = dU M^T,  dM = T^T dU
//   d log s_k = s_k sum_r dM[r][k] R[r][k],  dR = dM diag(s)
//   unit quaternion (w, v): R = (1 - 2|v|^2) I + 2 v v^T + 2 w [v]x, so with
//   a = (dR21 - dR12, dR02 - dR20, dR10 - dR01):
//     dw = 2 v.a,  dv = 2 (dR + dR^T) v - 4 tr(dR) v + 2 w a
//   raw quaternion: dq = (dq_unit - (dq_unit . q_unit) q_unit) / |q|
//   T = J W: dJ = dT W^T; J's entries fx/z, -fx tx/z^2 (and y) give the camera-space mean
//   (tx = x, or z * bound when clamped), plus the projected mean u = fx x/z + cx;
//   world mean: d mu = W^T d pc.

// Small row-major 2x3 helpers of the backward.
struct Rows2x3 {
    float r[2][3];
};

__device__ __forceinline__ float dot3(const float *a, const float *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Chain rule for Gaussian i of a model whose screen state / 2D gradients sit at entry o. The
// SH coefficients' gradients go to sh.grad, or (sh.m set) their Adam update is applied in place.
template <int kDeg, class Sink>
__device__ __forceinline__ void chainBody(
    int64_t i, int64_t o, const float *__restrict__ means, const float *__restrict__ scales,
    const float *__restrict__ quats, u3d_camera cam, float nearPlane, float eps2d,
    const int *__restrict__ radius, const uint8_t *__restrict__ flagsArr,
    const float4 *__restrict__ rec1, const float *__restrict__ grad2d, Sink &out, float *ndc,
    int *gradCount, const ShCoefs &sh, float3 campos, float ndcScale)
{
    if (radius[o] == 0)
        return;
    if (gradCount && radius[o] > 0)
        gradCount[i] += 1;
    float g[kGrad2dStride];
#pragma unroll
    for (int k = 0; k < kGrad2dStride; ++k)
        g[k] = grad2d[o * kGrad2dStride + k];
    const float du = g[0], dv = g[1];
    if (ndc) {
        // densification statistic: |dL/d mean2d| in NDC units
        const float hx = du * 0.5f * float(cam.width), hy = dv * 0.5f * float(cam.height);
        ndc[i] += sqrtf(hx * hx + hy * hy) * ndcScale;
    }
    const uint8_t flags = flagsArr[o];
    // colour: rgb = max(0.5 + C0 c, 0); the clamp stops the gradient
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
        if (!(flags & (4 << ch)))
            out.color(ch, kShC0 * g[5 + ch]);
    // opacity: alpha = sigmoid(o), d alpha / d o = alpha (1 - alpha)
    const float alpha = rec1[o].y;
    out.opacity(g[8] * alpha * (1.0f - alpha));

    const float mu[3] = {means[i * 3], means[i * 3 + 1], means[i * 3 + 2]};
    // View-dependent colour: coefficient gradients (or their fused Adam update) and the
    // gradient w.r.t. the unit view direction, from the coefficients before the update.
    const int nb = shTerms<kDeg>(sh);
    float dDir[3] = {0.0f, 0.0f, 0.0f}, viewD[3] = {0.0f, 0.0f, 0.0f}, viewLen = 1.0f;
    if (nb) {
        float b[15];
        float3 gb[15];
        viewDir(mu, campos, viewD, &viewLen);
        shBasis(viewD[0], viewD[1], viewD[2], kDeg > 0 ? kDeg : sh.degree, b, gb);
        // One channel at a time: its coefficients (and moments) are loaded together before
        // any of them is stored, so the loads of a channel overlap instead of each waiting for
        // the previous coefficient's stores (the arrays could alias as far as the compiler
        // knows).
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const bool on = !(flags & (4 << ch)); // a clamped channel passes no gradient
            const float gc = g[5 + ch];
            float cf[15];
#pragma unroll
            for (int k = 0; k < 15; ++k)
                if (k < nb)
                    cf[k] = sh.coef[i * sh.gs + int64_t(k * 3 + ch) * sh.cs];
#pragma unroll
            for (int k = 0; k < 15; ++k)
                if (k < nb && on) {
                    const float w = cf[k] * gc;
                    dDir[0] += w * gb[k].x;
                    dDir[1] += w * gb[k].y;
                    dDir[2] += w * gb[k].z;
                }
            if (sh.m) {
                float mm[15], vv[15];
#pragma unroll
                for (int k = 0; k < 15; ++k)
                    if (k < nb) {
                        const int64_t idx = i * sh.gs + int64_t(k * 3 + ch) * sh.cs;
                        mm[k] = sh.m[idx];
                        vv[k] = sh.v[idx];
                    }
#pragma unroll
                for (int k = 0; k < 15; ++k)
                    if (k < nb) {
                        const int64_t idx = i * sh.gs + int64_t(k * 3 + ch) * sh.cs;
                        float pp = cf[k];
                        adamScalar(pp, mm[k], vv[k], on ? b[k] * gc : 0.0f, sh.stepSize, sh.invSqrtBc2);
                        sh.coef[idx] = pp;
                        sh.m[idx] = mm[k];
                        sh.v[idx] = vv[k];
                    }
            } else if (sh.grad && on) {
#pragma unroll
                for (int k = 0; k < 15; ++k)
                    if (k < nb)
                        sh.grad[i * sh.gs + int64_t(k * 3 + ch) * sh.cs] += b[k] * gc;
            }
        }
    }
    const float qr[4] = {quats[i * 4], quats[i * 4 + 1], quats[i * 4 + 2], quats[i * 4 + 3]};
    const float ls[3] = {scales[i * 3], scales[i * 3 + 1], scales[i * 3 + 2]};
    Projection p;
    projectGaussian(mu, qr, ls, cam, nearPlane, eps2d, p);

    // dS = -Q dQ Q with Q = S^-1 = [[c, -b], [-b, a]] / det
    const float det = p.a * p.c - p.b * p.b;
    const float q00 = p.c / det, q01 = -p.b / det, q11 = p.a / det;
    const float h00 = g[2], h01 = 0.5f * g[3], h11 = g[4]; // dQ as a symmetric matrix
    // P = dQ Q, then dS = -Q P
    const float p00 = h00 * q00 + h01 * q01, p01 = h00 * q01 + h01 * q11;
    const float p10 = h01 * q00 + h11 * q01, p11 = h01 * q01 + h11 * q11;
    const float s00 = -(q00 * p00 + q01 * p10);
    const float s01 = -(q00 * p01 + q01 * p11);
    const float s11 = -(q01 * p01 + q11 * p11);

    // U = T M (rows: the two screen axes of the Gaussian's three principal directions)
    Rows2x3 T, U, dU, dT;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        T.r[0][k] = p.T0[k];
        T.r[1][k] = p.T1[k];
    }
#pragma unroll
    for (int row = 0; row < 2; ++row)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            U.r[row][k] = T.r[row][0] * p.M[k] + T.r[row][1] * p.M[3 + k] + T.r[row][2] * p.M[6 + k];
    // dU = 2 dS U
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        dU.r[0][k] = 2.0f * (s00 * U.r[0][k] + s01 * U.r[1][k]);
        dU.r[1][k] = 2.0f * (s01 * U.r[0][k] + s11 * U.r[1][k]);
    }
    // dT = dU M^T
#pragma unroll
    for (int row = 0; row < 2; ++row)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            dT.r[row][k] = dot3(dU.r[row], &p.M[k * 3]);
    // dM = T^T dU, folded straight into the scale and rotation gradients
    float dR[9];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float along = 0.0f; // sum_r dM[r][k] R[r][k]
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const float dm = T.r[0][r] * dU.r[0][k] + T.r[1][r] * dU.r[1][k];
            along += dm * p.Rq[r * 3 + k];
            dR[r * 3 + k] = dm * p.s[k];
        }
        out.scale(k, along * p.s[k]);
    }
    // rotation -> unit quaternion (vector form), then the normalisation's projection
    const float qw = p.q[0];
    const float v[3] = {p.q[1], p.q[2], p.q[3]};
    const float ax[3] = {dR[7] - dR[5], dR[2] - dR[6], dR[3] - dR[1]};
    const float tr = dR[0] + dR[4] + dR[8];
    float dq[4];
    dq[0] = 2.0f * dot3(v, ax);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        // ((dR + dR^T) v)_k
        const float sym = (dR[k * 3] + dR[k]) * v[0] + (dR[k * 3 + 1] + dR[3 + k]) * v[1]
                          + (dR[k * 3 + 2] + dR[6 + k]) * v[2];
        dq[1 + k] = 2.0f * sym - 4.0f * tr * v[k] + 2.0f * qw * ax[k];
    }
    const float radial = dq[0] * p.q[0] + dq[1] * p.q[1] + dq[2] * p.q[2] + dq[3] * p.q[3];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        out.quat(k, (dq[k] - radial * p.q[k]) / p.qNorm);

    // camera-space mean through J (dJ = dT W^T) and the projected centre
    const float *W = cam.R;
    const float z = p.pc[2];
    const float iz = 1.0f / z, iz2 = iz * iz;
    const float gJ00 = dot3(dT.r[0], &W[0]), gJ02 = dot3(dT.r[0], &W[6]);
    const float gJ11 = dot3(dT.r[1], &W[3]), gJ12 = dot3(dT.r[1], &W[6]);
    const float gtx = -cam.fx * iz2 * gJ02; // d/d tx of J02 = -fx tx / z^2
    const float gty = -cam.fy * iz2 * gJ12;
    float dpc[3];
    dpc[0] = cam.fx * iz * du;
    dpc[1] = cam.fy * iz * dv;
    dpc[2] = -(cam.fx * p.pc[0] * du + cam.fy * p.pc[1] * dv) * iz2
             - cam.fx * iz2 * gJ00 - cam.fy * iz2 * gJ11
             + 2.0f * iz2 * iz * (cam.fx * p.tx * gJ02 + cam.fy * p.ty * gJ12);
    if (p.clampX)
        dpc[2] += gtx * p.limTx;
    else
        dpc[0] += gtx;
    if (p.clampY)
        dpc[2] += gty * p.limTy;
    else
        dpc[1] += gty;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        out.mean(k, W[k] * dpc[0] + W[3 + k] * dpc[1] + W[6 + k] * dpc[2]);
    if (nb) {
        // d = (mu - c) / |mu - c|: project out the radial part
        const float along = dot3(dDir, viewD);
#pragma unroll
        for (int k = 0; k < 3; ++k)
            out.mean(k, (dDir[k] - along * viewD[k]) / viewLen);
    }
}

// clear: the 2D gradients are zeroed after use (only visible Gaussians' entries are ever
// written, so the buffer is all zero again for the next backward).
template <int kDeg>
__global__ __launch_bounds__(kThreads) void chainRuleKernel(
    int64_t n, const float *__restrict__ means, const float *__restrict__ scales,
    const float *__restrict__ quats, u3d_camera cam, float nearPlane, float eps2d,
    const int *__restrict__ radius, const uint8_t *__restrict__ flagsArr,
    const float4 *__restrict__ rec1, float *grad2d, float *gMeans,
    float *gScales, float *gQuats, float *gOpac, float *gColors, float *ndc, int *gradCount,
    ShCoefs sh, float3 campos, bool clear)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n)
        return;
    GradArrays out{gMeans, gScales, gQuats, gOpac, gColors, i};
    chainBody<kDeg>(i, i, means, scales, quats, cam, nearPlane, eps2d, radius, flagsArr, rec1,
                    grad2d, out, ndc, gradCount, sh, campos, 1.0f);
    if (clear && radius[i] != 0) {
#pragma unroll
        for (int k = 0; k < kGrad2dStride; ++k)
            grad2d[i * kGrad2dStride + k] = 0.0f;
    }
}

// Chain rule of Gaussian i into registers, then the Adam update of its 14 parameters from them.
// The registers start at zero and take the same terms in the same order as the zeroed gradient
// arrays would, so the update equals chainRuleKernel + adamKernel bit for bit.
template <int kDeg>
__global__ __launch_bounds__(kThreads) void chainAdamKernel(
    int64_t n, u3d_gaussians prm, u3d_camera cam, float nearPlane, float eps2d,
    const int *__restrict__ radius, const uint8_t *__restrict__ flagsArr,
    const float4 *__restrict__ rec1, float *grad2d, float *ndc, int *gradCount, ShCoefs sh,
    float3 campos, AdamStepArgs adam)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n)
        return;
    GradRegs gr;
    chainBody<kDeg>(i, i, prm.means, prm.scales, prm.quats, cam, nearPlane, eps2d, radius, flagsArr,
                    rec1, grad2d, gr, ndc, gradCount, sh, campos, 1.0f);
    if (radius[i] != 0) {
#pragma unroll
        for (int k = 0; k < kGrad2dStride; ++k)
            grad2d[i * kGrad2dStride + k] = 0.0f;
    } else if (kDeg != 0 && sh.m) {
        // not visible: its coefficients still take the step with a zero gradient
        const int nb = shTerms<kDeg>(sh);
        for (int c0 = 0; c0 < 3 * nb; c0 += 15) { // 15 coefficients' loads before their stores
            float pp[15], mm[15], vv[15];
#pragma unroll
            for (int k = 0; k < 15; ++k)
                if (c0 + k < 3 * nb) {
                    const int64_t idx = i * sh.gs + int64_t(c0 + k) * sh.cs;
                    pp[k] = sh.coef[idx];
                    mm[k] = sh.m[idx];
                    vv[k] = sh.v[idx];
                }
#pragma unroll
            for (int k = 0; k < 15; ++k)
                if (c0 + k < 3 * nb) {
                    const int64_t idx = i * sh.gs + int64_t(c0 + k) * sh.cs;
                    adamScalar(pp[k], mm[k], vv[k], 0.0f, sh.stepSize, sh.invSqrtBc2);
                    sh.coef[idx] = pp[k];
                    sh.m[idx] = mm[k];
                    sh.v[idx] = vv[k];
                }
        }
    }
    const float inv = adam.invSqrtBc2;
    // each group's parameter, first and second moment as one 12- or 16-byte access
    struct F3 {
        float x, y, z;
    };
    auto adam3 = [&](float *p, float *m, float *v, const float *g, float step) {
        F3 pp = *reinterpret_cast<const F3 *>(p + i * 3);
        F3 mm = *reinterpret_cast<const F3 *>(m + i * 3);
        F3 vv = *reinterpret_cast<const F3 *>(v + i * 3);
        adamScalar(pp.x, mm.x, vv.x, g[0], step, inv);
        adamScalar(pp.y, mm.y, vv.y, g[1], step, inv);
        adamScalar(pp.z, mm.z, vv.z, g[2], step, inv);
        *reinterpret_cast<F3 *>(p + i * 3) = pp;
        *reinterpret_cast<F3 *>(m + i * 3) = mm;
        *reinterpret_cast<F3 *>(v + i * 3) = vv;
    };
    adam3(prm.means, adam.m.means, adam.v.means, gr.means, adam.stepSize[0]);
    adam3(prm.scales, adam.m.scales, adam.v.scales, gr.scales, adam.stepSize[1]);
    {
        float4 pp = reinterpret_cast<const float4 *>(prm.quats)[i];
        float4 mm = reinterpret_cast<const float4 *>(adam.m.quats)[i];
        float4 vv = reinterpret_cast<const float4 *>(adam.v.quats)[i];
        adamScalar(pp.x, mm.x, vv.x, gr.quats[0], adam.stepSize[2], inv);
        adamScalar(pp.y, mm.y, vv.y, gr.quats[1], adam.stepSize[2], inv);
        adamScalar(pp.z, mm.z, vv.z, gr.quats[2], adam.stepSize[2], inv);
        adamScalar(pp.w, mm.w, vv.w, gr.quats[3], adam.stepSize[2], inv);
        reinterpret_cast<float4 *>(prm.quats)[i] = pp;
        reinterpret_cast<float4 *>(adam.m.quats)[i] = mm;
        reinterpret_cast<float4 *>(adam.v.quats)[i] = vv;
    }
    adamScalar(prm.opacities[i], adam.m.opacities[i], adam.v.opacities[i], gr.opac,
               adam.stepSize[3], inv);
    adam3(prm.colors, adam.m.colors, adam.v.colors, gr.colors, adam.stepSize[4]);
}

// Batched chain rule over scene groups: Gaussian i of scene blockIdx.y takes the 2D gradients
// of each of the scene's items in item order, then clears the entries it consumed (so the next
// step's backward accumulates into zeros without a separate clearing pass).
__global__ __launch_bounds__(kThreads) void chainBatchedKernel(const RenderItem *__restrict__ items,
                                                               const int2 *__restrict__ groups,
                                                               float nearPlane, float eps2d,
                                                               ScreenBuffers sb, float *grad2d)
{
    const int2 grp = groups[blockIdx.y];
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (grp.y <= 0 || i >= items[grp.x].n)
        return;
    for (int v = grp.x; v < grp.x + grp.y; ++v) {
        const RenderItem &it = items[v];
        const int64_t o = it.screenBase + i;
        if (sb.radius[o] == 0)
            continue;
        GradArrays out{it.gMeans, it.gScales, it.gQuats, it.gOpac, it.gColors, i};
        ShCoefs sh;
        sh.coef = it.shRest;
        sh.grad = it.gShRest;
        sh.gs = 1;
        sh.cs = it.n;
        sh.degree = it.shRest ? it.shDegree : 0;
        chainBody<-1>(i, o, it.means, it.scales, it.quats, it.cam, nearPlane, eps2d, sb.radius,
                      sb.flags, sb.rec1, grad2d, out, it.ndc, it.gradCount, sh,
                      cameraCentre(it.cam), it.ndcScale);
#pragma unroll
        for (int k = 0; k < kGrad2dStride; ++k)
            grad2d[o * kGrad2dStride + k] = 0.0f;
    }
}

// Batched chain rule fused with Adam, for scene groups without view-dependent colour. A block
// holds 64 Gaussians of scene group blockIdx.y; wavefront w takes the group's views w,
// w + kWaves, ... (a lane runs one (Gaussian, view) pair at a time, so a group of B views keeps
// up to B times the wavefronts of one view busy), adds the parameter gradients in registers and
// clears the 2D gradients it consumed. The wavefronts' sums meet in LDS and are added in
// wavefront order (a fixed order: the result does not depend on timing); each of the block's
// 64 x 14 parameters then takes adamBatchedKernel's update from its summed gradient, so the
// gradients never pass through memory and the gradient arrays stay zero. The densification
// statistics (|d mean2d| in NDC units, views that saw the Gaussian) are summed the same way.
template <int kWaves>
__global__ __launch_bounds__(64 * kWaves) void chainAdamViewsKernel(
    const RenderItem *__restrict__ items, const int2 *__restrict__ groups,
    const SceneStep *__restrict__ scenes, float nearPlane, float eps2d, ScreenBuffers sb,
    float *grad2d)
{
    constexpr int kVals = 16; // 14 parameter gradients, the NDC statistic, the views seen
    __shared__ float part[kWaves][kVals][64];
    const int2 grp = groups[blockIdx.y];
    const SceneStep &sc = scenes[blockIdx.y];
    const int64_t n = sc.n, base = int64_t(blockIdx.x) * 64;
    if (grp.y <= 0 || base >= n)
        return;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    const int64_t i = base + lane;
    const RenderItem &first = items[grp.x];
    float *const ndc = first.ndc;
    int *const gradCount = first.gradCount;
    GradRegs gr;
    float ndcSum = 0.0f, seen = 0.0f;
    if (i < n) {
        for (int v = grp.x + wave; v < grp.x + grp.y; v += kWaves) {
            const RenderItem &it = items[v];
            const int64_t o = it.screenBase + i;
            const int r = sb.radius[o];
            if (r == 0)
                continue;
            if (r > 0)
                seen += 1.0f;
            if (ndc) {
                const float hx = grad2d[o * kGrad2dStride] * 0.5f * float(it.cam.width);
                const float hy = grad2d[o * kGrad2dStride + 1] * 0.5f * float(it.cam.height);
                ndcSum += sqrtf(hx * hx + hy * hy) * it.ndcScale;
            }
            ShCoefs sh;
            chainBody<0>(i, o, it.means, it.scales, it.quats, it.cam, nearPlane, eps2d, sb.radius,
                         sb.flags, sb.rec1, grad2d, gr, nullptr, nullptr, sh,
                         cameraCentre(it.cam), it.ndcScale);
#pragma unroll
            for (int k = 0; k < kGrad2dStride; ++k)
                grad2d[o * kGrad2dStride + k] = 0.0f;
        }
    }
    // slots: means 0-2, scales 3-5, quats 6-9, opacity 10, colours 11-13, NDC 14, seen 15
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        part[wave][k][lane] = gr.means[k];
        part[wave][3 + k][lane] = gr.scales[k];
        part[wave][11 + k][lane] = gr.colors[k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        part[wave][6 + k][lane] = gr.quats[k];
    part[wave][10][lane] = gr.opac;
    part[wave][14][lane] = ndcSum;
    part[wave][15][lane] = seen;
    __syncthreads();

    // The block's parameters in the scene's layout [means 3n | scales 3n | quats 4n | opacity n |
    // colours 3n] (each group at its capacity offset): value f of the block is group g's entry
    // q = f - start[g], Gaussian q / comps, component q % comps.
    const int64_t c = sc.cap;
    const int start[6] = {0, 192, 384, 640, 704, 896};
    const int comps[5] = {3, 3, 4, 1, 3};
    const int slot0[5] = {0, 3, 6, 10, 11};
    const int64_t off[5] = {0, 3 * c, 6 * c, 10 * c, 11 * c};
    for (int f = threadIdx.x; f < 896; f += 64 * kWaves) {
        int g = 0;
#pragma unroll
        for (int k = 1; k < 5; ++k)
            g += f >= start[k] ? 1 : 0;
        const int q = f - start[g], gi = q / comps[g], comp = q - gi * comps[g];
        if (base + gi >= n)
            continue;
        float grad = part[0][slot0[g] + comp][gi];
#pragma unroll
        for (int w = 1; w < kWaves; ++w)
            grad += part[w][slot0[g] + comp][gi];
        const int64_t idx = off[g] + (base + gi) * comps[g] + comp;
        float *__restrict__ p = sc.params;
        float *__restrict__ m = sc.adamM;
        float *__restrict__ v = sc.adamV;
        const float mk = sc.beta1 * m[idx] + (1.0f - sc.beta1) * grad;
        const float vk = sc.beta2 * v[idx] + (1.0f - sc.beta2) * grad * grad;
        m[idx] = mk;
        v[idx] = vk;
        p[idx] = p[idx] - sc.stepSize[g] * mk / (sqrtf(vk) * sc.invSqrtBc2 + 1e-15f);
    }
    if (threadIdx.x < 64 && i < n) {
        float stat = part[0][14][lane], views = part[0][15][lane];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) {
            stat += part[w][14][lane];
            views += part[w][15][lane];
        }
        if (ndc)
            ndc[i] += stat;
        if (gradCount)
            gradCount[i] += int(views);
    }
}

__global__ __launch_bounds__(kThreads) void unpackKernel(int64_t n, ScreenBuffers sb,
                                                         float *mean2d, float *conic,
                                                         float *color, float *opacity)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n || sb.radius[i] <= 0)
        return;
    const float4 a = sb.rec0[i], b = sb.rec1[i];
    mean2d[i * 2] = a.x;
    mean2d[i * 2 + 1] = a.y;
    conic[i * 3] = a.z;
    conic[i * 3 + 1] = a.w;
    conic[i * 3 + 2] = b.x;
    opacity[i] = b.y;
    color[i * 3] = b.z;
    color[i * 3 + 1] = b.w;
    color[i * 3 + 2] = sb.rec2[i];
}

inline unsigned blocksFor(int64_t n) { return unsigned((n + kThreads - 1) / kThreads); }

} // namespace

int projectBlocks(int64_t n) { return n > 0 ? int(blocksFor(n)) : 0; }

// The kernel instance for the SH degree of sh (0 without coefficients).
template <template <int> class K>
inline auto byDegree(const ShCoefs &sh) -> decltype(K<0>::fn)
{
    switch (sh.coef ? sh.degree : 0) {
    case 1: return K<1>::fn;
    case 2: return K<2>::fn;
    case 3: return K<3>::fn;
    default: return K<0>::fn;
    }
}
template <int D> struct ProjectK { static constexpr auto fn = projectKernel<D>; };
template <int D> struct ChainRuleK { static constexpr auto fn = chainRuleKernel<D>; };
template <int D> struct ChainAdamK { static constexpr auto fn = chainAdamKernel<D>; };

int launchProject(hipStream_t s, const u3d_gaussians &m, const u3d_camera &cam,
                  const u3d_raster_settings &st, int tilesX, int tilesY, const ScreenBuffers &out,
                  const ShCoefs &sh, uint32_t *blockTiles)
{
    if (m.n <= 0)
        return U3D_OK;
    U3D_CHECK(sh.degree >= 0 && sh.degree <= 3, U3D_ERR_INVALID, "sh_degree must be 0..3");
    hipLaunchKernelGGL(byDegree<ProjectK>(sh), dim3(blocksFor(m.n)), dim3(kThreads), 0, s, m.n, m.means,
                       m.scales, m.quats, m.opacities, m.colors, cam, st.near_plane, st.eps2d,
                       st.tile_size, tilesX, tilesY, out, sh, cameraCentre(cam), blockTiles);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchChainRule(hipStream_t s, const u3d_gaussians &m, const u3d_camera &cam,
                    const u3d_raster_settings &st, const ScreenBuffers &sb, float *grad2d,
                    u3d_gaussians &grads, float *ndc, int *gradCount, const ShCoefs &sh,
                    bool clearGrad2d)
{
    if (m.n <= 0)
        return U3D_OK;
    U3D_CHECK(sh.degree >= 0 && sh.degree <= 3, U3D_ERR_INVALID, "sh_degree must be 0..3");
    hipLaunchKernelGGL(byDegree<ChainRuleK>(sh), dim3(blocksFor(m.n)), dim3(kThreads), 0, s, m.n,
                       m.means, m.scales, m.quats, cam, st.near_plane, st.eps2d, sb.radius,
                       sb.flags, sb.rec1, grad2d, grads.means, grads.scales, grads.quats,
                       grads.opacities, grads.colors, ndc, gradCount, sh, cameraCentre(cam),
                       clearGrad2d);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchChainAdam(hipStream_t s, const u3d_gaussians &params, const u3d_camera &cam,
                    const u3d_raster_settings &st, const ScreenBuffers &sb, float *grad2d,
                    float *ndc, int *gradCount, const ShCoefs &sh, const AdamStepArgs &adam)
{
    if (params.n <= 0)
        return U3D_OK;
    U3D_CHECK(adam.m.n == params.n && adam.v.n == params.n, U3D_ERR_INVALID,
              "Adam moments differ in size from the parameters");
    U3D_CHECK(sh.degree >= 0 && sh.degree <= 3, U3D_ERR_INVALID, "sh_degree must be 0..3");
    hipLaunchKernelGGL(byDegree<ChainAdamK>(sh), dim3(blocksFor(params.n)), dim3(kThreads), 0, s, params.n,
                       params, cam, st.near_plane, st.eps2d, sb.radius, sb.flags, sb.rec1, grad2d,
                       ndc, gradCount, sh, cameraCentre(cam), adam);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchProjectBatched(hipStream_t s, const RenderItem *items, int numItems, int64_t maxN,
                         const u3d_raster_settings &st, const ScreenBuffers &out, int *entryItem)
{
    if (numItems <= 0 || maxN <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(projectBatchedKernel, dim3(blocksFor(maxN), numItems), dim3(kThreads), 0, s,
                       items, st.near_plane, st.eps2d, st.tile_size, out, entryItem);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchChainBatched(hipStream_t s, const RenderItem *items, const int2 *groups, int numGroups,
                       int64_t maxN, const u3d_raster_settings &st, const ScreenBuffers &sb,
                       float *grad2d)
{
    if (numGroups <= 0 || maxN <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(chainBatchedKernel, dim3(blocksFor(maxN), numGroups), dim3(kThreads), 0, s,
                       items, groups, st.near_plane, st.eps2d, sb, grad2d);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchChainAdamBatched(hipStream_t s, const RenderItem *items, const int2 *groups,
                           const SceneStep *scenes, int numGroups, int64_t maxN, int maxViews,
                           const u3d_raster_settings &st, const ScreenBuffers &sb, float *grad2d)
{
    if (numGroups <= 0 || maxN <= 0)
        return U3D_OK;
    const dim3 grid(unsigned((maxN + 63) / 64), unsigned(numGroups));
    if (maxViews >= 4)
        hipLaunchKernelGGL(chainAdamViewsKernel<4>, grid, dim3(256), 0, s, items, groups, scenes,
                           st.near_plane, st.eps2d, sb, grad2d);
    else if (maxViews >= 2)
        hipLaunchKernelGGL(chainAdamViewsKernel<2>, grid, dim3(128), 0, s, items, groups, scenes,
                           st.near_plane, st.eps2d, sb, grad2d);
    else
        hipLaunchKernelGGL(chainAdamViewsKernel<1>, grid, dim3(64), 0, s, items, groups, scenes,
                           st.near_plane, st.eps2d, sb, grad2d);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchUnpackScreen(hipStream_t s, int64_t n, const ScreenBuffers &sb, float *mean2d,
                       float *conic, float *color, float *opacity)
{
    if (n <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(unpackKernel, dim3(blocksFor(n)), dim3(kThreads), 0, s, n, sb, mean2d,
                       conic, color, opacity);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

} // namespace u3d
