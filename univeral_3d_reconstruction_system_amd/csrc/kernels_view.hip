// Viewer support (app/src/SplatView.cpp:117-180, 349-397): per-splat display attributes and the
// front-to-back order of a view. All HBM-bound streaming passes; the order is the stable sort of
// 16-bit depth keys by the project's radix sort (two 8-bit passes). Compiled with contraction
// off: attributes, keys and order equal the reference's host computation bit for bit.
#include "kernels.h"
#include "projection.h"

namespace u3d {

namespace {

constexpr int kThreads = 256;

inline unsigned blocksFor(int64_t n) { return unsigned((n + kThreads - 1) / kThreads); }

// std::lround of a float (half away from zero) for the colour channels in [0, 255]; NaN -> 0.
__device__ __forceinline__ int roundHalfAway(float x)
{
    if (!(x == x))
        return 0;
    const float t = truncf(x);
    const float frac = x - t; // exact for |x| < 2^23
    return int(t) + (frac >= 0.5f ? 1 : 0) - (frac <= -0.5f ? 1 : 0);
}

__device__ __forceinline__ float clamp01(float v)
{
    // std::clamp(v, 0, 1): v < 0 ? 0 : (1 < v ? 1 : v)
    return v < 0.0f ? 0.0f : (1.0f < v ? 1.0f : v);
}

__global__ __launch_bounds__(kThreads) void viewAttributesKernel(u3d_gaussians m, uint8_t *rgba,
                                                                 float *cov6)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= m.n)
        return;
    uchar4 c;
    c.x = uint8_t(roundHalfAway(clamp01(0.5f + kShC0 * m.colors[i * 3]) * 255.0f));
    c.y = uint8_t(roundHalfAway(clamp01(0.5f + kShC0 * m.colors[i * 3 + 1]) * 255.0f));
    c.z = uint8_t(roundHalfAway(clamp01(0.5f + kShC0 * m.colors[i * 3 + 2]) * 255.0f));
    const float alpha = 1.0f / (1.0f + expfExact(-m.opacities[i]));
    c.w = uint8_t(roundHalfAway(alpha * 255.0f));
    reinterpret_cast<uchar4 *>(rgba)[i] = c;
    const float *q = &m.quats[i * 4];
    const float qn = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) + 1e-12f;
    const float w = q[0] / qn, x = q[1] / qn, y = q[2] / qn, z = q[3] / qn;
    float R[9];
    rotOfQuat(w, x, y, z, R);
    float M[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            M[r * 3 + k] = R[r * 3 + k] * expfExact(m.scales[i * 3 + k]);
    auto sigma = [&](int r, int k) {
        return M[r * 3] * M[k * 3] + M[r * 3 + 1] * M[k * 3 + 1] + M[r * 3 + 2] * M[k * 3 + 2];
    };
    float *o = cov6 + i * 6;
    o[0] = sigma(0, 0);
    o[1] = sigma(0, 1);
    o[2] = sigma(0, 2);
    o[3] = sigma(1, 1);
    o[4] = sigma(1, 2);
    o[5] = sigma(2, 2);
}

// Distance in front of the camera: z = -(v . (c, 1)), v the view matrix's third row.
__device__ __forceinline__ float viewDepth(const float *__restrict__ c, int64_t i, float4 v)
{
    return -(v.x * c[i * 3] + v.y * c[i * 3 + 1] + v.z * c[i * 3 + 2] + v.w);
}

// Per block of depths: the smallest and largest (std::min / std::max from +-FLT_MAX: NaN is
// never taken) at range[2 b], range[2 b + 1].
__global__ __launch_bounds__(kThreads) void depthRangeKernel(const float *__restrict__ centers,
                                                             int64_t n, float4 v, float *range)
{
    __shared__ float lo[kThreads / 64], hi[kThreads / 64];
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    float a = 3.402823466e38f, b = -3.402823466e38f;
    if (i < n) {
        const float z = viewDepth(centers, i, v);
        a = (z < a) ? z : a; // std::min(a, z)
        b = (b < z) ? z : b; // std::max(b, z)
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float oa = __shfl_xor(a, off, 64), ob = __shfl_xor(b, off, 64);
        a = (oa < a) ? oa : a;
        b = (b < ob) ? ob : b;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        lo[wave] = a;
        hi[wave] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kThreads / 64; ++k) {
            a = (lo[k] < a) ? lo[k] : a;
            b = (b < hi[k]) ? hi[k] : b;
        }
        range[2 * blockIdx.x] = a;
        range[2 * blockIdx.x + 1] = b;
    }
}

// 16-bit keys uint16((z - lo) * 65535 / (hi - lo)) (0 when hi == lo); every block first reduces
// the block ranges itself (min / max are exact, so any order gives the same lo, hi).
__global__ __launch_bounds__(kThreads) void depthKeysKernel(const float *__restrict__ centers,
                                                            int64_t n, float4 v,
                                                            const float *__restrict__ range,
                                                            int rangeBlocks, uint32_t *keys)
{
    __shared__ float lo4[kThreads / 64], hi4[kThreads / 64];
    float a = 3.402823466e38f, b = -3.402823466e38f;
    for (int k = int(threadIdx.x); k < rangeBlocks; k += kThreads) {
        a = (range[2 * k] < a) ? range[2 * k] : a;
        b = (b < range[2 * k + 1]) ? range[2 * k + 1] : b;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float oa = __shfl_xor(a, off, 64), ob = __shfl_xor(b, off, 64);
        a = (oa < a) ? oa : a;
        b = (b < ob) ? ob : b;
    }
    if ((threadIdx.x & 63) == 0) {
        lo4[threadIdx.x >> 6] = a;
        hi4[threadIdx.x >> 6] = b;
    }
    __syncthreads();
    float sLo = lo4[0], sHi = hi4[0];
    for (int k = 1; k < kThreads / 64; ++k) {
        sLo = (lo4[k] < sLo) ? lo4[k] : sLo;
        sHi = (sHi < hi4[k]) ? hi4[k] : sHi;
    }
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n)
        return;
    const float lo = sLo, hi = sHi;
    const float scale = hi > lo ? 65535.0f / (hi - lo) : 0.0f;
    const float z = viewDepth(centers, i, v);
    keys[i] = uint32_t(uint16_t(cvtTruncI32((z - lo) * scale)));
}

struct PackedRecord {
    float center[3];
    uint8_t color[4];
    float cov[6];
};
static_assert(sizeof(PackedRecord) == 40, "40-byte records");

__global__ __launch_bounds__(kThreads) void packViewKernel(const uint32_t *__restrict__ order,
                                                           int64_t n, const float *__restrict__ centers,
                                                           const uint8_t *__restrict__ rgba,
                                                           const float *__restrict__ cov6,
                                                           PackedRecord *out)
{
    const int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (k >= n)
        return;
    const int64_t i = order[k];
    PackedRecord r;
#pragma unroll
    for (int j = 0; j < 3; ++j)
        r.center[j] = centers[i * 3 + j];
    const uchar4 c = reinterpret_cast<const uchar4 *>(rgba)[i];
    r.color[0] = c.x;
    r.color[1] = c.y;
    r.color[2] = c.z;
    r.color[3] = c.w;
#pragma unroll
    for (int j = 0; j < 6; ++j)
        r.cov[j] = cov6[i * 6 + j];
    out[k] = r;
}

} // namespace

int launchViewAttributes(hipStream_t s, const u3d_gaussians &m, uint8_t *rgba, float *cov6)
{
    if (m.n <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(viewAttributesKernel, dim3(blocksFor(m.n)), dim3(kThreads), 0, s, m, rgba, cov6);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int viewRangeBlocks(int64_t n) { return n > 0 ? int(blocksFor(n)) : 0; }

int launchDepthKeys(hipStream_t s, const float *centers, int64_t n, const float view_row2[4],
                    float *range, uint32_t *keys)
{
    if (n <= 0)
        return U3D_OK;
    const float4 v = make_float4(view_row2[0], view_row2[1], view_row2[2], view_row2[3]);
    const int rb = viewRangeBlocks(n);
    hipLaunchKernelGGL(depthRangeKernel, dim3(rb), dim3(kThreads), 0, s, centers, n, v, range);
    U3D_LAUNCH_CHECK();
    hipLaunchKernelGGL(depthKeysKernel, dim3(blocksFor(n)), dim3(kThreads), 0, s, centers, n, v,
                       range, rb, keys);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchPackView(hipStream_t s, const uint32_t *order, int64_t n, const float *centers,
                   const uint8_t *rgba, const float *cov6, void *packed)
{
    if (n <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(packViewKernel, dim3(blocksFor(n)), dim3(kThreads), 0, s, order, n, centers,
                       rgba, cov6, static_cast<PackedRecord *>(packed));
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

} // namespace u3d
