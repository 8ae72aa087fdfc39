// Seed-point kernels: the kNN-distance initialisation (initGaussiansFromPoints,
// GaussianModel.cpp:68-144) and the learned initialiser's neighbourhood features.
//
// The reference buckets points in a hash map keyed by
//   key(x, y, z) = (int64(x) * 73856093) ^ (int64(y) * 19349663) ^ (int64(z) * 83492791)
// of the integer cell, and searches the buckets of the 27 cells around a point. Cells whose keys
// collide share a bucket, and a bucket reached from two of the 27 offsets is scanned twice. The
// device grid keeps exactly that: points sorted by the 64-bit key, a bucket is the run of equal
// keys found by binary search, and the 27 offsets are visited in the reference's order. The 3
// smallest squared distances are kept with the reference's insertion (swap) network, so the
// mean squared distance is bit-identical.
#include "kernels.h"

namespace u3d {

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ uint64_t cellKey(int x, int y, int z)
{
    const int64_t k = (int64_t(x) * 73856093) ^ (int64_t(y) * 19349663) ^ (int64_t(z) * 83492791);
    return uint64_t(k);
}

__device__ __forceinline__ void cellOf(const float *__restrict__ xyz, int64_t i, float3 lo,
                                       float cell, int c[3])
{
    c[0] = int(floorf((xyz[i * 3] - lo.x) / cell));
    c[1] = int(floorf((xyz[i * 3 + 1] - lo.y) / cell));
    c[2] = int(floorf((xyz[i * 3 + 2] - lo.z) / cell));
}

// First index in the sorted keys whose key is >= k (keys as (hi, lo) 32-bit halves).
__device__ __forceinline__ int64_t lowerBound(const uint32_t *__restrict__ hi,
                                              const uint32_t *__restrict__ lo, int64_t n, uint64_t k)
{
    int64_t a = 0, b = n;
    while (a < b) {
        const int64_t m = (a + b) >> 1;
        const uint64_t km = (uint64_t(hi[m]) << 32) | uint64_t(lo[m]);
        if (km < k)
            a = m + 1;
        else
            b = m;
    }
    return a;
}

__global__ __launch_bounds__(kThreads) void boundsPartial(const float *__restrict__ xyz, int64_t n,
                                                          float *partial)
{
    __shared__ float s[6][kThreads];
    float v[6] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f,
                  -3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kThreads)
        for (int k = 0; k < 3; ++k) {
            const float x = xyz[i * 3 + k];
            v[k] = fminf(v[k], x);
            v[3 + k] = fmaxf(v[3 + k], x);
        }
    for (int k = 0; k < 6; ++k)
        s[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if (int(threadIdx.x) < off)
            for (int k = 0; k < 6; ++k)
                s[k][threadIdx.x] = k < 3 ? fminf(s[k][threadIdx.x], s[k][threadIdx.x + off])
                                          : fmaxf(s[k][threadIdx.x], s[k][threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x < 6)
        partial[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void boundsFinal(const float *__restrict__ partial, int blocks, float *out)
{
    const int k = threadIdx.x;
    if (k >= 6)
        return;
    float v = partial[k];
    for (int b = 1; b < blocks; ++b)
        v = k < 3 ? fminf(v, partial[b * 6 + k]) : fmaxf(v, partial[b * 6 + k]);
    out[k] = v;
}

__global__ __launch_bounds__(kThreads) void cellKeysKernel(const float *__restrict__ xyz, int64_t n,
                                                           float3 lo, float cell, uint32_t *keyLo,
                                                           uint32_t *keyHi, uint32_t *idx)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n)
        return;
    int c[3];
    cellOf(xyz, i, lo, cell, c);
    const uint64_t k = cellKey(c[0], c[1], c[2]);
    keyLo[i] = uint32_t(k);
    keyHi[i] = uint32_t(k >> 32);
    idx[i] = uint32_t(i);
}

// initGaussiansFromPoints per point (GaussianModel.cpp:103-141).
__global__ __launch_bounds__(kThreads) void knnInitKernel(
    const float *__restrict__ xyz, const float *__restrict__ rgb, int rgbCount, int64_t n, float3 lo,
    float cell, const uint32_t *__restrict__ sHi, const uint32_t *__restrict__ sLo,
    const uint32_t *__restrict__ sIdx, float opacityLogit, float *means, float *quats, float *opac,
    float *colors, float *meanD2Out)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n)
        return;
    const float fmax = 3.402823466e38f;
    float best[3] = {fmax, fmax, fmax};
    int c[3];
    cellOf(xyz, i, lo, cell, c);
    const float xi = xyz[i * 3], yi = xyz[i * 3 + 1], zi = xyz[i * 3 + 2];
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                const uint64_t k = cellKey(c[0] + dx, c[1] + dy, c[2] + dz);
                for (int64_t p = lowerBound(sHi, sLo, n, k);
                     p < n && ((uint64_t(sHi[p]) << 32) | uint64_t(sLo[p])) == k; ++p) {
                    const int64_t j = int64_t(sIdx[p]);
                    if (j == i)
                        continue;
                    const float ex = xyz[j * 3] - xi, ey = xyz[j * 3 + 1] - yi, ez = xyz[j * 3 + 2] - zi;
                    float d2 = ex * ex + ey * ey + ez * ez;
#pragma unroll
                    for (int q = 0; q < 3; ++q)
                        if (d2 < best[q]) {
                            const float t = best[q];
                            best[q] = d2;
                            d2 = t;
                        }
                }
            }
    float sum = 0.0f;
    int count = 0;
#pragma unroll
    for (int q = 0; q < 3; ++q)
        if (best[q] < fmax) {
            sum += best[q];
            ++count;
        }
    const float meanD2 = count ? sum / float(count) : cell * cell;
    meanD2Out[i] = meanD2;
    // The log-scale std::log(std::sqrt(max(meanD2, 1e-14f))) is applied by the host with its
    // libm (launchKnnInit's caller), so it matches the reference bit for bit.
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        means[i * 3 + k] = xyz[i * 3 + k];
        const float col = rgbCount >= (i + 1) * 3 ? rgb[i * 3 + k] : 0.5f;
        colors[i * 3 + k] = (col - 0.5f) / kShC0;
    }
    quats[i * 4] = 1.0f;
    quats[i * 4 + 1] = 0.0f;
    quats[i * 4 + 2] = 0.0f;
    quats[i * 4 + 3] = 0.0f;
    opac[i] = opacityLogit;
}

// Learned-initialiser input of one seed point (kFeatureDim floats, zero padded), from its 16
// nearest neighbours among the same 27 buckets: per neighbour (nearest first) the offset and
// distance in units of the point's kNN scale s0 = sqrt(meanD2) (clamped to +-8), then the
// point's colour - 0.5 and the fraction of cameras that see it.
__global__ __launch_bounds__(kThreads) void knnFeaturesKernel(
    const float *__restrict__ xyz, const float *__restrict__ rgb, int rgbCount, int64_t n, float3 lo,
    float cell, const uint32_t *__restrict__ sHi, const uint32_t *__restrict__ sLo,
    const uint32_t *__restrict__ sIdx, const float *__restrict__ meanD2,
    const float *__restrict__ viewFrac, float *features)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n)
        return;
    constexpr int kK = kFeatureNeighbours;
    float bd[kK];
    uint32_t bj[kK];
#pragma unroll
    for (int q = 0; q < kK; ++q) {
        bd[q] = 3.402823466e38f;
        bj[q] = 0xffffffffu;
    }
    int c[3];
    cellOf(xyz, i, lo, cell, c);
    const float xi = xyz[i * 3], yi = xyz[i * 3 + 1], zi = xyz[i * 3 + 2];
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                const uint64_t k = cellKey(c[0] + dx, c[1] + dy, c[2] + dz);
                for (int64_t p = lowerBound(sHi, sLo, n, k);
                     p < n && ((uint64_t(sHi[p]) << 32) | uint64_t(sLo[p])) == k; ++p) {
                    const uint32_t j = sIdx[p];
                    if (int64_t(j) == i)
                        continue;
                    const float ex = xyz[size_t(j) * 3] - xi, ey = xyz[size_t(j) * 3 + 1] - yi,
                                ez = xyz[size_t(j) * 3 + 2] - zi;
                    float d2 = ex * ex + ey * ey + ez * ez;
                    if (!(d2 < bd[kK - 1]))
                        continue;
                    bool dup = false; // a bucket scanned twice must not fill two slots
#pragma unroll
                    for (int q = 0; q < kK; ++q)
                        dup |= bj[q] == j;
                    if (dup)
                        continue;
                    uint32_t jj = j;
#pragma unroll
                    for (int q = 0; q < kK; ++q)
                        if (d2 < bd[q]) {
                            const float t = bd[q];
                            const uint32_t tj = bj[q];
                            bd[q] = d2;
                            bj[q] = jj;
                            d2 = t;
                            jj = tj;
                        }
                }
            }
    const float s0 = sqrtf(fmaxf(meanD2[i], 1e-14f));
    const float inv = 1.0f / s0;
    float *f = features + i * kFeatureDim;
#pragma unroll
    for (int q = 0; q < kK; ++q) {
        float ox = 0.0f, oy = 0.0f, oz = 0.0f, d = 8.0f;
        if (bj[q] != 0xffffffffu) {
            const size_t j = bj[q];
            ox = fminf(fmaxf((xyz[j * 3] - xi) * inv, -8.0f), 8.0f);
            oy = fminf(fmaxf((xyz[j * 3 + 1] - yi) * inv, -8.0f), 8.0f);
            oz = fminf(fmaxf((xyz[j * 3 + 2] - zi) * inv, -8.0f), 8.0f);
            d = fminf(sqrtf(bd[q]) * inv, 8.0f);
        }
        f[q * 4] = ox;
        f[q * 4 + 1] = oy;
        f[q * 4 + 2] = oz;
        f[q * 4 + 3] = d;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
        f[kK * 4 + k] = (rgbCount >= (i + 1) * 3 ? rgb[i * 3 + k] : 0.5f) - 0.5f;
    f[kK * 4 + 3] = viewFrac ? viewFrac[i] : 1.0f;
    for (int k = kK * 4 + 4; k < kFeatureDim; ++k)
        f[k] = 0.0f;
}

// Fraction of the cameras whose image contains the point's projection (in front of the near
// plane): the seed point's view count of an SfM / AR track, normalised.
__global__ __launch_bounds__(kThreads) void viewFractionKernel(const float *__restrict__ xyz,
                                                               int64_t n,
                                                               const u3d_camera *__restrict__ cams,
                                                               int numCams, float nearPlane,
                                                               float *out)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n)
        return;
    const float x = xyz[i * 3], y = xyz[i * 3 + 1], z = xyz[i * 3 + 2];
    int seen = 0;
    for (int c = 0; c < numCams; ++c) {
        const u3d_camera &cam = cams[c];
        const float *R = cam.R;
        const float pz = R[6] * x + R[7] * y + R[8] * z + cam.t[2];
        if (pz < nearPlane)
            continue;
        const float px = R[0] * x + R[1] * y + R[2] * z + cam.t[0];
        const float py = R[3] * x + R[4] * y + R[5] * z + cam.t[1];
        const float u = cam.fx * px / pz + cam.cx, v = cam.fy * py / pz + cam.cy;
        if (u >= -0.5f && u < float(cam.width) - 0.5f && v >= -0.5f && v < float(cam.height) - 0.5f)
            ++seen;
    }
    out[i] = numCams > 0 ? float(seen) / float(numCams) : 1.0f;
}

inline int blocks(int64_t n) { return int((n + kThreads - 1) / kThreads); }

} // namespace

int launchPointBounds(hipStream_t s, const float *xyz, int64_t n, float *partial, float *out6)
{
    const int nb = int(std::min<int64_t>(std::max<int64_t>(blocks(n), 1), kBoundsBlocks));
    hipLaunchKernelGGL(boundsPartial, dim3(nb), dim3(kThreads), 0, s, xyz, n, partial);
    U3D_LAUNCH_CHECK();
    hipLaunchKernelGGL(boundsFinal, dim3(1), dim3(64), 0, s, partial, nb, out6);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchCellKeys(hipStream_t s, const float *xyz, int64_t n, const float lo[3], float cell,
                   uint32_t *keyLo, uint32_t *keyHi, uint32_t *idx)
{
    if (n <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(cellKeysKernel, dim3(blocks(n)), dim3(kThreads), 0, s, xyz, n,
                       make_float3(lo[0], lo[1], lo[2]), cell, keyLo, keyHi, idx);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchKnnInit(hipStream_t s, const float *xyz, const float *rgb, int rgbCount, int64_t n,
                  const float lo[3], float cell, const PointGrid &g, float opacityLogit,
                  u3d_gaussians &out, float *meanD2)
{
    if (n <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(knnInitKernel, dim3(blocks(n)), dim3(kThreads), 0, s, xyz, rgb, rgbCount, n,
                       make_float3(lo[0], lo[1], lo[2]), cell, g.keyHi, g.keyLo, g.idx,
                       opacityLogit, out.means, out.quats, out.opacities, out.colors, meanD2);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchKnnFeatures(hipStream_t s, const float *xyz, const float *rgb, int rgbCount, int64_t n,
                      const float lo[3], float cell, const PointGrid &g, const float *meanD2,
                      const float *viewFrac, float *features)
{
    if (n <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(knnFeaturesKernel, dim3(blocks(n)), dim3(kThreads), 0, s, xyz, rgb, rgbCount,
                       n, make_float3(lo[0], lo[1], lo[2]), cell, g.keyHi, g.keyLo, g.idx, meanD2,
                       viewFrac, features);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchViewFraction(hipStream_t s, const float *xyz, int64_t n, const u3d_camera *cams,
                       int numCams, float nearPlane, float *out)
{
    if (n <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(viewFractionKernel, dim3(blocks(n)), dim3(kThreads), 0, s, xyz, n, cams,
                       numCams, nearPlane, out);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

} // namespace u3d

namespace u3d {

namespace {

// Learned initialisation: the network's 11 outputs correct the kNN initialisation of the same
// point (log-scales, quaternion, opacity logit, DC colour).
__global__ __launch_bounds__(256) void applyCorrectionsKernel(const float *__restrict__ net, int ld,
                                                              int64_t n, float *scales, float *quats,
                                                              float *opac, float *colors)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n)
        return;
    const float *o = net + i * ld;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        scales[i * 3 + k] += o[k];
        colors[i * 3 + k] += o[8 + k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        quats[i * 4 + k] += o[3 + k];
    opac[i] += o[7];
}

} // namespace

int launchApplyCorrections(hipStream_t s, const float *net, int ld, int64_t n, u3d_gaussians &m)
{
    if (n <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(applyCorrectionsKernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, net,
                       ld, n, m.scales, m.quats, m.opacities, m.colors);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

} // namespace u3d
