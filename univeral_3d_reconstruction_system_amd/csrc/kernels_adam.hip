// Fused Adam (Adam.cpp:13-28, 68-78), opacity reset (Trainer.cpp:171-178) and densification as
// stream compaction (Trainer.cpp:105-169). All elementwise / scan-shaped and HBM-bound.
// Contraction is off: parameter updates and split offsets are bit-identical to the reference.
#include "kernels.h"
#include "projection.h"

namespace u3d {

namespace {

constexpr int kThreads = 256;

struct AdamGroups {
    float *p[5];
    float *g[5];
    float *m[5];
    float *v[5];
    int64_t len[5];
    float stepSize[5];
};

// One pass over the five parameter groups (means, scales, quats, opacities, colours).
// kClear: each gradient is zeroed once applied (the trainer's buffer is then all zero again).
template <bool kClear>
__global__ __launch_bounds__(kThreads) void adamKernel(AdamGroups a, float invSqrtBc2)
{
    const int64_t stride = int64_t(gridDim.x) * kThreads;
    const int64_t start = int64_t(blockIdx.x) * kThreads + threadIdx.x;
#pragma unroll
    for (int grp = 0; grp < 5; ++grp) {
        float *__restrict__ p = a.p[grp];
        float *__restrict__ g = a.g[grp];
        float *__restrict__ m = a.m[grp];
        float *__restrict__ v = a.v[grp];
        const float stepSize = a.stepSize[grp];
        for (int64_t k = start; k < a.len[grp]; k += stride) {
            float pk = p[k], mk = m[k], vk = v[k];
            adamScalar(pk, mk, vk, g[k], stepSize, invSqrtBc2);
            m[k] = mk;
            v[k] = vk;
            p[k] = pk;
            if (kClear)
                g[k] = 0.0f;
        }
    }
}

__device__ __forceinline__ void adamUpdate(float *__restrict__ p, float *__restrict__ g,
                                           float *__restrict__ m, float *__restrict__ v,
                                           int64_t len, float stepSize, float invSqrtBc2,
                                           float b1, float b2, int64_t start, int64_t stride)
{
    for (int64_t k = start; k < len; k += stride) {
        const float gk = g[k];
        const float mk = b1 * m[k] + (1.0f - b1) * gk;
        const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
        m[k] = mk;
        v[k] = vk;
        p[k] = p[k] - stepSize * mk / (sqrtf(vk) * invSqrtBc2 + 1e-15f);
        g[k] = 0.0f;
    }
}

// Batched Adam: scene blockIdx.y; the same update as adamKernel, and the gradients are cleared
// for the next step's accumulation.
__global__ __launch_bounds__(kThreads) void adamBatchedKernel(const SceneStep *__restrict__ scenes)
{
    const SceneStep &sc = scenes[blockIdx.y];
    const int64_t stride = int64_t(gridDim.x) * kThreads;
    const int64_t start = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    const int64_t c = sc.cap, n = sc.n;
    const int64_t off[5] = {0, 3 * c, 6 * c, 10 * c, 11 * c};
    const int64_t len[5] = {3 * n, 3 * n, 4 * n, n, 3 * n};
#pragma unroll
    for (int grp = 0; grp < 5; ++grp)
        adamUpdate(sc.params + off[grp], sc.grads + off[grp], sc.adamM + off[grp],
                   sc.adamV + off[grp], len[grp], sc.stepSize[grp], sc.invSqrtBc2, sc.beta1,
                   sc.beta2, start, stride);
    if (sc.shDegree > 0)
        adamUpdate(sc.shP, sc.shG, sc.shM, sc.shV, 45 * n, sc.shStepSize, sc.invSqrtBc2, sc.beta1,
                   sc.beta2, start, stride);
}

__global__ void fillKernel(float *p, float value, int64_t count)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += int64_t(gridDim.x) * blockDim.x)
        p[i] = value;
}

__global__ void fillI32Kernel(int *p, int value, int64_t count)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += int64_t(gridDim.x) * blockDim.x)
        p[i] = value;
}

__global__ __launch_bounds__(kThreads) void stepPrologueKernel(
    const uint4 *__restrict__ pinned, size_t slotWords, size_t tableWords, uint4 *__restrict__ table,
    const uint32_t *__restrict__ ring, uint32_t *zeroA, int countA, uint32_t *zeroB, int countB)
{
    const uint4 *src = pinned + size_t(*ring) * slotWords;
    for (size_t k = threadIdx.x; k < tableWords; k += kThreads)
        table[k] = src[k];
    if (int(threadIdx.x) < countA)
        zeroA[threadIdx.x] = 0u;
    if (int(threadIdx.x) < countB)
        zeroB[threadIdx.x] = 0u;
}

__global__ void opacityCapKernel(float *o, int64_t n, float cap)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        const float x = o[i];
        o[i] = (cap < x) ? cap : x; // std::min(x, cap)
    }
}

// Trainer.cpp:116-135 up to the budget test: keep flag and demand (0, 1 clone, 2 split).
__global__ __launch_bounds__(kThreads) void densifyDemandKernel(
    int64_t n, const float *__restrict__ scales, const float *__restrict__ opac,
    const float *__restrict__ accum, const int *__restrict__ count, float splitScale,
    float largeScale, float gradThreshold, float pruneOpacity, int wasReset, uint8_t *keep,
    uint32_t *demand)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n)
        return;
    // std::max({a, b, c}) keeps the first of equal maxima and follows operator<.
    float maxLog = scales[i * 3];
    if (maxLog < scales[i * 3 + 1])
        maxLog = scales[i * 3 + 1];
    if (maxLog < scales[i * 3 + 2])
        maxLog = scales[i * 3 + 2];
    const float maxScale = expfExact(maxLog);
    const float alpha = 1.0f / (1.0f + expfExact(-opac[i]));
    uint8_t k = 1;
    uint32_t d = 0;
    if (alpha < pruneOpacity || (wasReset && maxScale > largeScale)) {
        k = 0;
    } else if (count[i] != 0) {
        const float avg = accum[i] / float(count[i]);
        if (!(avg < gradThreshold))
            d = maxScale <= splitScale ? 1u : 2u;
    }
    keep[i] = k;
    demand[i] = d;
}

// Candidates whose running demand fits the budget are accepted (Trainer.cpp:130-137).
__global__ __launch_bounds__(kThreads) void densifyAcceptKernel(
    int64_t n, const uint32_t *__restrict__ demand, const uint32_t *__restrict__ scan,
    uint64_t budget, uint8_t *keep, uint8_t *kind, uint32_t *firstReject)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t d = demand[i];
    uint8_t kd = 0;
    if (d != 0) {
        if (uint64_t(scan[i]) + d <= budget)
            kd = uint8_t(d);
        else
            atomicMin(firstReject, uint32_t(i));
    }
    kind[i] = kd;
    if (kd == 2)
        keep[i] = 0;
}

// After the first rejection at running total budget - 1 (necessarily a split), exactly one
// more candidate fits: the first clone after it.
__global__ __launch_bounds__(kThreads) void densifyLateCloneKernel(
    int64_t n, const uint32_t *__restrict__ demand, const uint32_t *__restrict__ scan,
    uint64_t budget, const uint32_t *__restrict__ firstReject, uint32_t *firstClone)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    const uint32_t r = *firstReject;
    if (i >= n || r == 0xffffffffu || uint64_t(scan[r]) + 1 != budget)
        return;
    if (uint32_t(i) > r && demand[i] == 1)
        atomicMin(firstClone, uint32_t(i));
}

__global__ void densifyCountKernel(int64_t n, const uint8_t *__restrict__ keep, uint8_t *kind,
                                   const uint32_t *__restrict__ firstClone, uint32_t *counts)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    if (uint32_t(i) == *firstClone)
        kind[i] = 1;
    const uint8_t kd = kind[i];
    const uint64_t clones = __ballot(kd == 1), splits = __ballot(kd == 2), kept = __ballot(keep[i] != 0);
    if ((threadIdx.x & 63) == 0) {
        if (clones) atomicAdd(&counts[0], uint32_t(__popcll(clones)));
        if (splits) atomicAdd(&counts[1], uint32_t(__popcll(splits)));
        if (kept) atomicAdd(&counts[2], uint32_t(__popcll(kept)));
    }
}

__global__ void densifyFlagsKernel(int64_t n, const uint8_t *__restrict__ keep,
                                   const uint8_t *__restrict__ kind, uint32_t *keepFlag,
                                   uint32_t *extraCount, uint32_t *splitFlag)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    keepFlag[i] = keep[i] ? 1u : 0u;
    extraCount[i] = kind[i] == 1 ? 1u : (kind[i] == 2 ? 2u : 0u);
    splitFlag[i] = kind[i] == 2 ? 1u : 0u;
}

__device__ __forceinline__ void copyEntry(const u3d_gaussians &s, int64_t i, u3d_gaussians &d,
                                          int64_t j)
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        d.means[j * 3 + k] = s.means[i * 3 + k];
        d.scales[j * 3 + k] = s.scales[i * 3 + k];
        d.colors[j * 3 + k] = s.colors[i * 3 + k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        d.quats[j * 4 + k] = s.quats[i * 4 + k];
    d.opacities[j] = s.opacities[i];
}

// Kept entries in order, then clones and split children in index order
// (Trainer.cpp:133-168). Split child c of the s-th accepted split uses normals[6s + 3c + k].
__global__ __launch_bounds__(kThreads) void densifyScatterKernel(
    u3d_gaussians src, u3d_gaussians dst, int64_t n, const uint8_t *__restrict__ keep,
    const uint8_t *__restrict__ kind, const uint32_t *__restrict__ keepPos,
    const uint32_t *__restrict__ extraPos, const uint32_t *__restrict__ splitPos,
    const uint32_t *__restrict__ counts, const float *__restrict__ normals, float logShrink)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n)
        return;
    const int64_t nKeep = counts[2];
    if (keep[i])
        copyEntry(src, i, dst, keepPos[i]);
    const uint8_t kd = kind[i];
    if (kd == 1) {
        copyEntry(src, i, dst, nKeep + extraPos[i]);
    } else if (kd == 2) {
        const float q0 = src.quats[i * 4], q1 = src.quats[i * 4 + 1], q2 = src.quats[i * 4 + 2],
                    q3 = src.quats[i * 4 + 3];
        const float qn = sqrtf(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3) + 1e-12f;
        float R[9];
        rotOfQuat(q0 / qn, q1 / qn, q2 / qn, q3 / qn, R);
        const int64_t s = splitPos[i];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int64_t j = nKeep + extraPos[i] + c;
            copyEntry(src, i, dst, j);
            float local[3];
#pragma unroll
            for (int k = 0; k < 3; ++k)
                local[k] = normals[s * 6 + c * 3 + k] * expfExact(src.scales[i * 3 + k]);
#pragma unroll
            for (int r = 0; r < 3; ++r)
                dst.means[j * 3 + r] = dst.means[j * 3 + r]
                                       + (R[r * 3] * local[0] + R[r * 3 + 1] * local[1] + R[r * 3 + 2] * local[2]);
#pragma unroll
            for (int k = 0; k < 3; ++k)
                dst.scales[j * 3 + k] = dst.scales[j * 3 + k] - logShrink;
        }
    }
}

__global__ __launch_bounds__(kThreads) void compactZeroTailKernel(
    u3d_gaussians src, u3d_gaussians dst, int64_t n, const uint8_t *__restrict__ keep,
    const uint32_t *__restrict__ keepPos, const uint32_t *__restrict__ counts)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    const int64_t nKeep = counts[2];
    const int64_t total = nKeep + int64_t(counts[0]) + 2 * int64_t(counts[1]);
    if (i < n && keep[i])
        copyEntry(src, i, dst, keepPos[i]);
    // zero the appended tail [nKeep, total): grid covers max(n, total)
    const int64_t j = nKeep + i;
    if (j < total) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            dst.means[j * 3 + k] = 0.0f;
            dst.scales[j * 3 + k] = 0.0f;
            dst.colors[j * 3 + k] = 0.0f;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
            dst.quats[j * 4 + k] = 0.0f;
        dst.opacities[j] = 0.0f;
    }
}

// Adam over one extra array (the SH coefficients), same arithmetic as adamKernel.
__global__ __launch_bounds__(kThreads) void adamArrayKernel(float *p, const float *__restrict__ g,
                                                            float *m, float *v, int64_t len,
                                                            float stepSize, float invSqrtBc2)
{
    for (int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x; k < len;
         k += int64_t(gridDim.x) * kThreads) {
        const float gk = g[k];
        const float mk = 0.9f * m[k] + (1.0f - 0.9f) * gk;
        const float vk = 0.999f * v[k] + (1.0f - 0.999f) * gk * gk;
        m[k] = mk;
        v[k] = vk;
        p[k] = p[k] - stepSize * mk / (sqrtf(vk) * invSqrtBc2 + 1e-15f);
    }
}

// Densification of a per-Gaussian row array of `width` floats (e.g. SH coefficients): kept
// rows to keepPos, clones and both split children copy the parent row (zeroRows: write zeros
// for the appended rows instead, as for optimiser moments).
__global__ __launch_bounds__(kThreads) void scatterRowsKernel(
    const float *__restrict__ src, float *__restrict__ dst, int width, int64_t n,
    const uint8_t *__restrict__ keep, const uint8_t *__restrict__ kind,
    const uint32_t *__restrict__ keepPos, const uint32_t *__restrict__ extraPos,
    const uint32_t *__restrict__ counts, int zeroRows)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n)
        return;
    const int64_t nKeep = counts[2];
    const float *row = src + i * width;
    if (keep[i]) {
        float *o = dst + int64_t(keepPos[i]) * width;
        for (int k = 0; k < width; ++k)
            o[k] = row[k];
    }
    const int copies = kind[i] == 1 ? 1 : (kind[i] == 2 ? 2 : 0);
    for (int c = 0; c < copies; ++c) {
        float *o = dst + (nKeep + int64_t(extraPos[i]) + c) * width;
        for (int k = 0; k < width; ++k)
            o[k] = zeroRows ? 0.0f : row[k];
    }
}

// Densification of planes of n floats (plane stride n) into planes of stride newN: the kept
// entry i goes to keepPos[i], its extras (1 clone, 2 split children) after the kept ones.
__global__ __launch_bounds__(kThreads) void scatterPlanesKernel(
    const float *__restrict__ src, float *__restrict__ dst, int planes, int64_t n, int64_t newN,
    const uint8_t *__restrict__ keep, const uint8_t *__restrict__ kind,
    const uint32_t *__restrict__ keepPos, const uint32_t *__restrict__ extraPos,
    const uint32_t *__restrict__ counts, int zeroRows)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n)
        return;
    const int64_t nKeep = counts[2];
    const bool kept = keep[i] != 0;
    const int copies = kind[i] == 1 ? 1 : (kind[i] == 2 ? 2 : 0);
    const int64_t kp = kept ? int64_t(keepPos[i]) : 0, ep = nKeep + int64_t(extraPos[i]);
    for (int c = 0; c < planes; ++c) {
        const float v = src[int64_t(c) * n + i];
        float *plane = dst + int64_t(c) * newN;
        if (kept)
            plane[kp] = v;
        for (int k = 0; k < copies; ++k)
            plane[ep + k] = zeroRows ? 0.0f : v;
    }
}

inline unsigned blocksFor(int64_t n) { return unsigned((n + kThreads - 1) / kThreads); }

} // namespace

int launchAdamArray(hipStream_t s, float *p, const float *g, float *m, float *v, int64_t len,
                    float stepSize, float invSqrtBc2)
{
    if (len <= 0)
        return U3D_OK;
    unsigned blocks = blocksFor(len);
    if (blocks > 4096)
        blocks = 4096;
    hipLaunchKernelGGL(adamArrayKernel, dim3(blocks), dim3(kThreads), 0, s, p, g, m, v, len,
                       stepSize, invSqrtBc2);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchScatterRows(hipStream_t s, const float *src, float *dst, int width, int64_t n,
                      const uint8_t *keep, const uint8_t *kind, const uint32_t *keepPos,
                      const uint32_t *extraPos, const uint32_t *counts, bool zeroRows)
{
    if (n <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(scatterRowsKernel, dim3(blocksFor(n)), dim3(kThreads), 0, s, src, dst, width,
                       n, keep, kind, keepPos, extraPos, counts, zeroRows ? 1 : 0);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchScatterPlanes(hipStream_t s, const float *src, float *dst, int planes, int64_t n,
                        int64_t newN, const uint8_t *keep, const uint8_t *kind,
                        const uint32_t *keepPos, const uint32_t *extraPos,
                        const uint32_t *counts, bool zeroRows)
{
    if (n <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(scatterPlanesKernel, dim3(blocksFor(n)), dim3(kThreads), 0, s, src, dst, planes,
                       n, newN, keep, kind, keepPos, extraPos, counts, zeroRows ? 1 : 0);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchAdamStep(hipStream_t s, u3d_gaussians &p, const u3d_gaussians &g, u3d_gaussians &m,
                   u3d_gaussians &v, const float stepSize[5], float invSqrtBc2, bool clearGrads)
{
    if (p.n <= 0)
        return U3D_OK;
    AdamGroups a;
    float *pp[5] = {p.means, p.scales, p.quats, p.opacities, p.colors};
    float *gg[5] = {g.means, g.scales, g.quats, g.opacities, g.colors};
    float *mm[5] = {m.means, m.scales, m.quats, m.opacities, m.colors};
    float *vv[5] = {v.means, v.scales, v.quats, v.opacities, v.colors};
    const int dims[5] = {3, 3, 4, 1, 3};
    for (int k = 0; k < 5; ++k) {
        a.p[k] = pp[k];
        a.g[k] = gg[k];
        a.m[k] = mm[k];
        a.v[k] = vv[k];
        a.len[k] = p.n * dims[k];
        a.stepSize[k] = stepSize[k];
    }
    const int64_t maxLen = p.n * 4;
    unsigned blocks = blocksFor(maxLen);
    if (blocks > 4096)
        blocks = 4096;
    if (clearGrads)
        hipLaunchKernelGGL(adamKernel<true>, dim3(blocks), dim3(kThreads), 0, s, a, invSqrtBc2);
    else
        hipLaunchKernelGGL(adamKernel<false>, dim3(blocks), dim3(kThreads), 0, s, a, invSqrtBc2);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchAdamBatched(hipStream_t s, const SceneStep *scenes, int numScenes, int64_t maxN)
{
    if (numScenes <= 0 || maxN <= 0)
        return U3D_OK;
    int64_t blocks = (4 * maxN + kThreads - 1) / kThreads;
    blocks = blocks < 1 ? 1 : (blocks > 4096 ? 4096 : blocks);
    hipLaunchKernelGGL(adamBatchedKernel, dim3(unsigned(blocks), numScenes), dim3(kThreads), 0, s,
                       scenes);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchStepPrologue(hipStream_t s, const unsigned char *pinned, size_t slotBytes,
                       size_t tableBytes, unsigned char *table, const uint32_t *ring,
                       uint32_t *zeroA, int countA, uint32_t *zeroB, int countB)
{
    U3D_CHECK(slotBytes % 16 == 0 && tableBytes % 16 == 0 && tableBytes <= slotBytes &&
                  countA <= kThreads && countB <= kThreads,
              U3D_ERR_INVALID, "step table layout");
    hipLaunchKernelGGL(stepPrologueKernel, dim3(1), dim3(kThreads), 0, s,
                       reinterpret_cast<const uint4 *>(pinned), slotBytes / 16, tableBytes / 16,
                       reinterpret_cast<uint4 *>(table), ring, zeroA, countA, zeroB, countB);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchFill(hipStream_t s, float *p, float value, int64_t count)
{
    if (count <= 0)
        return U3D_OK;
    unsigned b = blocksFor(count);
    hipLaunchKernelGGL(fillKernel, dim3(b > 4096 ? 4096 : b), dim3(kThreads), 0, s, p, value, count);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchFillI32(hipStream_t s, int *p, int value, int64_t count)
{
    if (count <= 0)
        return U3D_OK;
    unsigned b = blocksFor(count);
    hipLaunchKernelGGL(fillI32Kernel, dim3(b > 4096 ? 4096 : b), dim3(kThreads), 0, s, p, value, count);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchOpacityCap(hipStream_t s, float *opacities, int64_t n, float cap)
{
    if (n <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(opacityCapKernel, dim3(blocksFor(n)), dim3(kThreads), 0, s, opacities, n, cap);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchDensifyDemand(hipStream_t s, const u3d_gaussians &m, const float *accum,
                        const int *count, float extent, float gradThreshold, float percentDense,
                        float pruneOpacity, int opacityWasReset, uint8_t *keep, uint32_t *demand)
{
    if (m.n <= 0)
        return U3D_OK;
    const float splitScale = percentDense * extent; // Trainer.cpp:108
    const float largeScale = 0.1f * extent;         // Trainer.cpp:120
    hipLaunchKernelGGL(densifyDemandKernel, dim3(blocksFor(m.n)), dim3(kThreads), 0, s, m.n,
                       m.scales, m.opacities, accum, count, splitScale, largeScale, gradThreshold,
                       pruneOpacity, opacityWasReset, keep, demand);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchDensifyAccept(hipStream_t s, int64_t n, const uint32_t *demand, const uint32_t *scan,
                        uint64_t budget, uint8_t *keep, uint8_t *kind, uint32_t *firstReject,
                        uint32_t *counts)
{
    // firstReject[0], firstClone = firstReject[1]; counts[0..3]
    U3D_HIP(hipMemsetAsync(firstReject, 0xff, 2 * sizeof(uint32_t), s));
    U3D_HIP(hipMemsetAsync(counts, 0, 4 * sizeof(uint32_t), s));
    if (n <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(densifyAcceptKernel, dim3(blocksFor(n)), dim3(kThreads), 0, s, n, demand,
                       scan, budget, keep, kind, firstReject);
    U3D_LAUNCH_CHECK();
    hipLaunchKernelGGL(densifyLateCloneKernel, dim3(blocksFor(n)), dim3(kThreads), 0, s, n,
                       demand, scan, budget, firstReject, firstReject + 1);
    U3D_LAUNCH_CHECK();
    hipLaunchKernelGGL(densifyCountKernel, dim3(blocksFor(n)), dim3(kThreads), 0, s, n, keep,
                       kind, firstReject + 1, counts);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchDensifyFlags(hipStream_t s, int64_t n, const uint8_t *keep, const uint8_t *kind,
                       uint32_t *keepFlag, uint32_t *extraCount, uint32_t *splitFlag)
{
    if (n <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(densifyFlagsKernel, dim3(blocksFor(n)), dim3(kThreads), 0, s, n, keep, kind,
                       keepFlag, extraCount, splitFlag);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchDensifyScatter(hipStream_t s, const u3d_gaussians &src, u3d_gaussians &dst,
                         int64_t n, const uint8_t *keep, const uint8_t *kind,
                         const uint32_t *keepPos, const uint32_t *extraPos,
                         const uint32_t *splitPos, const uint32_t *counts, const float *normals,
                         float logShrink)
{
    if (n <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(densifyScatterKernel, dim3(blocksFor(n)), dim3(kThreads), 0, s, src, dst, n,
                       keep, kind, keepPos, extraPos, splitPos, counts, normals, logShrink);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchCompactZeroTail(hipStream_t s, const u3d_gaussians &src, u3d_gaussians &dst,
                          int64_t n, const uint8_t *keep, const uint32_t *keepPos,
                          const uint32_t *counts)
{
    // the tail is at most 2n long
    const int64_t grid = 2 * n;
    if (grid <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(compactZeroTailKernel, dim3(blocksFor(grid)), dim3(kThreads), 0, s, src,
                       dst, n, keep, keepPos, counts);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

} // namespace u3d
