// The learned initialiser's network kernels: a BF16 MFMA GEMM with FP32 accumulation and fused
// bias / GELU / GELU-gradient epilogues, plus the small elementwise kernels of its training
// (MSE loss, column sums, Adam with a BF16 weight copy).
//
// GEMM: a 256-thread workgroup owns a 64 x 64 tile of C; each of its four waves computes a
// 32 x 32 sub-tile as 2 x 2 blocks of v_mfma_f32_16x16x16_bf16. Operand tiles of 64 x 32 are
// staged in LDS with k contiguous, so each lane fetches its four k-consecutive bf16 values of
// an A row or a B column with one 8-byte LDS read. Per lane l of a 16 x 16 x 16 MFMA: A(i, k) with
// i = l % 16, k = 4 (l / 16) + v; B(k, j) with j = l % 16, k = 4 (l / 16) + v; D(i, j) with
// j = l % 16, i = 4 (l / 16) + v, v = 0..3.
#include "kernels.h"

namespace u3d {

namespace {

constexpr int kBM = 64, kBN = 64, kBK = 32, kThreads = 256;
constexpr int kLdsK = kBK + 8; // padded row (bank spread)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t toBf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) // NaN stays NaN
        return uint16_t((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

__device__ __forceinline__ float geluF(float x)
{
    return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}

__device__ __forceinline__ float geluGrad(float x)
{
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * __expf(-0.5f * x * x);
    return cdf + x * pdf;
}

__global__ __launch_bounds__(kThreads) void gemmBf16Kernel(GemmArgs a, int kChunk)
{
    __shared__ __attribute__((aligned(16))) uint16_t As[kBM][kLdsK];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[kBN][kLdsK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
    const int kBegin = blockIdx.z * kChunk;
    const int kEnd = min(a.K, kBegin + kChunk);
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
            acc[i][j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    const bool aRowFast = a.sam == 1 && a.sak != 1; // m contiguous in memory
    const bool bColFast = a.sbn == 1 && a.sbk != 1; // n contiguous in memory
    for (int k0 = kBegin; k0 < kEnd; k0 += kBK) {
#pragma unroll
        for (int q = 0; q < kBM * kBK / kThreads; ++q) {
            const int e = tid + q * kThreads;
            const int r = aRowFast ? e % kBM : e / kBK, k = aRowFast ? e / kBM : e % kBK;
            const int gm = m0 + r, gk = k0 + k;
            As[r][k] = (gm < a.M && gk < kEnd) ? a.A[int64_t(gm) * a.sam + int64_t(gk) * a.sak] : uint16_t(0);
        }
#pragma unroll
        for (int q = 0; q < kBN * kBK / kThreads; ++q) {
            const int e = tid + q * kThreads;
            const int c = bColFast ? e % kBN : e / kBK, k = bColFast ? e / kBN : e % kBK;
            const int gn = n0 + c, gk = k0 + k;
            Bs[c][k] = (gn < a.N && gk < kEnd) ? a.B[int64_t(gk) * a.sbk + int64_t(gn) * a.sbn] : uint16_t(0);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kBK; kk += 16) {
            const int ko = kk + 4 * (lane >> 4);
            bf16x4 fa[2], fb[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                fa[i] = *reinterpret_cast<const bf16x4 *>(&As[wr + i * 16 + (lane & 15)][ko]);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                fb[j] = *reinterpret_cast<const bf16x4 *>(&Bs[wc + j * 16 + (lane & 15)][ko]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = m0 + wr + i * 16 + 4 * (lane >> 4) + v;
                const int n = n0 + wc + j * 16 + (lane & 15);
                if (m >= a.M || n >= a.N)
                    continue;
                float x = acc[i][j][v];
                if (a.splitK > 1) {
                    unsafeAtomicAdd(&a.C[int64_t(m) * a.ldc + n], x);
                    continue;
                }
                if (a.bias)
                    x += a.bias[n];
                if (a.gate)
                    x *= geluGrad(a.gate[int64_t(m) * a.ldg + n]);
                if (a.C)
                    a.C[int64_t(m) * a.ldc + n] = x;
                if (a.Y)
                    a.Y[int64_t(m) * a.ldy + n] = toBf16(a.gelu ? geluF(x) : x);
            }
}

__global__ void toBf16Kernel(const float *__restrict__ in, uint16_t *out, int64_t count)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += int64_t(gridDim.x) * blockDim.x)
        out[i] = toBf16(in[i]);
}

__global__ __launch_bounds__(kThreads) void columnSumsKernel(const float *__restrict__ in, int64_t ld,
                                                             int M, int N, float *out)
{
    // blockIdx.x: column block of 64, blockIdx.y: row chunk; 4 waves split the rows.
    const int n = blockIdx.x * 64 + (threadIdx.x & 63);
    const int rows = (M + gridDim.y - 1) / gridDim.y;
    const int r0 = blockIdx.y * rows, r1 = min(M, r0 + rows);
    float s = 0.0f;
    if (n < N)
        for (int m = r0 + (threadIdx.x >> 6); m < r1; m += 4)
            s += in[int64_t(m) * ld + n];
    __shared__ float part[4][64];
    part[threadIdx.x >> 6][threadIdx.x & 63] = s;
    __syncthreads();
    if (threadIdx.x < 64 && n < N)
        unsafeAtomicAdd(&out[n], part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x]
                                     + part[3][threadIdx.x]);
}

__global__ __launch_bounds__(kThreads) void mseKernel(const float *__restrict__ y,
                                                      const float *__restrict__ t,
                                                      const float *__restrict__ w, int M,
                                                      int width, int64_t ld, float scale,
                                                      float *dy, float *loss)
{
    const int64_t total = int64_t(M) * ld;
    float s = 0.0f;
    for (int64_t e = int64_t(blockIdx.x) * kThreads + threadIdx.x; e < total;
         e += int64_t(gridDim.x) * kThreads) {
        const int col = int(e % ld);
        float g = 0.0f;
        if (col < width) {
            const float d = y[e] - t[e];
            s += w[col] * d * d;
            g = 2.0f * w[col] * d * scale;
        }
        dy[e] = g;
    }
    for (int off = 32; off > 0; off >>= 1)
        s += __shfl_xor(s, off, 64);
    __shared__ float ws[kThreads / 64];
    if ((threadIdx.x & 63) == 0)
        ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        unsafeAtomicAdd(loss, (ws[0] + ws[1] + ws[2] + ws[3]) * scale);
}

__global__ void adamBf16Kernel(float *p, const float *__restrict__ g, float *m, float *v, int64_t n,
                               float stepSize, float invSqrtBc2, uint16_t *pb)
{
    for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
         k += int64_t(gridDim.x) * blockDim.x) {
        const float gk = g[k];
        const float mk = 0.9f * m[k] + 0.1f * gk;
        const float vk = 0.999f * v[k] + 0.001f * gk * gk;
        m[k] = mk;
        v[k] = vk;
        const float pk = p[k] - stepSize * mk / (sqrtf(vk) * invSqrtBc2 + 1e-8f);
        p[k] = pk;
        if (pb)
            pb[k] = toBf16(pk);
    }
}

inline unsigned gridFor(int64_t n, int per = 256)
{
    const int64_t b = (n + per - 1) / per;
    return unsigned(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

} // namespace

int launchGemmBf16(hipStream_t s, const GemmArgs &a)
{
    if (a.M <= 0 || a.N <= 0)
        return U3D_OK;
    U3D_CHECK(a.splitK >= 1 && (a.splitK == 1 || a.C), U3D_ERR_INVALID, "bad GEMM split");
    U3D_CHECK((a.M + kBM - 1) / kBM <= 65535, U3D_ERR_INVALID, "GEMM M too large");
    const int kChunk = ((a.K + a.splitK - 1) / a.splitK + kBK - 1) / kBK * kBK;
    const int splits = a.K > 0 ? (a.K + kChunk - 1) / kChunk : 1;
    GemmArgs b = a;
    b.splitK = a.splitK > 1 ? splits : 1;
    if (a.splitK > 1 && splits == 1)
        b.splitK = 2; // still accumulate atomically into C, as the caller zeroed it for that
    const dim3 grid((a.N + kBN - 1) / kBN, (a.M + kBM - 1) / kBM, splits);
    hipLaunchKernelGGL(gemmBf16Kernel, grid, dim3(kThreads), 0, s, b, kChunk);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchToBf16(hipStream_t s, const float *in, uint16_t *out, int64_t count)
{
    if (count <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(toBf16Kernel, dim3(gridFor(count)), dim3(256), 0, s, in, out, count);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchColumnSums(hipStream_t s, const float *in, int64_t ld, int M, int N, float *out)
{
    if (M <= 0 || N <= 0)
        return U3D_OK;
    const int chunks = std::max(1, std::min(256, M / 256));
    hipLaunchKernelGGL(columnSumsKernel, dim3((N + 63) / 64, chunks), dim3(kThreads), 0, s, in, ld,
                       M, N, out);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchMseLoss(hipStream_t s, const float *y, const float *target, const float *w, int M,
                  int width, int64_t ld, float *dy, float *loss)
{
    if (M <= 0)
        return U3D_OK;
    const float scale = 1.0f / float(M);
    hipLaunchKernelGGL(mseKernel, dim3(gridFor(int64_t(M) * ld)), dim3(kThreads), 0, s, y, target,
                       w, M, width, ld, scale, dy, loss);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchAdamBf16(hipStream_t s, float *p, const float *g, float *m, float *v, int64_t n,
                   float stepSize, float invSqrtBc2, uint16_t *pBf16)
{
    if (n <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(adamBf16Kernel, dim3(gridFor(n)), dim3(256), 0, s, p, g, m, v, n, stepSize,
                       invSqrtBc2, pBf16);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

} // namespace u3d
