// Fused photometric loss (L1 + lambda * D-SSIM) value and gradient, Loss.cpp:47-185.
//
// The reference runs four full-image passes over 11 planar scratch maps. Here one workgroup
// owns a 32x32 output block and keeps the whole dependency cone in LDS: render/target with a
// 10-pixel halo, the five horizontally blurred moment maps, the three SSIM partial-derivative
// maps with a 5-pixel halo, and their horizontal blur; the vertical blur then yields the
// gradient. Only render, target and grad touch HBM (36 B per pixel). Each blur sums its taps in
// the reference's order over the reference's truncated (zero-padded) window, with contraction
// off, so the gradient is bit-identical; the L1 / SSIM / squared-error sums are per-block
// doubles reduced in a fixed order.
#include "kernels.h"

namespace u3d {

namespace {

constexpr int kR = 5;          // 11-tap window radius
constexpr int kTaps = 2 * kR + 1;
constexpr int kTW = 32, kTH = 32;
constexpr int kIW = kTW + 4 * kR, kIH = kTH + 4 * kR; // inputs: 10-pixel halo
constexpr int kHW = kTW + 2 * kR, kHH = kIH;          // horizontal moment blur
constexpr int kAW = kTW + 2 * kR, kAH = kTH + 2 * kR; // SSIM partials
constexpr int kGW = kTW, kGH = kAH;                   // horizontal blur of partials
// LDS row strides, padded so that lanes walking different rows hit different banks
constexpr int kInStride = kIW + 1, kHStride = kHW + 1, kAStride = kAW + 1, kGStride = kGW + 1;
// Outputs per thread along a row (horizontal blurs) or a column (vertical blurs), computed from
// a sliding window of kSeg + 10 inputs held in registers.
constexpr int kH1Seg = 6, kV1Seg = 4, kH2Seg = 3, kV2Seg = 2;
constexpr int kThreads = 512;
constexpr float kC1 = 0.01f * 0.01f;
constexpr float kC2 = 0.03f * 0.03f;
static_assert(kIH * ((kHW + kH1Seg - 1) / kH1Seg) <= kThreads, "one H1 task per thread");
static_assert(kAW * ((kAH + kV1Seg - 1) / kV1Seg) <= kThreads, "one V1 task per thread");
static_assert(kGH * ((kGW + kH2Seg - 1) / kH2Seg) <= kThreads, "one H2 task per thread");
static_assert(kTW * ((kTH + kV2Seg - 1) / kV2Seg) <= kThreads, "one output task per thread");
static_assert(kTW * (kTH / kV2Seg) == kThreads && kTH % kV2Seg == 0,
              "every thread owns kV2Seg pixels of one column");

struct Window {
    float w[kTaps];
};

__device__ __forceinline__ double blockSumD(double v, double *red)
{
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    const int tid = threadIdx.x;
    if ((tid & 63) == 0)
        red[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (tid == 0)
        for (int w = 0; w < kThreads / 64; ++w)
            s += red[w];
    __syncthreads();
    return s;
}

// First output of segment s of kSeg outputs along a line of len (the last segment is moved
// back to end at len, so it overlaps the one before and recomputes the same values).
__device__ __forceinline__ int segStart(int s, int seg, int len)
{
    return s * seg < len - seg ? s * seg : len - seg;
}

// One 32x32 block of one channel; writes its three double partials at partials[0..2].
// gradScale multiplies the stored gradient (kScaled only: views averaged in a batched step).
//
// Every blur sums all 11 taps in the reference's order over maps that are exactly zero outside
// the image, which equals the reference's truncated window bit for bit: an added term w * 0 is
// +0, and a running sum that starts at +0 never becomes -0 in round-to-nearest, so x + (+0) = x.
// The inputs are zero outside the image, so H1 of an outside row is +0; the SSIM partials are
// stored as 0 outside the image, so H2 of an outside row is +0 as well. Each thread computes a
// run of consecutive outputs from a window of inputs in registers; every output still sums its
// own 11 terms in tap order from +0.
template <bool kU8, bool kScaled>
__device__ __forceinline__ void lossBlock(const float *__restrict__ render,
                                          const float *__restrict__ targetF,
                                          const uint8_t *__restrict__ targetU8, int W, int H,
                                          float lambda, float invCount, float gradScale,
                                          const Window &win, int blocksX, int blockId, int c,
                                          float *__restrict__ grad, double *__restrict__ partials)
{
    // LDS (67 KB): the inputs (render, target) are dead once H1 is done, so the SSIM partials
    // reuse their space; the H1 maps are dead once V1 is done, so H2 reuses theirs.
    constexpr int kInBytes = kIH * kInStride * 8, kABytes = 3 * kAH * kAStride * 4;
    constexpr int kHBytes = 5 * kHH * kHStride * 4, kGBytes = 3 * kGH * kGStride * 4;
    __shared__ __align__(16) unsigned char ldsA[kInBytes > kABytes ? kInBytes : kABytes];
    __shared__ __align__(16) unsigned char ldsB[kHBytes > kGBytes ? kHBytes : kGBytes];
    __shared__ double red[3][kThreads / 64];
    __shared__ float sV[kTH][kTW]; // SSIM of the block's pixels (summed in the split order)
    auto sIn = reinterpret_cast<float2(*)[kInStride]>(ldsA);
    auto sA = reinterpret_cast<float(*)[kAH][kAStride]>(ldsA);
    auto sH = reinterpret_cast<float(*)[kHH][kHStride]>(ldsB);
    auto sG = reinterpret_cast<float(*)[kGH][kGStride]>(ldsB);
    const int tid = threadIdx.x;
    float w[kTaps];
#pragma unroll
    for (int t = 0; t < kTaps; ++t)
        w[t] = win.w[t];
    const int ox = (blockId % blocksX) * kTW, oy = (blockId / blocksX) * kTH;
    const bool ssimOn = lambda > 0.0f;
    const float dLds = -lambda * invCount;
    const float l1Scale = (1.0f - lambda) * invCount;
    double l1 = 0.0, ssim = 0.0, sq = 0.0;
    auto tgt = [&](size_t idx) -> float {
        if (kU8)
            return float(targetU8[idx]) * (1.0f / 255.0f);
        return targetF[idx];
    };
    auto inImage = [&](int gx, int gy) { return gx >= 0 && gx < W && gy >= 0 && gy < H; };
    // L1 term and the pixel's gradient (Loss.cpp:64-74, 160-177); g = the blurred partials
    auto finish = [&](int gx, int gy, const float g[3]) {
        const size_t idx = (size_t(gy) * size_t(W) + size_t(gx)) * 3 + size_t(c);
        const float r = render[idx], t = tgt(idx);
        const float d = r - t;
        l1 += double(fabsf(d));
        const double dd = double(r) - double(t);
        sq += dd * dd;
        float gval = l1Scale * (d > 0 ? 1.0f : (d < 0 ? -1.0f : 0.0f));
        if (ssimOn)
            gval += g[0] + 2.0f * r * g[1] + t * g[2];
        if (kScaled)
            gval *= gradScale;
        grad[idx] = gval;
    };
    if (ssimOn) {
        // 1. inputs with a 10-pixel halo (all loads issued before the LDS stores)
        constexpr int kIn = kIH * kIW, kInIter = (kIn + kThreads - 1) / kThreads;
        float2 v[kInIter];
#pragma unroll
        for (int j = 0; j < kInIter; ++j) {
            const int k = tid + j * kThreads;
            const int ly = k / kIW, lx = k - ly * kIW;
            const int gx = ox - 2 * kR + lx, gy = oy - 2 * kR + ly;
            v[j] = make_float2(0.0f, 0.0f);
            if (k < kIn && inImage(gx, gy)) {
                const size_t idx = (size_t(gy) * size_t(W) + size_t(gx)) * 3 + size_t(c);
                v[j] = make_float2(render[idx], tgt(idx));
            }
        }
#pragma unroll
        for (int j = 0; j < kInIter; ++j) {
            const int k = tid + j * kThreads;
            if (k < kIn)
                sIn[k / kIW][k % kIW] = v[j];
        }
        __syncthreads();
        // 2. horizontal blur of x, y, x^2, y^2, xy (Loss.cpp:77-94): row ly, kH1Seg outputs
        if (tid < kIH * ((kHW + kH1Seg - 1) / kH1Seg)) {
            const int ly = tid % kIH, x0 = segStart(tid / kIH, kH1Seg, kHW);
            float2 in[kH1Seg + 2 * kR];
#pragma unroll
            for (int i = 0; i < kH1Seg + 2 * kR; ++i)
                in[i] = sIn[ly][x0 + i];
#pragma unroll
            for (int o = 0; o < kH1Seg; ++o) {
                float sx = 0, sy = 0, sxx = 0, syy = 0, sxy = 0;
#pragma unroll
                for (int t = 0; t < kTaps; ++t) {
                    const float a = in[o + t].x, b = in[o + t].y;
                    sx += w[t] * a;
                    sy += w[t] * b;
                    sxx += w[t] * a * a;
                    syy += w[t] * b * b;
                    sxy += w[t] * a * b;
                }
                sH[0][ly][x0 + o] = sx;
                sH[1][ly][x0 + o] = sy;
                sH[2][ly][x0 + o] = sxx;
                sH[3][ly][x0 + o] = syy;
                sH[4][ly][x0 + o] = sxy;
            }
        }
        __syncthreads();
        // 3. vertical blur, SSIM and its partial derivatives (Loss.cpp:110-139); 0 outside.
        //    Column lx, kV1Seg outputs.
        if (tid < kAW * ((kAH + kV1Seg - 1) / kV1Seg)) {
            const int lx = tid % kAW, y0 = segStart(tid / kAW, kV1Seg, kAH);
            float m[5][kV1Seg];
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                float col[kV1Seg + 2 * kR];
#pragma unroll
                for (int i = 0; i < kV1Seg + 2 * kR; ++i)
                    col[i] = sH[j][y0 + i][lx];
#pragma unroll
                for (int o = 0; o < kV1Seg; ++o) {
                    float acc = 0.0f;
#pragma unroll
                    for (int t = 0; t < kTaps; ++t)
                        acc += w[t] * col[o + t];
                    m[j][o] = acc;
                }
            }
            // the segment's outputs are read by nobody else before the next barrier, and the
            // inputs sA overwrites were all read before the last one
#pragma unroll
            for (int o = 0; o < kV1Seg; ++o) {
                const int ly = y0 + o;
                const int gx = ox - kR + lx, gy = oy - kR + ly;
                float p0 = 0.0f, p1 = 0.0f, p2 = 0.0f;
                if (inImage(gx, gy)) {
                    const float mx = m[0][o], my = m[1][o];
                    const float sx = m[2][o] - mx * mx, sy = m[3][o] - my * my;
                    const float sxy = m[4][o] - mx * my;
                    const float a1 = 2 * mx * my + kC1, a2 = 2 * sxy + kC2;
                    const float b1 = mx * mx + my * my + kC1, b2 = sx + sy + kC2;
                    const float vv = (a1 * a2) / (b1 * b2);
                    if (lx >= kR && lx < kR + kTW && ly >= kR && ly < kR + kTH)
                        sV[ly - kR][lx - kR] = vv;
                    const float dmu = (2 * my * a2 - 2 * my * a1) / (b1 * b2) - vv * (2 * mx / b1 - 2 * mx / b2);
                    p0 = dLds * dmu;
                    p1 = dLds * (-vv / b2);
                    p2 = dLds * (2 * a1 / (b1 * b2));
                }
                sA[0][ly][lx] = p0;
                sA[1][ly][lx] = p1;
                sA[2][ly][lx] = p2;
            }
        }
        __syncthreads();
        {
            // the block's SSIM sum, in the order of ssimBlock
            const int lx = tid % kTW, y0 = (tid / kTW) * kV2Seg;
#pragma unroll
            for (int o = 0; o < kV2Seg; ++o)
                if (ox + lx < W && oy + y0 + o < H)
                    ssim += double(sV[y0 + o][lx]);
        }
        // 4. horizontal blur of the partials (Loss.cpp:142-157): row ly, kH2Seg outputs
        if (tid < kGH * ((kGW + kH2Seg - 1) / kH2Seg)) {
            const int ly = tid % kGH, x0 = segStart(tid / kGH, kH2Seg, kGW);
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                float row[kH2Seg + 2 * kR];
#pragma unroll
                for (int i = 0; i < kH2Seg + 2 * kR; ++i)
                    row[i] = sA[j][ly][x0 + i];
#pragma unroll
                for (int o = 0; o < kH2Seg; ++o) {
                    float acc = 0.0f;
#pragma unroll
                    for (int t = 0; t < kTaps; ++t)
                        acc += w[t] * row[o + t];
                    sG[j][ly][x0 + o] = acc; // sG aliases the H1 maps, dead since the last barrier
                }
            }
        }
        __syncthreads();
        // 5. vertical blur of the partials, L1 term and chain rule: column lx, kV2Seg pixels
        {
            const int lx = tid % kTW, y0 = segStart(tid / kTW, kV2Seg, kTH);
            float g[3][kV2Seg];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                float col[kV2Seg + 2 * kR];
#pragma unroll
                for (int i = 0; i < kV2Seg + 2 * kR; ++i)
                    col[i] = sG[j][y0 + i][lx];
#pragma unroll
                for (int o = 0; o < kV2Seg; ++o) {
                    float acc = 0.0f;
#pragma unroll
                    for (int t = 0; t < kTaps; ++t)
                        acc += w[t] * col[o + t];
                    g[j][o] = acc;
                }
            }
#pragma unroll
            for (int o = 0; o < kV2Seg; ++o) {
                const int gx = ox + lx, gy = oy + y0 + o;
                if (gx < W && gy < H) {
                    const float gp[3] = {g[0][o], g[1][o], g[2][o]};
                    finish(gx, gy, gp);
                }
            }
        }
    } else {
        const int lx = tid % kTW, y0 = (tid / kTW) * kV2Seg;
#pragma unroll
        for (int o = 0; o < kV2Seg; ++o) {
            const int gx = ox + lx, gy = oy + y0 + o;
            const float gp[3] = {0.0f, 0.0f, 0.0f};
            if (gx < W && gy < H)
                finish(gx, gy, gp);
        }
    }
    // block sums of the three doubles (fixed order)
    double v3[3] = {l1, ssim, sq};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            v3[j] += __shfl_xor(v3[j], off, 64);
    }
    if ((tid & 63) == 0) {
#pragma unroll
        for (int j = 0; j < 3; ++j)
            red[j][tid >> 6] = v3[j];
    }
    __syncthreads();
    if (tid < 3) {
        double sum = 0.0;
#pragma unroll
        for (int wv = 0; wv < kThreads / 64; ++wv)
            sum += red[tid][wv];
        partials[tid] = sum;
    }
}

// ---- The loss as two launches: SSIM partials through memory.
//
// The fused block recomputes the moment blurs and the SSIM partials over a 10-pixel halo (the
// partials' own 5-pixel halo, blurred again): 2.1x the horizontal moment blurs and 1.7x the
// SSIM points of its 32x32 pixels. With a scratch of nine floats per pixel the work splits at
// the partials: ssimBlock computes them once per pixel (moment blurs over a 5-pixel halo) and
// stores them, planar per channel; gradBlock blurs them (again a 5-pixel halo) into the
// gradient. Same expressions in the same order, so the gradient is bit-identical to the fused
// block's; the block sums are taken in the same thread order as there (the fused block routes
// its SSIM values through LDS for that), so the loss values are identical too.
constexpr int kSH = 3, kSV = 2;                         // outputs per thread: horizontal, vertical
constexpr int kMW = kTW + 2 * kR, kMH = kTH + 2 * kR;   // 42 x 42: a block with a 5-pixel halo
constexpr int kMStride = kMW + 1, kBStride = kTW + 1;   // padded LDS rows
constexpr int kHSegs = (kTW + kSH - 1) / kSH;           // horizontal runs per row
static_assert(kMH * kHSegs <= kThreads, "one horizontal run per thread");
static_assert(kTW * (kTH / kSV) == kThreads && kSV == kV2Seg, "vertical runs as the fused V2");
// Window loads: thread t < kLoadRows * kMW reads window column t % kMW of rows
// t / kMW + kLoadRows * q (one division per thread, a row step per pass).
constexpr int kLoadRows = kThreads / kMW;                        // 12
constexpr int kLoadPasses = (kMH + kLoadRows - 1) / kLoadRows;   // 4

// Map j (0..2) of channel c: maps[(3 c + j) * P + pixel].
__device__ __forceinline__ size_t mapIndex(int c, int j, size_t P, size_t pix)
{
    return size_t(3 * c + j) * P + pix;
}

// The same memory as a global-segment pointer: accesses through it use the global
// instructions with a uniform base and a 32-bit offset (a pointer read from an item table is
// otherwise a generic one, accessed with flat instructions that also count as LDS traffic in
// the waits). Offsets: images of 3 * W * H < 2^32 values (checked by the launchers).
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T *glob(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

template <bool kU8>
__device__ __forceinline__ void ssimBlock(const float *__restrict__ render,
                                          const float *__restrict__ targetF,
                                          const uint8_t *__restrict__ targetU8, int W, int H,
                                          float lambda, float invCount, const Window &win,
                                          int blocksX, int blockId, int c,
                                          float *__restrict__ maps, double *__restrict__ partials)
{
    __shared__ float2 sIn[kMH][kMStride];
    __shared__ float sH[5][kMH][kBStride];
    __shared__ double red[kThreads / 64];
    const int tid = threadIdx.x;
    float w[kTaps];
#pragma unroll
    for (int t = 0; t < kTaps; ++t)
        w[t] = win.w[t];
    const int ox = (blockId % blocksX) * kTW, oy = (blockId / blocksX) * kTH;
    const size_t P = size_t(W) * size_t(H);
    const float dLds = -lambda * invCount;
    // 1. inputs with a 5-pixel halo, zero outside the image (every load is issued: an
    //    outside position reads the image's first value and is then replaced by zero)
    {
        const auto gR = glob(render);
        const auto gT = glob(targetF);
        const auto gU = glob(targetU8);
        const int wc = tid % kMW, wr = tid / kMW;
        const int gx = ox - kR + wc;
        const bool colIn = wr < kLoadRows && gx >= 0 && gx < W;
        float2 v[kLoadPasses];
#pragma unroll
        for (int q = 0; q < kLoadPasses; ++q) {
            const int gy = oy - kR + wr + kLoadRows * q;
            const bool in = colIn && wr + kLoadRows * q < kMH && gy >= 0 && gy < H;
            const uint32_t idx = in ? (uint32_t(gy) * uint32_t(W) + uint32_t(gx)) * 3u + uint32_t(c) : 0u;
            const float r = gR[idx];
            const float t = kU8 ? float(gU[idx]) * (1.0f / 255.0f) : gT[idx];
            v[q] = in ? make_float2(r, t) : make_float2(0.0f, 0.0f);
        }
#pragma unroll
        for (int q = 0; q < kLoadPasses; ++q)
            if (wr < kLoadRows && wr + kLoadRows * q < kMH)
                sIn[wr + kLoadRows * q][wc] = v[q];
    }
    __syncthreads();
    // 2. horizontal blur of x, y, x^2, y^2, xy (Loss.cpp:77-94): window row ly, block columns
    //    x0 .. x0 + kSH - 1
    if (tid < kMH * kHSegs) {
        const int ly = tid % kMH, x0 = segStart(tid / kMH, kSH, kTW);
        float2 in[kSH + 2 * kR];
#pragma unroll
        for (int i = 0; i < kSH + 2 * kR; ++i)
            in[i] = sIn[ly][x0 + i];
#pragma unroll
        for (int o = 0; o < kSH; ++o) {
            float sx = 0, sy = 0, sxx = 0, syy = 0, sxy = 0;
#pragma unroll
            for (int t = 0; t < kTaps; ++t) {
                const float a = in[o + t].x, b = in[o + t].y;
                sx += w[t] * a;
                sy += w[t] * b;
                sxx += w[t] * a * a;
                syy += w[t] * b * b;
                sxy += w[t] * a * b;
            }
            sH[0][ly][x0 + o] = sx;
            sH[1][ly][x0 + o] = sy;
            sH[2][ly][x0 + o] = sxx;
            sH[3][ly][x0 + o] = syy;
            sH[4][ly][x0 + o] = sxy;
        }
    }
    __syncthreads();
    // 3. vertical blur, SSIM and its partial derivatives (Loss.cpp:110-139) at block column lx,
    //    rows y0, y0 + 1
    const int lx = tid % kTW, y0 = (tid / kTW) * kSV;
    float m[5][kSV];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        float col[kSV + 2 * kR];
#pragma unroll
        for (int i = 0; i < kSV + 2 * kR; ++i)
            col[i] = sH[j][y0 + i][lx];
#pragma unroll
        for (int o = 0; o < kSV; ++o) {
            float acc = 0.0f;
#pragma unroll
            for (int t = 0; t < kTaps; ++t)
                acc += w[t] * col[o + t];
            m[j][o] = acc;
        }
    }
    double ssim = 0.0;
#pragma unroll
    for (int o = 0; o < kSV; ++o) {
        const int gx = ox + lx, gy = oy + y0 + o;
        if (gx < W && gy < H) {
            const float mx = m[0][o], my = m[1][o];
            const float sx = m[2][o] - mx * mx, sy = m[3][o] - my * my;
            const float sxy = m[4][o] - mx * my;
            const float a1 = 2 * mx * my + kC1, a2 = 2 * sxy + kC2;
            const float b1 = mx * mx + my * my + kC1, b2 = sx + sy + kC2;
            const float vv = (a1 * a2) / (b1 * b2);
            ssim += double(vv);
            const float dmu = (2 * my * a2 - 2 * my * a1) / (b1 * b2) - vv * (2 * mx / b1 - 2 * mx / b2);
            const uint32_t pix = uint32_t(gy) * uint32_t(W) + uint32_t(gx);
            const auto m0 = glob(maps + mapIndex(c, 0, P, 0));
            m0[pix] = dLds * dmu;
            m0[uint32_t(P) + pix] = dLds * (-vv / b2);
            m0[2u * uint32_t(P) + pix] = dLds * (2 * a1 / (b1 * b2));
        }
    }
    ssim = blockSumD(ssim, red);
    if (tid == 0)
        partials[1] = ssim;
}

template <bool kU8, bool kScaled>
__device__ __forceinline__ void gradBlock(const float *__restrict__ render,
                                          const float *__restrict__ targetF,
                                          const uint8_t *__restrict__ targetU8, int W, int H,
                                          float lambda, float invCount, float gradScale,
                                          const Window &win, int blocksX, int blockId, int c,
                                          const float *__restrict__ maps, float *__restrict__ grad,
                                          double *__restrict__ partials)
{
    __shared__ float sA[3][kMH][kMStride];
    __shared__ float sG[3][kMH][kBStride];
    __shared__ double red[2][kThreads / 64];
    const int tid = threadIdx.x;
    float w[kTaps];
#pragma unroll
    for (int t = 0; t < kTaps; ++t)
        w[t] = win.w[t];
    const int ox = (blockId % blocksX) * kTW, oy = (blockId / blocksX) * kTH;
    const size_t P = size_t(W) * size_t(H);
    const bool ssimOn = lambda > 0.0f;
    const float l1Scale = (1.0f - lambda) * invCount;
    const int lx = tid % kTW, y0 = (tid / kTW) * kSV;
    // the thread's pixels (read first: their loads overlap the partials' below)
    float rv[kSV], tv[kSV];
    uint32_t pidx[kSV];
#pragma unroll
    for (int o = 0; o < kSV; ++o) {
        const int gx = ox + lx, gy = oy + y0 + o;
        pidx[o] = gx < W && gy < H ? (uint32_t(gy) * uint32_t(W) + uint32_t(gx)) * 3u + uint32_t(c) : 0u;
        rv[o] = glob(render)[pidx[o]];
        tv[o] = kU8 ? float(glob(targetU8)[pidx[o]]) * (1.0f / 255.0f) : glob(targetF)[pidx[o]];
    }
    float g[3][kSV] = {};
    if (ssimOn) {
        // 1. the partials with a 5-pixel halo, zero outside the image (as the fused block's);
        //    all loads are issued before the first LDS store
        const int wc = tid % kMW, wr = tid / kMW;
        const int gx = ox - kR + wc;
        const bool colIn = wr < kLoadRows && gx >= 0 && gx < W;
        const auto m0 = glob(maps + mapIndex(c, 0, P, 0));
        float v[3][kLoadPasses];
#pragma unroll
        for (int q = 0; q < kLoadPasses; ++q) {
            const int gy = oy - kR + wr + kLoadRows * q;
            const bool in = colIn && wr + kLoadRows * q < kMH && gy >= 0 && gy < H;
            const uint32_t pix = in ? uint32_t(gy) * uint32_t(W) + uint32_t(gx) : 0u;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const float x = m0[uint32_t(j) * uint32_t(P) + pix];
                v[j][q] = in ? x : 0.0f;
            }
        }
#pragma unroll
        for (int q = 0; q < kLoadPasses; ++q)
            if (wr < kLoadRows && wr + kLoadRows * q < kMH) {
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    sA[j][wr + kLoadRows * q][wc] = v[j][q];
            }
        __syncthreads();
        // 2. horizontal blur of the partials (Loss.cpp:142-157)
        if (tid < kMH * kHSegs) {
            const int ly = tid % kMH, x0 = segStart(tid / kMH, kSH, kTW);
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                float row[kSH + 2 * kR];
#pragma unroll
                for (int i = 0; i < kSH + 2 * kR; ++i)
                    row[i] = sA[j][ly][x0 + i];
#pragma unroll
                for (int o = 0; o < kSH; ++o) {
                    float acc = 0.0f;
#pragma unroll
                    for (int t = 0; t < kTaps; ++t)
                        acc += w[t] * row[o + t];
                    sG[j][ly][x0 + o] = acc;
                }
            }
        }
        __syncthreads();
        // 3. vertical blur of the partials
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            float col[kSV + 2 * kR];
#pragma unroll
            for (int i = 0; i < kSV + 2 * kR; ++i)
                col[i] = sG[j][y0 + i][lx];
#pragma unroll
            for (int o = 0; o < kSV; ++o) {
                float acc = 0.0f;
#pragma unroll
                for (int t = 0; t < kTaps; ++t)
                    acc += w[t] * col[o + t];
                g[j][o] = acc;
            }
        }
    }
    // 4. L1 term and chain rule (Loss.cpp:64-74, 160-177)
    double l1 = 0.0, sq = 0.0;
#pragma unroll
    for (int o = 0; o < kSV; ++o) {
        const int gx = ox + lx, gy = oy + y0 + o;
        if (gx < W && gy < H) {
            const float r = rv[o], t = tv[o];
            const float d = r - t;
            l1 += double(fabsf(d));
            const double dd = double(r) - double(t);
            sq += dd * dd;
            float gval = l1Scale * (d > 0 ? 1.0f : (d < 0 ? -1.0f : 0.0f));
            if (ssimOn)
                gval += g[0][o] + 2.0f * r * g[1][o] + t * g[2][o];
            if (kScaled)
                gval *= gradScale;
            glob(grad)[pidx[o]] = gval;
        }
    }
    // block sums (the fused block's reduction, component by component)
    double v2[2] = {l1, sq};
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            v2[j] += __shfl_xor(v2[j], off, 64);
    }
    if ((tid & 63) == 0) {
        red[0][tid >> 6] = v2[0];
        red[1][tid >> 6] = v2[1];
    }
    __syncthreads();
    if (tid < 2) {
        double sum = 0.0;
#pragma unroll
        for (int wv = 0; wv < kThreads / 64; ++wv)
            sum += red[tid][wv];
        partials[tid == 0 ? 0 : 2] = sum;
    }
}

// Workgroup -> (32x32 block, channel). Workgroups are dealt to the 8 XCDs in turn, so the
// three channels of block b get workgroups 24 (b / 8) + 8 c + b % 8: one XCD, dispatched
// together, so their reads of the interleaved images and their partial-line writes of the
// gradient meet in one L2 (HBM traffic of one pass instead of three).
constexpr int kLossXcds = 8;
inline int lossWorkgroups(int blocks)
{
    return 3 * ((blocks + kLossXcds - 1) / kLossXcds) * kLossXcds;
}

// With prep.segCount set, an extra first workgroup runs the segmented backward's work-item scan
// instead, dispatched ahead of the loss blocks so that it never extends the launch.
template <bool kU8>
__global__ __launch_bounds__(kThreads) void lossKernel(const float *__restrict__ render,
                                                       const float *__restrict__ targetF,
                                                       const uint8_t *__restrict__ targetU8,
                                                       int W, int H, float lambda, float invCount,
                                                       Window win, int blocksX, int blocks,
                                                       float *__restrict__ grad,
                                                       double *__restrict__ partials,
                                                       SegmentPrep prep)
{
    const int skip = prep.segCount ? 1 : 0;
    if (int(blockIdx.x) < skip) {
        segmentPrepBlock(prep);
        return;
    }
    const int x = int(blockIdx.x) - skip;
    const int grp = x / (3 * kLossXcds), r = x % (3 * kLossXcds);
    const int c = r / kLossXcds, block = grp * kLossXcds + r % kLossXcds;
    if (block >= blocks)
        return;
    lossBlock<kU8, false>(render, targetF, targetU8, W, H, lambda, invCount, 1.0f, win, blocksX,
                          block, c, grad, partials + (size_t(c) * size_t(blocks) + size_t(block)) * 3);
}

// The two launches of the split loss, with lossKernel's workgroup order (and its optional
// first workgroup running the segmented backward's work-item scan, on the first launch).
template <bool kU8>
__global__ __launch_bounds__(kThreads) void lossSsimKernel(const float *__restrict__ render,
                                                           const float *__restrict__ targetF,
                                                           const uint8_t *__restrict__ targetU8,
                                                           int W, int H, float lambda,
                                                           float invCount, Window win, int blocksX,
                                                           int blocks, float *__restrict__ maps,
                                                           double *__restrict__ partials,
                                                           SegmentPrep prep)
{
    const int skip = prep.segCount ? 1 : 0;
    if (int(blockIdx.x) < skip) {
        segmentPrepBlock(prep);
        return;
    }
    const int x = int(blockIdx.x) - skip;
    const int grp = x / (3 * kLossXcds), r = x % (3 * kLossXcds);
    const int c = r / kLossXcds, block = grp * kLossXcds + r % kLossXcds;
    if (block >= blocks)
        return;
    ssimBlock<kU8>(render, targetF, targetU8, W, H, lambda, invCount, win, blocksX, block, c, maps,
                   partials + (size_t(c) * size_t(blocks) + size_t(block)) * 3);
}

template <bool kU8>
__global__ __launch_bounds__(kThreads) void lossGradKernel(const float *__restrict__ render,
                                                           const float *__restrict__ targetF,
                                                           const uint8_t *__restrict__ targetU8,
                                                           int W, int H, float lambda,
                                                           float invCount, Window win, int blocksX,
                                                           int blocks, const float *__restrict__ maps,
                                                           float *__restrict__ grad,
                                                           double *__restrict__ partials)
{
    const int x = int(blockIdx.x);
    const int grp = x / (3 * kLossXcds), r = x % (3 * kLossXcds);
    const int c = r / kLossXcds, block = grp * kLossXcds + r % kLossXcds;
    if (block >= blocks)
        return;
    gradBlock<kU8, false>(render, targetF, targetU8, W, H, lambda, invCount, 1.0f, win, blocksX,
                          block, c, maps, grad,
                          partials + (size_t(c) * size_t(blocks) + size_t(block)) * 3);
}

// Batched split loss: item blockIdx.z; blockIdx.x -> (32x32 block, channel) in lossKernel's
// order, so the three channels of a block share an XCD's L2 (interleaved images).
__global__ __launch_bounds__(kThreads) void lossSsimBatchedKernel(const LossItem *__restrict__ items,
                                                                  float lambda, Window win)
{
    const LossItem &it = items[blockIdx.z];
    const int x = int(blockIdx.x);
    const int grp = x / (3 * kLossXcds), r = x % (3 * kLossXcds);
    const int c = r / kLossXcds, block = grp * kLossXcds + r % kLossXcds;
    if (block >= it.numBlocks)
        return;
    ssimBlock<true>(it.render, nullptr, it.target, it.width, it.height, lambda, it.invCount, win,
                    it.blocksX, block, c, it.maps,
                    it.partials + (size_t(c) * it.numBlocks + block) * 3);
}

__global__ __launch_bounds__(kThreads) void lossGradBatchedKernel(const LossItem *__restrict__ items,
                                                                  float lambda, Window win)
{
    const LossItem &it = items[blockIdx.z];
    const int x = int(blockIdx.x);
    const int grp = x / (3 * kLossXcds), r = x % (3 * kLossXcds);
    const int c = r / kLossXcds, block = grp * kLossXcds + r % kLossXcds;
    if (block >= it.numBlocks)
        return;
    gradBlock<true, true>(it.render, nullptr, it.target, it.width, it.height, lambda, it.invCount,
                          it.gradScale, win, it.blocksX, block, c, it.maps, it.grad,
                          it.partials + (size_t(c) * it.numBlocks + block) * 3);
}

// Batched: item blockIdx.z, channel blockIdx.y; blocks past the item's count exit.
__global__ __launch_bounds__(kThreads) void lossBatchedKernel(const LossItem *__restrict__ items,
                                                              float lambda, Window win)
{
    const LossItem &it = items[blockIdx.z];
    if (int(blockIdx.x) >= it.numBlocks)
        return;
    lossBlock<true, true>(it.render, nullptr, it.target, it.width, it.height, lambda, it.invCount,
                          it.gradScale, win, it.blocksX, int(blockIdx.x), int(blockIdx.y),
                          it.grad, it.partials + (size_t(blockIdx.y) * it.numBlocks + blockIdx.x) * 3);
}

__global__ __launch_bounds__(kThreads) void lossFinalize(LossFinalize f) { lossFinalizeBlock(f); }

__global__ __launch_bounds__(kThreads) void lossFinalizeBatched(const LossItem *__restrict__ items,
                                                                float lambda, StepStatus status)
{
    if (blockIdx.x == 0 && threadIdx.x == 0 && status.ring) {
        // the step's pair counts are final (the binning ran before the compositing)
        const uint32_t slot = *status.ring;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            __hip_atomic_store(&status.host[4 * slot + k], status.total[k], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_SYSTEM);
        *status.ring = (slot + 1) % uint32_t(status.ringSize);
    }
    const LossItem &it = items[blockIdx.x];
    LossFinalize f;
    f.partials = it.partials;
    f.numBlocks = 3 * it.numBlocks;
    f.lambda = lambda;
    f.invCount = it.invCount;
    f.count = double(it.width) * double(it.height) * 3.0;
    f.result = it.result;
    lossFinalizeBlock(f);
}

// Squared error / absolute error sums for evaluation passes.
template <int kMode> // 0: squared error, 1: absolute error
__global__ __launch_bounds__(kThreads) void errSumKernel(const float *__restrict__ a,
                                                         const float *__restrict__ b,
                                                         const uint8_t *__restrict__ bU8,
                                                         int64_t count, double *partials)
{
    __shared__ double red[kThreads / 64];
    double s = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < count;
         i += int64_t(gridDim.x) * kThreads) {
        const double x = double(a[i]);
        const double y = bU8 ? double(float(bU8[i]) * (1.0f / 255.0f)) : double(b[i]);
        const double d = x - y;
        s += kMode == 0 ? d * d : fabs(d);
    }
    s = blockSumD(s, red);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = s;
}

template <int kMode>
__global__ __launch_bounds__(kThreads) void errFinalize(const double *__restrict__ partials,
                                                        int numBlocks, double count, float *result)
{
    __shared__ double red[kThreads / 64];
    double s = 0.0;
    for (int i = threadIdx.x; i < numBlocks; i += kThreads)
        s += partials[i];
    s = blockSumD(s, red);
    if (threadIdx.x == 0) {
        const double m = s / (count > 1.0 ? count : 1.0);
        if (kMode == 0)
            result[0] = m <= 1e-12 ? 99.0f : float(-10.0 * log10(m));
        else
            result[0] = float(m);
    }
}

Window makeWindow()
{
    // Loss.cpp:15-27: w_k = exp(-k^2 / (2 * 1.5^2)) normalised, in float with the host expf.
    static const uint64_t table[32] = U3D_EXP2_TABLE;
    Window w;
    float sum = 0.0f;
    for (int k = -kR; k <= kR; ++k) {
        const float x = -float(k * k) / (2.0f * 1.5f * 1.5f);
        w.w[k + kR] = u3dmath::expf_core(x, table);
        sum += w.w[k + kR];
    }
    for (float &v : w.w)
        v /= sum;
    return w;
}

constexpr int kErrBlocks = 1024;

} // namespace

size_t lossPartialsSize(int width, int height)
{
    const int bx = (width + kTW - 1) / kTW, by = (height + kTH - 1) / kTH;
    return size_t(bx) * size_t(by) * 9;
}

int launchPhotometricLoss(hipStream_t s, const float *render, const float *target,
                          const uint8_t *targetU8, int width, int height, float lambda,
                          float *grad, double *partials, size_t partialsCap, float *result,
                          LossFinalize *deferred, const SegmentPrep *prep, float *ssimMaps)
{
    const int bx = (width + kTW - 1) / kTW, by = (height + kTH - 1) / kTH;
    const int nb = bx * by; // per channel
    U3D_CHECK(size_t(nb) * 9 <= partialsCap, U3D_ERR_CAPACITY, "loss partials too small");
    const size_t P = size_t(width) * size_t(height);
    const float invCount = 1.0f / float(P * 3);
    static const Window win = makeWindow();
    const SegmentPrep side = prep ? *prep : SegmentPrep{};
    const int wgs = (nb > 0 ? lossWorkgroups(nb) : 0) + (prep ? 1 : 0);
    // the split kernels index an image with 32-bit offsets
    if (ssimMaps && lambda > 0.0f && nb > 0 && P * 3 < (size_t(1) << 32)) {
        if (targetU8)
            hipLaunchKernelGGL(lossSsimKernel<true>, dim3(wgs), dim3(kThreads), 0, s, render,
                               target, targetU8, width, height, lambda, invCount, win, bx, nb,
                               ssimMaps, partials, side);
        else
            hipLaunchKernelGGL(lossSsimKernel<false>, dim3(wgs), dim3(kThreads), 0, s, render,
                               target, targetU8, width, height, lambda, invCount, win, bx, nb,
                               ssimMaps, partials, side);
        U3D_LAUNCH_CHECK();
        const int wg2 = lossWorkgroups(nb);
        if (targetU8)
            hipLaunchKernelGGL(lossGradKernel<true>, dim3(wg2), dim3(kThreads), 0, s, render,
                               target, targetU8, width, height, lambda, invCount, win, bx, nb,
                               ssimMaps, grad, partials);
        else
            hipLaunchKernelGGL(lossGradKernel<false>, dim3(wg2), dim3(kThreads), 0, s, render,
                               target, targetU8, width, height, lambda, invCount, win, bx, nb,
                               ssimMaps, grad, partials);
        U3D_LAUNCH_CHECK();
    } else if (wgs > 0) {
        if (targetU8)
            hipLaunchKernelGGL(lossKernel<true>, dim3(wgs), dim3(kThreads), 0, s, render,
                               target, targetU8, width, height, lambda, invCount, win, bx, nb,
                               grad, partials, side);
        else
            hipLaunchKernelGGL(lossKernel<false>, dim3(wgs), dim3(kThreads), 0, s, render,
                               target, targetU8, width, height, lambda, invCount, win, bx, nb,
                               grad, partials, side);
        U3D_LAUNCH_CHECK();
    }
    LossFinalize f;
    f.partials = partials;
    f.numBlocks = 3 * nb;
    f.lambda = lambda;
    f.invCount = invCount;
    f.count = double(P * 3);
    f.result = result;
    if (deferred) {
        *deferred = f;
        return U3D_OK;
    }
    return launchLossFinalize(s, f);
}

int launchLossFinalize(hipStream_t s, const LossFinalize &f)
{
    hipLaunchKernelGGL(lossFinalize, dim3(1), dim3(kThreads), 0, s, f);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

void lossItemGeometry(int width, int height, int *blocksX, int *numBlocks)
{
    const int bx = (width + kTW - 1) / kTW, by = (height + kTH - 1) / kTH;
    *blocksX = bx;
    *numBlocks = bx * by;
}

int launchPhotometricLossBatched(hipStream_t s, const LossItem *items, int numItems,
                                 int maxBlocks, float lambda, const StepStatus &status,
                                 bool splitMaps)
{
    if (numItems <= 0)
        return U3D_OK;
    static const Window win = makeWindow();
    if (maxBlocks > 0 && splitMaps && lambda > 0.0f) {
        const dim3 grid(unsigned(lossWorkgroups(maxBlocks)), 1, unsigned(numItems));
        hipLaunchKernelGGL(lossSsimBatchedKernel, grid, dim3(kThreads), 0, s, items, lambda, win);
        U3D_LAUNCH_CHECK();
        hipLaunchKernelGGL(lossGradBatchedKernel, grid, dim3(kThreads), 0, s, items, lambda, win);
        U3D_LAUNCH_CHECK();
    } else if (maxBlocks > 0) {
        hipLaunchKernelGGL(lossBatchedKernel, dim3(maxBlocks, 3, numItems), dim3(kThreads), 0, s,
                           items, lambda, win);
        U3D_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(lossFinalizeBatched, dim3(numItems), dim3(kThreads), 0, s, items, lambda,
                       status);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

size_t psnrPartialsSize(int64_t) { return kErrBlocks; }

static int launchErr(hipStream_t s, int mode, const float *a, const float *b, const uint8_t *bU8,
                     int64_t count, double *partials, size_t cap, float *result)
{
    U3D_CHECK(cap >= size_t(kErrBlocks), U3D_ERR_CAPACITY, "error partials too small");
    int nb = int((count + kThreads - 1) / kThreads);
    nb = nb < 1 ? 1 : (nb > kErrBlocks ? kErrBlocks : nb);
    if (mode == 0) {
        hipLaunchKernelGGL(errSumKernel<0>, dim3(nb), dim3(kThreads), 0, s, a, b, bU8, count, partials);
        U3D_LAUNCH_CHECK();
        hipLaunchKernelGGL(errFinalize<0>, dim3(1), dim3(kThreads), 0, s, partials, nb, double(count), result);
    } else {
        hipLaunchKernelGGL(errSumKernel<1>, dim3(nb), dim3(kThreads), 0, s, a, b, bU8, count, partials);
        U3D_LAUNCH_CHECK();
        hipLaunchKernelGGL(errFinalize<1>, dim3(1), dim3(kThreads), 0, s, partials, nb, double(count), result);
    }
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchPsnr(hipStream_t s, const float *a, const float *b, const uint8_t *bU8, int64_t count,
               double *partials, size_t partialsCap, float *result)
{
    return launchErr(s, 0, a, b, bU8, count, partials, partialsCap, result);
}

int launchMeanAbs(hipStream_t s, const float *a, const uint8_t *bU8, int64_t count,
                  double *partials, size_t partialsCap, float *result)
{
    return launchErr(s, 1, a, nullptr, bU8, count, partials, partialsCap, result);
}

} // namespace u3d
