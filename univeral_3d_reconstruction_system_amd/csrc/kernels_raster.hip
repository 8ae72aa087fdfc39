// Tile compositing, forward (Rasterizer.cpp:263-330) and backward stage 1 (Rasterizer.cpp:
// 334-427), for gfx950.
//
// The single-image forward with 16-pixel tiles is rasterForwardQuadKernel (a workgroup per 8x8
// quadrant whose wavefronts split the walk, forwardQuadrant below). Elsewhere (larger tiles, the
// batched engine, the backward): one 256-thread workgroup per 16x16 pixel block; each wave64
// owns one 8x8 quadrant, one pixel per lane. A logical tile of tile_size (a multiple of 16) is
// covered by (tile_size/16)^2 blocks that share its list. The tile's Gaussians are staged
// through LDS in batches of records (36 B each, gathered once per batch, one per thread).
//
// Exact culling. While staging, every Gaussian gets a 4-bit mask of the quadrants in which it
// can contribute: the minimum of its quadratic form over the quadrant's pixel rectangle
// (interior or the closest point of one of the four edges) must bring the power above
// log(min_alpha / alpha) - 1e-3, with a rounding margin. A quadrant's wave then visits only
// the masked Gaussians, in list order, iterating set bits of a ballot. Culled pairs are pairs
// the reference skips anyway (alpha < min_alpha), so the sequence of contributing pairs, and
// every result, is unchanged.
//
// Forward: front-to-back with the reference's exact arithmetic (contraction off, host-identical
// expf), so images, final transmittance and contributor counts are bit-identical. Pairs below
// the per-Gaussian power threshold are rejected before any exp. The workgroup stops once every
// pixel has saturated.
//
// Backward: reverse walk from each wave's largest contributor count; per Gaussian the nine 2D
// gradient terms of the wave's 64 pixels are summed with DPP (only when a lane contributes;
// a 16-lane-row variant feeding LDS atomics measured slower because of atomic conflicts), the
// four waves combine in an LDS batch accumulator, and each batch is flushed with one global
// float atomic per non-zero term per (block, Gaussian).
#include "kernels.h"

namespace u3d {

namespace {

constexpr int kThreads = 256;

constexpr int kEdge = 16;
constexpr int kPlainFewLanes = 0; // lane threshold of the plain backward (see backwardBlock)
// Lane threshold of the segmented backward: a Gaussian reached by at most this many lanes of a
// wavefront has their terms added to LDS directly instead of reduced over the wavefront
// (measured: 4, 8, 16 lanes -> 121, 120, 131 us against 123 us without).
constexpr int kSegFewLanes = 8;

template <int CTRL, int ROW_MASK> __device__ __forceinline__ float dppMove(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROW_MASK, 0xf, false));
}

// Sum over the 64 lanes of the wave (every lane must be active): inclusive row scans with DPP
// row shifts, then row broadcasts; lane 63 holds the total.
__device__ __forceinline__ float waveSum(float x)
{
    x += dppMove<0x111, 0xf>(x); // row_shr:1
    x += dppMove<0x112, 0xf>(x); // row_shr:2
    x += dppMove<0x114, 0xf>(x); // row_shr:4
    x += dppMove<0x118, 0xf>(x); // row_shr:8
    x += dppMove<0x142, 0xa>(x); // row_bcast:15 into rows 1, 3
    x += dppMove<0x143, 0xc>(x); // row_bcast:31 into rows 2, 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}

// Sums of nine per-lane values over the 64 lanes of the wavefront (every lane active) by
// recursive halving: at each lane distance a lane hands half of its partial sums to its partner
// and keeps the other half (v_permlane32_swap / v_permlane16_swap for distances 32 and 16, DPP
// row rotate by 8, half-row mirror and quad permutes below), 25 instructions for all nine.
// Lane l with (l & 3) == 0 ends with the total of value k = bit5 + 2 bit4 + 4 bit3 of l when
// bit 2 of l is clear; lane 4 ends with value 8; *index is k, 8 or -1.
__device__ __forceinline__ float pairSum32(float x, float y)
{
    // lanes 0-31: x[l] + x[l + 32]; lanes 32-63: y[l - 32] + y[l]
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

__device__ __forceinline__ float pairSum16(float x, float y)
{
    // rows 0, 2: x summed with the next row; rows 1, 3: y summed with the previous row
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// DPP exchange with a partner whose `upper` flag is the opposite: keeps x where !upper, y where
// upper, each summed with the partner's copy.
template <int CTRL> __device__ __forceinline__ float pairSumDpp(float x, float y, bool upper)
{
    const float send = upper ? x : y, keep = upper ? y : x;
    return keep + dppMove<CTRL, 0xf>(send);
}

__device__ __forceinline__ float waveReduceScatter9(const float v[9], int lane, int *index)
{
    const float p0 = pairSum32(v[0], v[1]), p1 = pairSum32(v[2], v[3]);
    const float p2 = pairSum32(v[4], v[5]), p3 = pairSum32(v[6], v[7]);
    float e = pairSum32(v[8], v[8]);
    const float q0 = pairSum16(p0, p1), q1 = pairSum16(p2, p3);
    e = pairSum16(e, e);
    const bool b3 = (lane >> 3) & 1, b2 = (lane >> 2) & 1;
    const float r0 = pairSumDpp<0x128>(q0, q1, b3); // row_ror:8 = lane xor 8
    e += dppMove<0x128, 0xf>(e);
    float t = pairSumDpp<0x141>(r0, e, b2);         // row_half_mirror = lane xor 7
    t += dppMove<0x4e, 0xf>(t);                     // quad_perm [2,3,0,1] = lane xor 2
    t += dppMove<0xb1, 0xf>(t);                     // quad_perm [1,0,3,2] = lane xor 1
    *index = (lane & 3) ? -1 : (b2 ? (lane == 4 ? 8 : -1) : ((lane >> 5) & 1) | (((lane >> 4) & 1) << 1) | (int(b3) << 2));
    return t;
}

// State of one pixel's reverse walk (Rasterizer.cpp:370-430): transmittance and the colour
// behind the next contributor. The reference blends the previous contributor's colour into the
// colour behind when it reaches the next one; here the blend (the same expression, so the same
// value) is taken right after the contributor is walked, which keeps four state registers
// instead of eight.
struct BackState {
    float T, acc0, acc1, acc2;
};

// One Gaussian of the reverse walk for every lane of the wavefront (uniform control flow):
// lanes with `active` and a contribution update their state with the reference's arithmetic
// (kept by selects) and the Gaussian's nine screen-space gradient terms are summed over the
// wavefront into accRow (LDS). kFewLanes > 0: when at most that many lanes contribute they add
// their terms to accRow directly instead.
// kFastExp (min_alpha >= 1e-37, so every power that passes the threshold test lies in
// (-86, 0]): the gradient is held to 1e-4, not to bit-exactness, so G comes from the hardware
// exp2 (relative error below 6e-6 for these powers) and 1 / (1 - alpha) from the hardware
// reciprocal (1 ulp). Which Gaussians contribute and which are clamped must still be the
// forward's decisions: a lane whose alpha lies within 2e-5 (relative) of min_alpha or of the
// 0.99 clamp takes the exact exp instead.
// kSkipIdle: return at once when no lane is in range. The segmented walk, whose entries come
// from the forward's cull, rarely meets such a Gaussian and runs faster without the branch
// (the compiler copies the walk state around it); the whole-list walk culls less and keeps it.
template <int kFewLanes, bool kFastExp, bool kSkipIdle = false>
__device__ __forceinline__ void backStep(BackState &st, bool active, const float4 &a, const float4 &b,
                                         float thr, float c2, float fpx, float fpy, float dp0,
                                         float dp1, float dp2, float fT, float bgDot, float minAlpha,
                                         const uint64_t *stab, int lane, float *accRow)
{
    const float dx = a.x - fpx, dy = a.y - fpy;
    const float power = -0.5f * (a.z * dx * dx + b.x * dy * dy) - a.w * dx * dy;
    // bitwise tests: short-circuit ones become branches whose merges copy the walk state
    const bool inRange = active & !(power > 0.0f) & !(power < thr);
    if (kSkipIdle && __ballot(inRange) == 0)
        return;
    const float op = b.y;
    float G, rawAlpha;
    if (kFastExp) {
        G = __builtin_amdgcn_exp2f(power * 1.44269504088896341f);
        rawAlpha = op * G;
        const bool nearCut = inRange & ((fabsf(rawAlpha - minAlpha) <= 2e-5f * minAlpha) |
                                        (fabsf(rawAlpha - kMaxAlpha) <= 2e-5f * kMaxAlpha));
        if (__ballot(nearCut) != 0 && nearCut) {
            G = expfExactLdsIf(power, stab, true);
            rawAlpha = op * G;
        }
    } else {
        G = expfExactLdsIf(power, stab, false);
        rawAlpha = op * G;
    }
    const float alpha = (rawAlpha < kMaxAlpha) ? rawAlpha : kMaxAlpha;
    const bool contrib = inRange & !(alpha < minAlpha);
    // A lane in range almost always contributes (the threshold is within 1e-3 of min_alpha),
    // so only the few-lane path tests for no contribution; otherwise the selects below leave
    // the state unchanged.
    // (no early exit when no lane contributes: the selects below keep the state and the
    // few-lane path then adds nothing; a branch here made the compiler copy the walk state)
    const uint64_t who = kFewLanes > 0 ? __ballot(contrib) : 0;
    const float c0 = b.z, c1 = b.w;
    const float oneMinus = 1.0f - alpha;
    const float inv = kFastExp ? __builtin_amdgcn_rcpf(oneMinus) : 0.0f;
    const float Tn = kFastExp ? st.T * inv : st.T / oneMinus;
    const float w = alpha * Tn;
    // (the reference's sum starts from +0; starting from the first term differs only in the
    // sign of a zero sum, which no later sum keeps)
    float dAlpha = (c0 - st.acc0) * dp0;
    dAlpha += (c1 - st.acc1) * dp1;
    dAlpha += (c2 - st.acc2) * dp2;
    dAlpha *= Tn;
    if (bgDot != 0.0f || bgDot != bgDot) // the background term; adding +-0 changes nothing
        dAlpha += kFastExp ? -fT * inv * bgDot : -fT / (1.0f - alpha) * bgDot;
    const bool grads = contrib & !(rawAlpha > kMaxAlpha);
    const float dPower = op * dAlpha * G;
    const float A = a.z, B = a.w, C = b.x;
    float v[9];
    v[0] = grads ? dPower * -(A * dx + B * dy) : 0.0f;
    v[1] = grads ? dPower * -(C * dy + B * dx) : 0.0f;
    v[2] = grads ? dPower * -0.5f * dx * dx : 0.0f;
    v[3] = grads ? dPower * -dx * dy : 0.0f;
    v[4] = grads ? dPower * -0.5f * dy * dy : 0.0f;
    v[5] = contrib ? w * dp0 : 0.0f;
    v[6] = contrib ? w * dp1 : 0.0f;
    v[7] = contrib ? w * dp2 : 0.0f;
    v[8] = grads ? G * dAlpha : 0.0f;
    st.T = contrib ? Tn : st.T;
    st.acc0 = contrib ? alpha * c0 + oneMinus * st.acc0 : st.acc0;
    st.acc1 = contrib ? alpha * c1 + oneMinus * st.acc1 : st.acc1;
    st.acc2 = contrib ? alpha * c2 + oneMinus * st.acc2 : st.acc2;
    if (kFewLanes > 0 && __popcll(who) <= kFewLanes) {
        if (contrib) {
#pragma unroll
            for (int k = 0; k < 9; ++k)
                atomicAdd(&accRow[k], v[k]);
        }
        return;
    }
    int idx;
    const float tot = waveReduceScatter9(v, lane, &idx);
    if (idx >= 0)
        atomicAdd(&accRow[idx], tot);
}

// log(min_alpha / alpha) - 1e-3: below it, min(0.99, alpha * exp(power)) < min_alpha holds for
// the rounded float result as well (logf error and the two roundings are < 1e-5).
// kFast (min_alpha >= 1e-37): from the hardware log2 (within 1 ulp), ln 2 (log2 min_alpha -
// log2 alpha) - 1e-3; its error, below 3e-5 for any float alpha, stays far inside the 1e-3
// margin, so the threshold still lies below the exact one. A denormal or zero alpha gives +inf
// (no contribution: alpha below min_alpha).
template <bool kFast>
__device__ __forceinline__ float skipThreshold(float minAlpha, float alpha)
{
    if (kFast)
        return (__builtin_amdgcn_logf(minAlpha) - __builtin_amdgcn_logf(alpha)) * 0.693147180559945309f - 1e-3f;
    return minAlpha > 0.0f ? logf(minAlpha / alpha) - 1e-3f : -INFINITY;
}

// With minAlpha >= 1e-37 every threshold is above ln(1e-37) - 1e-3 > -86 (opacity <= 1), so a
// power that passes the threshold test lies in (-86, 0] or is NaN: no special exp inputs.
__device__ __forceinline__ bool noSpecialExp(float minAlpha) { return minAlpha >= 1e-37f; }

// Minimum of Q(d) = A dx^2 + 2B dx dy + C dy^2 over the box d in [x0, x1] x [y0, y1] (Q positive
// definite); also returns a bound on the magnitude of the terms for the rounding margin.
__device__ __forceinline__ float boxMinQuad(float A, float B, float C, float x0, float x1, float y0,
                                            float y1, float *scale)
{
    const float dmax = fmaxf(fmaxf(fabsf(x0), fabsf(x1)), fmaxf(fabsf(y0), fabsf(y1)));
    *scale = (fabsf(A) + 2.0f * fabsf(B) + fabsf(C)) * dmax * dmax;
    if (x0 <= 0.0f && 0.0f <= x1 && y0 <= 0.0f && 0.0f <= y1)
        return 0.0f;
    auto q = [&](float dx, float dy) { return A * dx * dx + 2.0f * B * dx * dy + C * dy * dy; };
    auto clampf = [](float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); };
    // edges dx = x0, x1 (minimise over dy) and dy = y0, y1 (minimise over dx). Any point of an
    // edge gives an upper bound of the minimum; a minimiser off by a few ulps raises it by
    // ~1e-13 * scale, far inside the caller's 1e-6 * scale margin, so the hardware reciprocal
    // (1 ulp) serves for the two ratios (A, C > 0 here).
    const float bc = B * __builtin_amdgcn_rcpf(C), ba = B * __builtin_amdgcn_rcpf(A);
    const float ya = clampf(-bc * x0, y0, y1), yb = clampf(-bc * x1, y0, y1);
    const float xa = clampf(-ba * y0, x0, x1), xb = clampf(-ba * y1, x0, x1);
    return fminf(fminf(q(x0, ya), q(x1, yb)), fminf(q(xa, y0), q(xb, y1)));
}

// Whether one staged Gaussian may contribute to a pixel of the 8x8 quadrant with top-left
// pixel (qx0, qy0): false only when the power bound over the quadrant is certainly below the
// threshold.
__device__ __forceinline__ bool quadrantHit(float4 a, float C, float thr, int qx0, int qy0)
{
    // Only positive-definite conics have the box minimum used below; anything else is kept.
    if (!(thr > -INFINITY) || !(a.z > 0.0f && C > 0.0f && a.z * C - a.w * a.w > 0.0f))
        return true;
    const float px0 = float(qx0), py0 = float(qy0);
    // d = mean - pixel over pixel centres [p0, p0 + 7]
    float scale;
    const float qmin = boxMinQuad(a.z, a.w, C, a.x - (px0 + 7.0f), a.x - px0, a.y - (py0 + 7.0f),
                                  a.y - py0, &scale);
    const float powerMax = -0.5f * qmin + 1e-6f * scale + 1e-6f;
    return !(powerMax < thr); // keep unless certainly below the threshold (NaN keeps)
}

// Whether one staged Gaussian may contribute to a pixel of the box with pixel centres
// [x0, x1] x [y0, y1]: false only when the power bound over the box is certainly below the
// threshold (quadrantHit for any box).
__device__ __forceinline__ bool boxHit(float4 a, float C, float thr, int x0, int x1, int y0, int y1)
{
    if (!(thr > -INFINITY) || !(a.z > 0.0f && C > 0.0f && a.z * C - a.w * a.w > 0.0f))
        return true;
    float scale;
    const float qmin = boxMinQuad(a.z, a.w, C, a.x - float(x1), a.x - float(x0), a.y - float(y1),
                                  a.y - float(y0), &scale);
    const float powerMax = -0.5f * qmin + 1e-6f * scale + 1e-6f;
    return !(powerMax < thr);
}

// Quadrant mask of one staged Gaussian for the block at (bx0, by0): bit w set when the
// Gaussian may contribute to a pixel of quadrant w (8x8, w = qy * 2 + qx).
__device__ __forceinline__ uint32_t quadrantMask(float4 a, float C, float thr, int bx0, int by0)
{
    // Only positive-definite conics have the box minimum used below; anything else is kept.
    if (!(thr > -INFINITY) || !(a.z > 0.0f && C > 0.0f && a.z * C - a.w * a.w > 0.0f))
        return 0xfu;
    uint32_t mask = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const float px0 = float(bx0 + (w & 1) * 8), py0 = float(by0 + (w >> 1) * 8);
        // d = mean - pixel over pixel centres [p0, p0 + 7]
        float scale;
        const float qmin = boxMinQuad(a.z, a.w, C, a.x - (px0 + 7.0f), a.x - px0,
                                      a.y - (py0 + 7.0f), a.y - py0, &scale);
        const float powerMax = -0.5f * qmin + 1e-6f * scale + 1e-6f;
        // keep unless certainly below the threshold (NaN keeps)
        if (!(powerMax < thr))
            mask |= 1u << w;
    }
    return mask;
}

// One 16x16 pixel block (bx, by) of an image whose tiles start at global tile id tileBase.
// Each wavefront walks the list for its own 8x8 quadrant, staging 64 entries at a time in its
// own LDS slice: no workgroup barrier couples the four walks, and a wavefront whose pixels are
// all done stops. kBatch: the checkpoint segment length (entries); kCkpt: seg.ckpt is set
// (segment checkpoints for the segmented backward are written).
template <int kBatch, bool kCkpt, bool kFastExp>
__device__ __forceinline__ void forwardBlock(
    int bx, int by, int tileBase, int W, int H, int ts, int tilesX, const int *__restrict__ tileOffsets,
    const int *__restrict__ tileList, const float4 *__restrict__ rec0,
    const float4 *__restrict__ rec1, const float *__restrict__ rec2, float bg0, float bg1,
    float bg2, float minAlpha, float minT, float *__restrict__ rgb, float *__restrict__ finalT,
    int *__restrict__ nContrib, const SegmentState &seg)
{
    constexpr int kStep = 64;
    static_assert(kBatch % kStep == 0, "segments are whole steps");
    __shared__ float4 s0[kThreads / 64][kStep];
    __shared__ float4 s1[kThreads / 64][kStep];
    __shared__ float s2[kThreads / 64][kStep];
    __shared__ float sthr[kThreads / 64][kStep];
    __shared__ uint64_t stab[32];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (tid < 32)
        stab[tid] = kExp2Table[tid];
    __syncthreads();
    const int bx0 = bx * kEdge, by0 = by * kEdge;
    const int qx0 = bx0 + (wave & 1) * 8, qy0 = by0 + (wave >> 1) * 8;
    const int px = qx0 + (lane & 7), py = qy0 + (lane >> 3);
    const int tile = tileBase + (by0 / ts) * tilesX + bx0 / ts;
    const bool inside = px < W && py < H;
    const int begin = tileOffsets[tile], end = tileOffsets[tile + 1];
    const float fpx = float(px), fpy = float(py);
    float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    float sr = 0.0f, sg = 0.0f, sb = 0.0f; // colour of the current segment (checkpoints only)
    int last = 0, batches = 0;             // batches: checkpoint segments started
    bool done = !inside;
    const size_t slot0 = size_t(begin / kBatch + tile);
    float4 *ls0 = s0[wave], *ls1 = s1[wave];
    float *ls2 = s2[wave], *lthr = sthr[wave];
    // The next step's records (and the ids of the one after) are loaded while the current
    // step is walked.
    auto loadId = [&](int b0) { return b0 + lane < end ? tileList[b0 + lane] : -1; };
    int gCur = loadId(begin), gAhead = loadId(begin + kStep);
    float4 ra = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rb = ra;
    float rc = 0.0f;
    if (gCur >= 0) {
        ra = rec0[gCur];
        rb = rec1[gCur];
        rc = rec2[gCur];
    }
    for (int base = begin; base < end; base += kStep) {
        if (__ballot(!done) == 0)
            break;
        if ((base - begin) % kBatch == 0) {
            if (kCkpt && base > begin) { // T before this segment and the previous one's colour
                if (inside)
                    seg.ckpt[(slot0 + size_t(batches)) * kThreads + tid] = make_float4(T, sr, sg, sb);
                sr = sg = sb = 0.0f;
            }
            ++batches;
        }
        bool hit = false;
        if (gCur >= 0) {
            const float thr = skipThreshold<kFastExp>(minAlpha, rb.y);
            ls0[lane] = ra;
            ls1[lane] = rb;
            ls2[lane] = rc;
            lthr[lane] = thr;
            hit = quadrantHit(ra, rb.x, thr, qx0, qy0);
        }
        uint64_t bits = __ballot(hit);
        gCur = gAhead;
        if (gCur >= 0) {
            ra = rec0[gCur];
            rb = rec1[gCur];
            rc = rec2[gCur];
        }
        gAhead = loadId(base + 2 * kStep);
        waveSync();
        {
            // One Gaussian for every pixel of the wavefront, without divergent branches: the
            // contribution is computed for all lanes and kept by selects where the reference
            // composites (same expressions, so the same values); lanes outside the window,
            // below the thresholds or already done keep their state. The pixel-local part
            // (power, exp, alpha and the reference's skip tests) does not depend on the walk's
            // state, so several Gaussians' parts are computed together and their long
            // dependency chains overlap; the sequential transmittance / colour update follows.
            struct Pre {
                float alpha, c0, c1, c2;
                bool cand;
            };
            auto prepare = [&](const float4 &a, const float4 &b, float thr, float bl) -> Pre {
                const float dx = a.x - fpx, dy = a.y - fpy;
                const float power = -0.5f * (a.z * dx * dx + b.x * dy * dy) - a.w * dx * dy;
                const float raw = b.y * expfExactLdsIf(power, stab, kFastExp);
                const float alpha = (raw < kMaxAlpha) ? raw : kMaxAlpha;
                const bool cand = !(power > 0.0f) && !(power < thr) && !(alpha < minAlpha);
                return Pre{alpha, b.z, b.w, bl, cand};
            };
            auto composite = [&](const Pre &p, int j) {
                const bool contrib = p.cand && !done;
                const float nextT = T * (1.0f - p.alpha);
                const bool stop = contrib && nextT < minT;
                const bool take = contrib && !stop;
                const float w = p.alpha * T;
                const float wr = p.c0 * w, wg = p.c1 * w, wb = p.c2 * w;
                cr = take ? cr + wr : cr;
                cg = take ? cg + wg : cg;
                cb = take ? cb + wb : cb;
                if (kCkpt) {
                    sr = take ? sr + wr : sr;
                    sg = take ? sg + wg : sg;
                    sb = take ? sb + wb : sb;
                }
                T = take ? nextT : T;
                last = take ? base - begin + j + 1 : last;
                done = done || stop;
            };
            // Groups of kGroup Gaussians: their pixel-local parts are independent dependency
            // chains computed side by side, then composited in list order.
            constexpr int kGroup = 4;
            while (bits) {
                int jj[kGroup];
                Pre pp[kGroup];
#pragma unroll
                for (int k = 0; k < kGroup; ++k) {
                    const bool has = bits != 0; // wave-uniform
                    jj[k] = has ? __builtin_ctzll(bits) : -1;
                    bits &= bits - 1;
                    const int jl = has ? jj[k] : 0;
                    pp[k] = prepare(ls0[jl], ls1[jl], lthr[jl], ls2[jl]);
                }
#pragma unroll
                for (int k = 0; k < kGroup; ++k)
                    if (jj[k] >= 0)
                        composite(pp[k], jj[k]);
            }
        }
        waveSync(); // the slice is restaged next step
    }
    if (inside) {
        const size_t pix = size_t(py) * size_t(W) + size_t(px);
        rgb[pix * 3] = cr + T * bg0;
        rgb[pix * 3 + 1] = cg + T * bg1;
        rgb[pix * 3 + 2] = cb + T * bg2;
        finalT[pix] = T;
        nContrib[pix] = last;
        if (kCkpt) {
            // Reverse pass over this pixel's checkpoints: replace each segment colour by the
            // colour of everything from that boundary on (suffix sums of positive terms, so
            // the backward needs no cancelling difference).
            float ar = sr, ag = sg, ab = sb;
            for (int k = batches - 1; k >= 1; --k) {
                float4 *c = seg.ckpt + (slot0 + size_t(k)) * kThreads + tid;
                const float4 v = *c;
                *c = make_float4(v.x, ar, ag, ab);
                ar += v.y;
                ag += v.z;
                ab += v.w;
            }
        }
    }
    if (kCkpt) {
        // segments of this tile the backward visits: ceil(max contributor count / kBatch)
        __shared__ int sMax[kThreads / 64];
        int m = last;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            m = max(m, __shfl_xor(m, off, 64));
        if (lane == 0)
            sMax[wave] = m;
        __syncthreads();
        if (tid == 0)
            seg.segCount[tile] = uint32_t((max(max(sMax[0], sMax[1]), max(sMax[2], sMax[3])) + kBatch - 1) / kBatch);
    }
}

// Forward of one 8x8 quadrant of a 16x16 tile by a whole 256-thread workgroup. A wavefront
// walking a long list alone issues ~90 instructions per Gaussian, most of them the pixel-local
// part (power, exact exp, alpha, the skip tests), which does not depend on the walk's state;
// the sequential update is ~20. Here wavefront 0 only composites; wavefronts 1-3 take turns
// staging the records of each step of kQStep list entries, and the two not staging compute the
// pixel-local parts of the next step's entries into LDS, half each. Between two barriers: stage
// step s + 2, prepare step s + 1, composite step s. The values and the update are the same
// expressions in the same order as forwardBlock's, so every output (colour, final
// transmittance, contributor count, checkpoints) is identical to it.
//
// Staging keeps only the entries that pass the quadrant cull, packed in list order (rank k):
// pos[k] = mean and conic (x, y, A, B), shape[k] = (C, opacity, skip threshold, -),
// col[k] = (colour, list position + 1 as int bits). An entry that is not a candidate for a
// pixel (power above 0 or below the threshold, alpha below min_alpha) is passed as alpha = NaN:
// a candidate's alpha is never NaN, since a NaN product fails `raw < 0.99` and becomes 0.99.
// The ranks are padded with NaN up to a multiple of kCGroup, the compositor's group.
//
// A pixel that is done (saturated) takes no further contribution, so the cull of a step only
// needs the box of the pixels still taking contributions: after each step the compositor
// publishes that box, and the steps staged later are culled against it (pixels only ever
// leave the set, so the box is never too small).
constexpr int kQStep = 64;
struct QuadLds {
    float4 pos[2][kQStep];   // by step parity: read while preparing, written while staging
    float4 shape[2][kQStep];
    float4 col[3][kQStep];   // by step % 3: also read while compositing
    float alpha[2][kQStep][64];
    int count[3];
    uint64_t tab[32];
    int stop[2];  // by step parity: a flag is rewritten only after every wave has read it
    int alive[2]; // by step parity: box of the pixels still taking contributions (4 x 4 bits)
};

template <bool kCkpt, bool kFastExp>
__device__ __forceinline__ void forwardQuadrant(
    int bx, int by, int quad, int W, int H, int tilesX, const int *__restrict__ tileOffsets,
    const int *__restrict__ tileList, const float4 *__restrict__ rec0,
    const float4 *__restrict__ rec1, const float *__restrict__ rec2, float bg0, float bg1,
    float bg2, float minAlpha, float minT, float *__restrict__ rgb, float *__restrict__ finalT,
    int *__restrict__ nContrib, const SegmentState &seg, unsigned long long *evalCount,
    uint32_t *walkCost)
{
    constexpr int kBatch = kSegEntries; // checkpoint segment, list entries
    static_assert(kBatch % kQStep == 0, "segments are whole steps");
    constexpr int kGroup = 4, kCGroup = 8, kPrep = 3;
    constexpr int kFullBox = 0x7070; // x in [0, 7], y in [0, 7]
    __shared__ QuadLds L;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int qx0 = bx * kEdge + (quad & 1) * 8, qy0 = by * kEdge + (quad >> 1) * 8;
    const int px = qx0 + (lane & 7), py = qy0 + (lane >> 3);
    const int tile = by * tilesX + bx;
    const bool inside = px < W && py < H;
    const int begin = tileOffsets[tile], end = tileOffsets[tile + 1];
    const int steps = (end - begin + kQStep - 1) / kQStep;
    const float fpx = float(px), fpy = float(py);
    if (tid < 32)
        L.tab[tid] = kExp2Table[tid];
    if (wave == 0)
        __builtin_amdgcn_s_setprio(3); // the sequential walk is the critical path

    // Staging: wavefront w (1-3) stages the steps s with s % 3 == w - 1; the records of its
    // next step and the ids of the one after are in registers (loaded kPrep steps ahead).
    int gCur = -1, gNext = -1;
    float4 ra = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rb = ra;
    float rc = 0.0f;
    auto loadId = [&](int st) {
        const int e = begin + st * kQStep + lane;
        return e < end ? tileList[e] : -1;
    };
    auto loadRec = [&]() {
        if (gCur >= 0) {
            ra = rec0[gCur];
            rb = rec1[gCur];
            rc = rec2[gCur];
        }
    };
    auto stage = [&](int st, int box) {
        const int b = st % 3;
        bool hit = false;
        float thr = 0.0f;
        if (gCur >= 0) {
            thr = skipThreshold<kFastExp>(minAlpha, rb.y);
            hit = boxHit(ra, rb.x, thr, qx0 + (box & 15), qx0 + ((box >> 4) & 15),
                         qy0 + ((box >> 8) & 15), qy0 + ((box >> 12) & 15));
        }
        const uint64_t bits = __ballot(hit);
        if (kCkpt && seg.hitBits && lane == 0 && st < steps)
            seg.hitBits[(size_t(begin / kQStep + tile) + size_t(st)) * 4 + quad] = bits;
        if (hit) {
            const int k = int(__builtin_amdgcn_mbcnt_hi(uint32_t(bits >> 32),
                                                        __builtin_amdgcn_mbcnt_lo(uint32_t(bits), 0u)));
            L.pos[st & 1][k] = ra;
            L.shape[st & 1][k] = make_float4(rb.x, rb.y, thr, 0.0f);
            L.col[b][k] = make_float4(rb.z, rb.w, rc, __int_as_float(st * kQStep + lane + 1));
        }
        if (lane == 0)
            L.count[b] = __popcll(bits);
        gCur = gNext;
        loadRec();
        gNext = loadId(st + 2 * kPrep);
    };
    auto stages = [&](int st) { return st % 3 == wave - 1; };

    // Wavefronts 1-3 except the one staging alongside (wavefront 1 + (st + 1) % 3, which
    // stages step st + 1 while step st is prepared): alpha (NaN when not a candidate) of the
    // ranks k = idx (mod 2) of step st, padded to a multiple of kCGroup; idx numbers the two.
    auto prepareStep = [&](int st) {
        const int b = st % 3, a = st & 1;
        const int stager = 1 + (st + 1) % 3;
        if (wave == stager)
            return;
        const int idx = wave < stager ? wave - 1 : wave - 2;
        const int n = __builtin_amdgcn_readfirstlane(L.count[b]);
        const int nPad = (n + kCGroup - 1) / kCGroup * kCGroup;
        for (int k0 = idx; k0 < nPad; k0 += 2 * kGroup) {
            float al[kGroup];
#pragma unroll
            for (int g = 0; g < kGroup; ++g) {
                const int k = k0 + 2 * g;
                const bool valid = k < n; // wave-uniform
                const int kl = valid ? k : 0;
                const float4 ga = L.pos[a][kl], gs = L.shape[a][kl];
                const float dx = ga.x - fpx, dy = ga.y - fpy;
                const float power = -0.5f * (ga.z * dx * dx + gs.x * dy * dy) - ga.w * dx * dy;
                const float raw = gs.y * expfExactLdsIf(power, L.tab, kFastExp);
                const float alpha = (raw < kMaxAlpha) ? raw : kMaxAlpha;
                // bitwise: a short-circuit test lets the compiler branch around each exp,
                // which serialises the group's dependency chains
                const bool cand = valid & !(power > 0.0f) & !(power < gs.z) & !(alpha < minAlpha);
                al[g] = cand ? alpha : __builtin_nanf("");
            }
#pragma unroll
            for (int g = 0; g < kGroup; ++g)
                if (k0 + 2 * g < nPad)
                    L.alpha[a][k0 + 2 * g][lane] = al[g];
        }
    };

    // Wavefront 0's walk state.
    float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    float sr = 0.0f, sg = 0.0f, sb = 0.0f; // colour of the current segment (checkpoints only)
    int last = 0, batches = 0;             // batches: checkpoint segments started
    int lastSlot = 0;                      // the pixel's last checkpoint written
    // +0 while the pixel takes contributions, NaN once it is done (saturated or outside the
    // image): alpha + doneNaN is NaN exactly when the entry is no contribution.
    float doneNaN = inside ? 0.0f : __builtin_nanf("");
    const size_t slot0 = size_t(begin / kBatch + tile);
    const int ckTid = quad * 64 + lane; // the pixel's index in the tile's checkpoint rows
    auto compositeStep = [&](int st) {
        const int b = st % 3, a = st & 1;
        const int n = __builtin_amdgcn_readfirstlane(L.count[b]);
        for (int k0 = 0; k0 < n; k0 += kCGroup) {
            float al[kCGroup];
            float4 cv[kCGroup];
#pragma unroll
            for (int g = 0; g < kCGroup; ++g) {
                al[g] = L.alpha[a][k0 + g][lane];
                cv[g] = L.col[b][k0 + g];
            }
            // Bitwise logic on the flags (no short-circuit branches). A skipped term is added
            // as +0, which leaves every sum unchanged: the sums start at +0 and cannot become
            // -0, and x + 0 == x for every other x.
#pragma unroll
            for (int g = 0; g < kCGroup; ++g) {
                const bool contrib = !__builtin_isnan(al[g] + doneNaN);
                const float nextT = T * (1.0f - al[g]);
                const bool stop = contrib & (nextT < minT);
                const bool take = contrib & !stop;
                const float w = al[g] * T;
                const float wr = take ? cv[g].x * w : 0.0f;
                const float wg = take ? cv[g].y * w : 0.0f;
                const float wb = take ? cv[g].z * w : 0.0f;
                cr += wr;
                cg += wg;
                cb += wb;
                if (kCkpt) {
                    sr += wr;
                    sg += wg;
                    sb += wb;
                }
                T = take ? nextT : T;
                last = take ? __float_as_int(cv[g].w) : last;
                doneNaN = stop ? __builtin_nanf("") : doneNaN;
            }
        }
    };

    if (wave > 0) {
        gCur = loadId(wave - 1);
        gNext = loadId(wave - 1 + kPrep);
        loadRec();
        if (wave <= 2)
            stage(wave - 1, kFullBox);
    }
    if (tid == 0)
        L.alive[1] = kFullBox; // the box the stagers of the first step read
    __syncthreads();
    if (wave > 0 && steps > 0)
        prepareStep(0);
    __syncthreads();
    for (int st = 0; st < steps; ++st) {
        if (wave == 0) {
            const bool allDone = __ballot(!__builtin_isnan(doneNaN)) == 0;
            if (lane == 0)
                L.stop[st & 1] = allDone ? 1 : 0;
            if (!allDone) {
                if ((st * kQStep) % kBatch == 0) {
                    if (kCkpt && st > 0) {
                        // T before this segment and the previous one's colour, for a pixel
                        // still walking: the backward reads a boundary's checkpoint only for
                        // pixels with more contributors than the boundary. A done pixel keeps
                        // its last segment's colour (later segments add nothing to it).
                        const bool walking = !__builtin_isnan(doneNaN);
                        if (inside && walking)
                            seg.ckpt[(slot0 + size_t(batches)) * 256 + ckTid] = make_float4(T, sr, sg, sb);
                        sr = walking ? 0.0f : sr;
                        sg = walking ? 0.0f : sg;
                        sb = walking ? 0.0f : sb;
                        lastSlot = walking ? batches : lastSlot;
                    }
                    ++batches;
                }
                compositeStep(st);
                // box of the pixels still taking contributions (rows and columns of the mask)
                const uint64_t live = __ballot(!__builtin_isnan(doneNaN));
                uint32_t cols = uint32_t(live | (live >> 32));
                cols |= cols >> 16;
                cols |= cols >> 8;
                cols &= 0xffu;
                uint32_t rows = 0;
#pragma unroll
                for (int r = 0; r < 8; ++r)
                    rows |= ((live >> (8 * r)) & 0xffu) ? (1u << r) : 0u;
                const int box = live ? (__builtin_ctz(cols) | ((31 - __builtin_clz(cols)) << 4) |
                                        (__builtin_ctz(rows) << 8) | ((31 - __builtin_clz(rows)) << 12))
                                     : kFullBox;
                if (lane == 0)
                    L.alive[st & 1] = box;
            }
        } else {
            if (stages(st + 2))
                stage(st + 2, __builtin_amdgcn_readfirstlane(L.alive[(st + 1) & 1]));
            if (st + 1 < steps)
                prepareStep(st + 1);
        }
        __syncthreads();
        if (L.stop[st & 1])
            break;
    }
    if (wave != 0)
        return;
    if (inside) {
        const size_t pix = size_t(py) * size_t(W) + size_t(px);
        rgb[pix * 3] = cr + T * bg0;
        rgb[pix * 3 + 1] = cg + T * bg1;
        rgb[pix * 3 + 2] = cb + T * bg2;
        finalT[pix] = T;
        nContrib[pix] = last;
        if (kCkpt) {
            // Reverse pass over this pixel's checkpoints: each segment colour becomes the colour
            // of everything from that boundary on (as forwardBlock; the same values in every
            // checkpoint written). Eight checkpoints are loaded before any is rewritten, so
            // their round trips overlap.
            float ar = sr, ag = sg, ab = sb;
            constexpr int kAhead = 8;
            for (int k = lastSlot; k >= 1; k -= kAhead) {
                float4 v[kAhead];
#pragma unroll
                for (int u = 0; u < kAhead; ++u)
                    if (k - u >= 1)
                        v[u] = seg.ckpt[(slot0 + size_t(k - u)) * 256 + ckTid];
#pragma unroll
                for (int u = 0; u < kAhead; ++u) {
                    if (k - u >= 1) {
                        seg.ckpt[(slot0 + size_t(k - u)) * 256 + ckTid] = make_float4(v[u].x, ar, ag, ab);
                        ar += v[u].y;
                        ag += v[u].z;
                        ab += v[u].w;
                    }
                }
            }
        }
    }
    if (kCkpt) {
        // this quadrant's segments: ceil(max contributor count / kBatch); the tile's count is
        // the largest of its four quadrants'
        int m = last;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            m = max(m, __shfl_xor(m, off, 64));
        if (lane == 0)
            seg.segCount[size_t(tile) * kSegCountsPerTile + quad] = uint32_t((m + kBatch - 1) / kBatch);
    }
    if (walkCost) {
        // how far this quadrant walked (its largest contributor count), for the ordering of
        // the next forward of the same view
        int m = inside ? last : 0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            m = max(m, __shfl_xor(m, off, 64));
        if (lane == 0 && m > 0)
            atomicMax(&walkCost[tile], unsigned(m));
    }
    if (evalCount) {
        // evaluated (pixel, Gaussian) pairs of this quadrant: the sum of its contributor counts,
        // added to one of 64 counters so that the quadrants do not all hit one address
        unsigned sum = inside ? unsigned(last) : 0u;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            sum += __shfl_xor(sum, off, 64);
        if (lane == 0)
            atomicAdd(&evalCount[blockIdx.x & 63], (unsigned long long)sum);
    }
}

// Reverse walk over one segment [lo, lo + kBatch) of a 16x16 tile's list. A pixel whose last
// contributor lies in the segment starts from its final state, as backwardBlock; a pixel that
// continues past the segment starts from the forward's checkpoint at the segment's end:
// transmittance T there and the colour composited from there on, which divided by T is the
// running "colour behind" of the reference's reverse recursion. Segments of a tile are
// independent work items.
template <int kBatch, bool kFastExp>
__device__ __forceinline__ void backwardSegment(
    int bx, int by, int tileBase, int segIdx, int W, int H, int tilesX,
    const int *__restrict__ tileOffsets, const int *__restrict__ tileList,
    const float4 *__restrict__ rec0, const float4 *__restrict__ rec1, const float *__restrict__ rec2,
    float bg0, float bg1, float bg2, float minAlpha, const float *__restrict__ dl,
    const float *__restrict__ finalTArr, const int *__restrict__ nContribArr,
    const float4 *__restrict__ ckpt, float *__restrict__ grad2d,
    const uint64_t *__restrict__ hitBits = nullptr)
{
    __shared__ float4 s0[kBatch];
    __shared__ float4 s1[kBatch];
    __shared__ float s2[kBatch];
    __shared__ float sthr[kBatch];
    __shared__ int sid[kBatch];
    __shared__ uint32_t smask[kBatch];
    __shared__ float acc[kBatch][kGrad2dStride];
    __shared__ uint64_t stab[32];
    __shared__ int waveMax[4];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (tid < 32)
        stab[tid] = kExp2Table[tid];
    for (int k = tid; k < kBatch * kGrad2dStride; k += kThreads)
        (&acc[0][0])[k] = 0.0f;
    const int bx0 = bx * kEdge, by0 = by * kEdge;
    const int px = bx0 + (wave & 1) * 8 + (lane & 7), py = by0 + (wave >> 1) * 8 + (lane >> 3);
    const int tile = tileBase + by * tilesX + bx;
    const bool inside = px < W && py < H;
    const int begin = tileOffsets[tile];
    const int lo = segIdx * kBatch, hi = lo + kBatch;
    // The segment's records are loaded before the walk's extent is known (their latency
    // overlaps the per-pixel loads and the barrier below).
    const int segLen = min(kBatch, tileOffsets[tile + 1] - begin - lo);
    int g = -1;
    float4 ra = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rb = ra;
    float rc = 0.0f;
    if (tid < segLen) {
        g = tileList[begin + lo + tid];
        ra = rec0[g];
        rb = rec1[g];
        rc = rec2[g];
    }
    const size_t pix = inside ? size_t(py) * size_t(W) + size_t(px) : 0;
    const float fT = inside ? finalTArr[pix] : 0.0f;
    const int nc = inside ? nContribArr[pix] : 0;
    const int jEnd = nc < hi ? nc : hi; // this pixel walks j = jEnd - 1 ... lo
    const float dp0 = inside ? dl[pix * 3] : 0.0f;
    const float dp1 = inside ? dl[pix * 3 + 1] : 0.0f;
    const float dp2 = inside ? dl[pix * 3 + 2] : 0.0f;
    const float bgDot = bg0 * dp0 + bg1 * dp1 + bg2 * dp2;
    const float fpx = float(px), fpy = float(py);
    BackState bs{fT, 0.0f, 0.0f, 0.0f};
    if (nc > hi) {
        const float4 ck = ckpt[(size_t(begin / kBatch + tile) + size_t(segIdx + 1)) * kThreads + tid];
        bs.T = ck.x;
        if (bs.T > 0.0f) {
            bs.acc0 = ck.y / bs.T;
            bs.acc1 = ck.z / bs.T;
            bs.acc2 = ck.w / bs.T;
        }
    }
    {
        int m = jEnd;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            m = max(m, __shfl_xor(m, off, 64));
        if (lane == 0)
            waveMax[wave] = m;
    }
    __syncthreads();
    const int wm0 = waveMax[0], wm1 = waveMax[1], wm2 = waveMax[2], wm3 = waveMax[3];
    const int maxN = max(max(wm0, wm1), max(wm2, wm3));
    if (maxN <= lo)
        return; // uniform: no pixel of the block reaches this segment
    const int cnt = maxN - lo; // <= segLen: contributor counts do not exceed the list
    uint32_t m = 0;
    if (tid < cnt) {
        const int j = lo + tid;
        const float thr = skipThreshold<kFastExp>(minAlpha, rb.y);
        sid[tid] = g;
        s0[tid] = ra;
        s1[tid] = rb;
        s2[tid] = rc;
        sthr[tid] = thr;
        if (hitBits) {
            // the forward's cull of this entry (valid for every pixel that walks it here: a
            // pixel with more than j contributors was live when the forward culled entry j)
            const uint64_t *w = hitBits + (size_t(begin / 64 + tile) + size_t(j / 64)) * 4;
            const int bit = j & 63;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                m |= uint32_t((w[q] >> bit) & 1u) << q;
        } else {
            m = quadrantMask(ra, rb.x, thr, bx0, by0);
        }
        m &= (j < wm0 ? 1u : 0u) | (j < wm1 ? 2u : 0u) | (j < wm2 ? 4u : 0u) | (j < wm3 ? 8u : 0u);
    }
    if (tid < kBatch)
        smask[tid] = m;
    __syncthreads();
    for (int c = kBatch / 64 - 1; c >= 0; --c) {
        uint64_t bits = __ballot((smask[c * 64 + lane] >> wave) & 1u);
        if (!bits)
            continue;
        // Reverse walk over the set bits, two register sets in turn: the next Gaussian's LDS
        // records are loaded while the current one is processed (same Gaussians, same order).
        auto step = [&](const float4 &ra, const float4 &rb, float rthr, float rc2, int jj) {
            backStep<kSegFewLanes, kFastExp>(bs, lo + jj < jEnd, ra, rb, rthr, rc2, fpx, fpy, dp0,
                                             dp1, dp2, fT, bgDot, minAlpha, stab, lane, acc[jj]);
        };
        int jA = c * 64 + 63 - __builtin_clzll(bits);
        bits &= ~(uint64_t(1) << (jA - c * 64));
        float4 aA = s0[jA], bA = s1[jA];
        float thrA = sthr[jA], cA = s2[jA];
        for (;;) {
            if (!bits) {
                step(aA, bA, thrA, cA, jA);
                break;
            }
            const int jB = c * 64 + 63 - __builtin_clzll(bits);
            bits &= ~(uint64_t(1) << (jB - c * 64));
            const float4 aB = s0[jB], bB = s1[jB];
            const float thrB = sthr[jB], cB = s2[jB];
            step(aA, bA, thrA, cA, jA);
            if (!bits) {
                step(aB, bB, thrB, cB, jB);
                break;
            }
            jA = c * 64 + 63 - __builtin_clzll(bits);
            bits &= ~(uint64_t(1) << (jA - c * 64));
            aA = s0[jA];
            bA = s1[jA];
            thrA = sthr[jA];
            cA = s2[jA];
            step(aB, bB, thrB, cB, jB);
        }
    }
    __syncthreads();
    if (tid < cnt) {
        float *dst = grad2d + size_t(sid[tid]) * kGrad2dStride;
#pragma unroll
        for (int k = 0; k < kGrad2dStride; ++k) {
            const float x = acc[tid][k];
            if (x != 0.0f)
                unsafeAtomicAdd(&dst[k], x);
        }
    }
}

// kFewLanes: when at most this many lanes of a wave contribute to a Gaussian, they add their
// values to its LDS accumulator directly instead of paying nine 64-lane reductions (0: always
// reduce).
template <int kBatch, int kFewLanes, bool kFastExp>
__device__ __forceinline__ void backwardBlock(
    int bx, int by, int tileBase, int W, int H, int ts, int tilesX, const int *__restrict__ tileOffsets,
    const int *__restrict__ tileList, const float4 *__restrict__ rec0,
    const float4 *__restrict__ rec1, const float *__restrict__ rec2, float bg0, float bg1,
    float bg2, float minAlpha, const float *__restrict__ dl, const float *__restrict__ finalTArr,
    const int *__restrict__ nContribArr, float *__restrict__ grad2d)
{
    __shared__ float4 s0[kBatch];
    __shared__ float4 s1[kBatch];
    __shared__ float s2[kBatch];
    __shared__ float sthr[kBatch];
    __shared__ int sid[kBatch];
    __shared__ uint32_t smask[kBatch];
    __shared__ float acc[kBatch][kGrad2dStride];
    __shared__ uint64_t stab[32];
    __shared__ int waveMax[4];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (tid < 32)
        stab[tid] = kExp2Table[tid];
    for (int k = tid; k < kBatch * kGrad2dStride; k += kThreads)
        (&acc[0][0])[k] = 0.0f;
    const int bx0 = bx * kEdge, by0 = by * kEdge;
    const int px = bx0 + (wave & 1) * 8 + (lane & 7), py = by0 + (wave >> 1) * 8 + (lane >> 3);
    const int tile = tileBase + (by0 / ts) * tilesX + bx0 / ts;
    const bool inside = px < W && py < H;
    const int begin = tileOffsets[tile];
    const size_t pix = inside ? size_t(py) * size_t(W) + size_t(px) : 0;
    const float fT = inside ? finalTArr[pix] : 0.0f;
    const int nc = inside ? nContribArr[pix] : 0;
    const float dp0 = inside ? dl[pix * 3] : 0.0f;
    const float dp1 = inside ? dl[pix * 3 + 1] : 0.0f;
    const float dp2 = inside ? dl[pix * 3 + 2] : 0.0f;
    const float bgDot = bg0 * dp0 + bg1 * dp1 + bg2 * dp2;
    const float fpx = float(px), fpy = float(py);
    BackState bs{fT, 0.0f, 0.0f, 0.0f};
    {
        int m = nc;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            m = max(m, __shfl_xor(m, off, 64));
        if (lane == 0)
            waveMax[wave] = m;
    }
    __syncthreads();
    const int wm0 = waveMax[0], wm1 = waveMax[1], wm2 = waveMax[2], wm3 = waveMax[3];
    const int maxN = max(max(wm0, wm1), max(wm2, wm3));
    for (int hi = maxN; hi > 0; hi -= kBatch) {
        const int lo = hi > kBatch ? hi - kBatch : 0;
        const int cnt = hi - lo;
        uint32_t m = 0;
        if (tid < cnt) {
            const int j = lo + tid;
            const int g = tileList[begin + j];
            const float4 a = rec0[g];
            const float4 b = rec1[g];
            const float thr = skipThreshold<kFastExp>(minAlpha, b.y);
            sid[tid] = g;
            s0[tid] = a;
            s1[tid] = b;
            s2[tid] = rec2[g];
            sthr[tid] = thr;
            m = quadrantMask(a, b.x, thr, bx0, by0);
            m &= (j < wm0 ? 1u : 0u) | (j < wm1 ? 2u : 0u) | (j < wm2 ? 4u : 0u) | (j < wm3 ? 8u : 0u);
        }
        if (tid < kBatch)
            smask[tid] = m;
        __syncthreads();
        for (int c = kBatch / 64 - 1; c >= 0; --c) {
            uint64_t bits = __ballot((smask[c * 64 + lane] >> wave) & 1u);
            if (!bits)
                continue;
            // Reverse walk over the set bits. Measured against backwardSegment's double-buffered
            // walk: this plain loop needs 30 fewer register copies per Gaussian and is faster
            // here (the batched engine keeps more wavefronts per CU to hide the LDS latency),
            // slower there.
            while (bits) {
                const int jj = c * 64 + 63 - __builtin_clzll(bits);
                bits &= ~(uint64_t(1) << (jj - c * 64));
                backStep<kFewLanes, kFastExp, true>(bs, lo + jj < nc, s0[jj], s1[jj], sthr[jj], s2[jj], fpx,
                                              fpy, dp0, dp1, dp2, fT, bgDot, minAlpha, stab, lane,
                                              acc[jj]);
            }
        }
        __syncthreads();
        if (tid < cnt) {
            float *dst = grad2d + size_t(sid[tid]) * kGrad2dStride;
#pragma unroll
            for (int k = 0; k < kGrad2dStride; ++k) {
                const float x = acc[tid][k];
                if (x != 0.0f) {
                    unsafeAtomicAdd(&dst[k], x);
                    acc[tid][k] = 0.0f;
                }
            }
        }
        __syncthreads();
    }
}

// Index of the item owning global raster block b (items sorted by blockBase).
__device__ __forceinline__ int itemOfBlock(const RenderItem *__restrict__ items, int numItems,
                                           int64_t b)
{
    int lo = 0, hi = numItems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].blockBase <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// 16-pixel tiles: four workgroups per tile, one per quadrant, the tiles in blockOrder's order
// when given. Workgroup x runs on XCD x % 8; the four quadrants of a tile get workgroups
// x, x + 8, x + 16, x + 24 (one XCD), so their gathers of the tile's records share an L2. The
// grid is 32 * ceil(tiles / 8) workgroups; those past the last tile exit.
template <bool kCkpt, bool kFastExp>
__global__ __launch_bounds__(kThreads) void rasterForwardQuadKernel(
    int W, int H, int tilesX, int numTiles, const int *__restrict__ tileOffsets,
    const int *__restrict__ tileList, const float4 *__restrict__ rec0,
    const float4 *__restrict__ rec1, const float *__restrict__ rec2, float bg0, float bg1,
    float bg2, float minAlpha, float minT, float *__restrict__ rgb, float *__restrict__ finalT,
    int *__restrict__ nContrib, SegmentState seg, const int *__restrict__ blockOrder,
    unsigned long long *evalCount, uint32_t *walkCost)
{
    const int x = int(blockIdx.x), k = x >> 3;
    const int b = (k >> 2) * 8 + (x & 7);
    if (b >= numTiles)
        return;
    const int blk = blockOrder ? blockOrder[b] : b;
    forwardQuadrant<kCkpt, kFastExp>(blk % tilesX, blk / tilesX, k & 3, W, H, tilesX,
                                     tileOffsets, tileList, rec0, rec1, rec2, bg0, bg1, bg2,
                                     minAlpha, minT, rgb, finalT, nContrib, seg, evalCount,
                                     walkCost);
}

template <bool kCkpt, bool kFastExp>
__global__ __launch_bounds__(kThreads) void rasterForwardKernel(
    int W, int H, int ts, int tilesX, int blocksX, const int *__restrict__ tileOffsets,
    const int *__restrict__ tileList, const float4 *__restrict__ rec0,
    const float4 *__restrict__ rec1, const float *__restrict__ rec2, float bg0, float bg1,
    float bg2, float minAlpha, float minT, float *__restrict__ rgb, float *__restrict__ finalT,
    int *__restrict__ nContrib, SegmentState seg, const int *__restrict__ blockOrder)
{
    const int blk = blockOrder ? blockOrder[blockIdx.x] : int(blockIdx.x);
    forwardBlock<kSegEntries, kCkpt, kFastExp>(blk % blocksX, blk / blocksX, 0, W, H, ts,
                                       tilesX, tileOffsets, tileList, rec0, rec1, rec2, bg0, bg1,
                                       bg2, minAlpha, minT, rgb, finalT, nContrib, seg);
}

// Tiles of any size (RasterSettings::tileSize is free in the reference, Rasterizer.cpp:297-300):
// one thread per pixel walks its own tile's list with the records read from memory -- the
// lanes of a wavefront may lie in different tiles, so nothing is staged in LDS. The forward
// takes forwardBlock's expressions in the same order (same image, final transmittance and
// contributor counts); the backward adds each contribution with float atomics.
__device__ __forceinline__ bool anyTilePixel(int W, int H, int blocksX, int &px, int &py)
{
    const int bx = int(blockIdx.x) % blocksX, by = int(blockIdx.x) / blocksX;
    px = bx * kEdge + int(threadIdx.x & 15);
    py = by * kEdge + int(threadIdx.x >> 4);
    return px < W && py < H;
}

__global__ __launch_bounds__(kThreads) void rasterForwardAnyKernel(
    int W, int H, int ts, int tilesX, int blocksX, const int *__restrict__ tileOffsets,
    const int *__restrict__ tileList, const float4 *__restrict__ rec0,
    const float4 *__restrict__ rec1, const float *__restrict__ rec2, float bg0, float bg1,
    float bg2, float minAlpha, float minT, float *__restrict__ rgb, float *__restrict__ finalT,
    int *__restrict__ nContrib)
{
    __shared__ uint64_t stab[32];
    if (threadIdx.x < 32)
        stab[threadIdx.x] = kExp2Table[threadIdx.x];
    __syncthreads();
    int px, py;
    if (!anyTilePixel(W, H, blocksX, px, py))
        return;
    const int tile = (py / ts) * tilesX + px / ts;
    const int begin = tileOffsets[tile], end = tileOffsets[tile + 1];
    const float fpx = float(px), fpy = float(py);
    float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    int last = 0;
    for (int j = begin; j < end; ++j) {
        const int g = tileList[j];
        const float4 a = rec0[g], b = rec1[g];
        const float dx = a.x - fpx, dy = a.y - fpy;
        const float power = -0.5f * (a.z * dx * dx + b.x * dy * dy) - a.w * dx * dy;
        if (power > 0.0f)
            continue;
        const float raw = b.y * expfExactLdsIf(power, stab, false);
        const float alpha = (raw < kMaxAlpha) ? raw : kMaxAlpha;
        if (alpha < minAlpha)
            continue;
        const float nextT = T * (1.0f - alpha);
        if (nextT < minT)
            break;
        const float w = alpha * T;
        cr = cr + b.z * w;
        cg = cg + b.w * w;
        cb = cb + rec2[g] * w;
        T = nextT;
        last = j - begin + 1;
    }
    const size_t pix = size_t(py) * size_t(W) + size_t(px);
    rgb[pix * 3] = cr + T * bg0;
    rgb[pix * 3 + 1] = cg + T * bg1;
    rgb[pix * 3 + 2] = cb + T * bg2;
    finalT[pix] = T;
    nContrib[pix] = last;
}

__global__ __launch_bounds__(kThreads) void rasterBackwardAnyKernel(
    int W, int H, int ts, int tilesX, int blocksX, const int *__restrict__ tileOffsets,
    const int *__restrict__ tileList, const float4 *__restrict__ rec0,
    const float4 *__restrict__ rec1, const float *__restrict__ rec2, float bg0, float bg1,
    float bg2, float minAlpha, const float *__restrict__ dl, const float *__restrict__ finalTArr,
    const int *__restrict__ nContribArr, float *__restrict__ grad2d)
{
    __shared__ uint64_t stab[32];
    if (threadIdx.x < 32)
        stab[threadIdx.x] = kExp2Table[threadIdx.x];
    __syncthreads();
    int px, py;
    if (!anyTilePixel(W, H, blocksX, px, py))
        return;
    const int tile = (py / ts) * tilesX + px / ts;
    const int begin = tileOffsets[tile];
    const size_t pix = size_t(py) * size_t(W) + size_t(px);
    const float fT = finalTArr[pix];
    const float dp0 = dl[pix * 3], dp1 = dl[pix * 3 + 1], dp2 = dl[pix * 3 + 2];
    const float bgDot = bg0 * dp0 + bg1 * dp1 + bg2 * dp2;
    const float fpx = float(px), fpy = float(py);
    // the walk state of the reference's reverse pass (Rasterizer.cpp:379-417)
    float T = fT, acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f;
    float last0 = 0.0f, last1 = 0.0f, last2 = 0.0f, lastAlpha = 0.0f;
    for (int j = nContribArr[pix] - 1; j >= 0; --j) {
        const int g = tileList[begin + j];
        const float4 a = rec0[g], b = rec1[g];
        const float c2 = rec2[g];
        const float dx = a.x - fpx, dy = a.y - fpy;
        const float power = -0.5f * (a.z * dx * dx + b.x * dy * dy) - a.w * dx * dy;
        if (power > 0.0f)
            continue;
        const float G = expfExactLdsIf(power, stab, false);
        const float rawAlpha = b.y * G;
        const float alpha = (rawAlpha < kMaxAlpha) ? rawAlpha : kMaxAlpha;
        if (alpha < minAlpha)
            continue;
        T = T / (1.0f - alpha);
        const float w = alpha * T;
        acc0 = lastAlpha * last0 + (1.0f - lastAlpha) * acc0;
        acc1 = lastAlpha * last1 + (1.0f - lastAlpha) * acc1;
        acc2 = lastAlpha * last2 + (1.0f - lastAlpha) * acc2;
        last0 = b.z;
        last1 = b.w;
        last2 = c2;
        float dAlpha = 0.0f;
        dAlpha += (b.z - acc0) * dp0;
        dAlpha += (b.w - acc1) * dp1;
        dAlpha += (c2 - acc2) * dp2;
        dAlpha *= T;
        lastAlpha = alpha;
        dAlpha += -fT / (1.0f - alpha) * bgDot;
        float *dst = grad2d + size_t(g) * kGrad2dStride;
        const float v5 = w * dp0, v6 = w * dp1, v7 = w * dp2;
        if (v5 != 0.0f)
            atomicAdd(dst + 5, v5);
        if (v6 != 0.0f)
            atomicAdd(dst + 6, v6);
        if (v7 != 0.0f)
            atomicAdd(dst + 7, v7);
        if (rawAlpha > kMaxAlpha)
            continue; // clamped: no gradient through alpha
        const float dPower = b.y * dAlpha * G;
        const float v[6] = {dPower * -(a.z * dx + a.w * dy), dPower * -(b.x * dy + a.w * dx),
                            dPower * -0.5f * dx * dx, dPower * -dx * dy, dPower * -0.5f * dy * dy,
                            G * dAlpha};
#pragma unroll
        for (int k = 0; k < 5; ++k)
            if (v[k] != 0.0f)
                atomicAdd(dst + k, v[k]);
        if (v[5] != 0.0f)
            atomicAdd(dst + 8, v[5]);
    }
}

// One workgroup: segStart = exclusive scan of segCount over the tiles (tiles + 1 entries),
// *totalItems, the owning tile of every work item (itemTile[segStart[t] + k] = t for
// k < segCount[t]) and the work counter cleared.
constexpr int kPrepThreads = 1024;
// Workgroup 1 (when launched) finishes the photometric loss of the same step instead.
__global__ __launch_bounds__(kPrepThreads) void segmentPrepKernel(SegmentPrep prep,
                                                                  LossFinalize fin)
{
    if (blockIdx.x == 1) {
        lossFinalizeBlock(fin);
        return;
    }
    segmentPrepBlock(prep);
}

// Segmented backward of one image (16x16 tiles): workgroups take (tile, segment) items from
// an atomic counter until *totalItems (itemTile maps an item to its tile, segStart is the
// exclusive scan of the per-tile segment counts written by the forward). An item is claimed
// only when the workgroup is free, so no item waits behind a busy workgroup.
template <bool kFastExp>
__global__ __launch_bounds__(kThreads) void rasterSegmentedKernel(
    int W, int H, int tilesX, const int *__restrict__ itemTile,
    const uint32_t *__restrict__ segStart, const uint32_t *__restrict__ totalItems,
    unsigned *counter, const int *__restrict__ tileOffsets, const int *__restrict__ tileList,
    const float4 *__restrict__ rec0, const float4 *__restrict__ rec1,
    const float *__restrict__ rec2, float bg0, float bg1, float bg2, float minAlpha,
    const float *__restrict__ dl, const float *__restrict__ finalTArr,
    const int *__restrict__ nContribArr, const float4 *__restrict__ ckpt,
    float *__restrict__ grad2d, LossFinalize fin, const uint64_t *__restrict__ hitBits)
{
    __shared__ uint32_t sWork;
    if (fin.partials && blockIdx.x == 0) // the step's loss value, before taking work items
        lossFinalizeBlock(fin);
    const uint32_t total = *totalItems;
    // The first item of each workgroup is its own index; later ones come from the counter (so
    // the launch does not start with every workgroup queueing on one atomic).
    for (uint32_t round = 0;; ++round) {
        if (round > 0) {
            if (threadIdx.x == 0)
                sWork = atomicAdd(counter, 1u) + gridDim.x;
            __syncthreads();
        }
        const uint32_t w = round > 0 ? sWork : blockIdx.x;
        __syncthreads();
        if (w >= total)
            break;
        const int tile = itemTile[w];
        backwardSegment<kSegEntries, kFastExp>(tile % tilesX, tile / tilesX, 0, int(w - segStart[tile]),
                                               W, H, tilesX, tileOffsets, tileList, rec0, rec1, rec2,
                                               bg0, bg1, bg2, minAlpha, dl, finalTArr, nContribArr,
                                               ckpt, grad2d, hitBits);
    }
}

template <bool kFastExp>
__global__ __launch_bounds__(kThreads) void rasterBackwardKernel(
    int W, int H, int ts, int tilesX, int blocksX, const int *__restrict__ tileOffsets,
    const int *__restrict__ tileList, const float4 *__restrict__ rec0,
    const float4 *__restrict__ rec1, const float *__restrict__ rec2, float bg0, float bg1,
    float bg2, float minAlpha, const float *__restrict__ dl, const float *__restrict__ finalTArr,
    const int *__restrict__ nContribArr, float *__restrict__ grad2d)
{
    backwardBlock<256, kPlainFewLanes, kFastExp>(blockIdx.x % blocksX, blockIdx.x / blocksX, 0, W, H, ts, tilesX, tileOffsets,
                  tileList, rec0, rec1, rec2, bg0, bg1, bg2, minAlpha, dl, finalTArr, nContribArr,
                  grad2d);
}

// Persistent batched compositing: every workgroup takes the next block of the work order from
// an atomic counter until the list is exhausted. The order lists heavy tiles first.
// kCkpt (forward: checkpoints for a segmented backward) and kFastExp are template parameters
// chosen on the host: each instantiation holds one block function, so its LDS arrays are those
// of one variant (runtime branches between variants had added up the variants' static LDS,
// 42 KB for the forward). At most three wavefronts per SIMD: with the register budget that
// leaves (up to 168), the compiler keeps the walks' independent chains apart; aiming at the
// five wavefronts the smaller LDS would allow (88-94 registers) measured 5% (backward) and 17%
// (forward) slower, and the whole step 2-3% slower with four wavefronts, 24% with two (8 views
// of 640x480, 50k Gaussians).
template <bool kBackward, int kBatch, int kFewLanes, bool kCkpt, bool kFastExp>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1, 3))) void rasterBatchedKernel(
    const RenderItem *__restrict__ items, int numItems, const uint32_t *__restrict__ order,
    int64_t totalBlocks, unsigned *counter, const int *__restrict__ tileOffsets,
    const int *__restrict__ tileList, const float4 *__restrict__ rec0,
    const float4 *__restrict__ rec1, const float *__restrict__ rec2,
    const u3d_raster_settings st, float *__restrict__ grad2d, SegmentState seg)
{
    __shared__ int64_t sWork;
    for (;;) {
        if (threadIdx.x == 0)
            sWork = int64_t(atomicAdd(counter, 1u));
        __syncthreads();
        const int64_t w = sWork;
        __syncthreads();
        if (w >= totalBlocks)
            break;
        const int64_t b = order ? int64_t(order[w]) : w;
        const int v = itemOfBlock(items, numItems, b);
        const RenderItem &it = items[v];
        const int local = int(b - it.blockBase);
        const int bx = local % it.blocksX, by = local / it.blocksX;
        const int *offs = tileOffsets + it.offsetShift;
#define U3D_FWD(CK, FE)                                                                          \
    forwardBlock<kBatch, CK, FE>(bx, by, it.tileBase, it.cam.width, it.cam.height, st.tile_size,   \
                                 it.tilesX, offs, tileList, rec0, rec1, rec2,                      \
                                 st.background[0], st.background[1], st.background[2],             \
                                 st.min_alpha, st.min_transmittance, it.rgb, it.finalT,            \
                                 it.nContrib, seg)
#define U3D_BWD(FE)                                                                              \
    backwardBlock<kBatch, kFewLanes, FE>(bx, by, it.tileBase, it.cam.width, it.cam.height,         \
                                         st.tile_size, it.tilesX, offs, tileList, rec0,            \
                                         rec1, rec2, st.background[0], st.background[1],           \
                                         st.background[2], st.min_alpha, it.dl, it.finalT,         \
                                         it.nContrib, grad2d)
        if constexpr (kBackward)
            U3D_BWD(kFastExp);
        else
            U3D_FWD(kCkpt, kFastExp);
#undef U3D_BWD
#undef U3D_FWD
    }
}

} // namespace

template <bool kBackward, int kBatch, int kFewLanes>
static void launchBatchedVariant(hipStream_t s, int numWorkgroups, const RenderItem *items,
                                 int numItems, const uint32_t *order, int64_t totalBlocks,
                                 unsigned *counter, const int *tileOffsets, const int *tileList,
                                 const ScreenBuffers &sb, const u3d_raster_settings &st,
                                 float *grad2d, const SegmentState &seg)
{
#define U3D_VARIANT(CK, FE)                                                                      \
    hipLaunchKernelGGL((rasterBatchedKernel<kBackward, kBatch, kFewLanes, CK, FE>),               \
                       dim3(numWorkgroups), dim3(kThreads), 0, s, items, numItems, order,         \
                       totalBlocks, counter, tileOffsets, tileList, sb.rec0, sb.rec1, sb.rec2, st, \
                       grad2d, seg)
    const bool fast = st.min_alpha >= 1e-37f; // noSpecialExp, on the host
    if constexpr (kBackward) {
        if (fast)
            U3D_VARIANT(false, true);
        else
            U3D_VARIANT(false, false);
    } else if (!seg.ckpt) {
        if (fast)
            U3D_VARIANT(false, true);
        else
            U3D_VARIANT(false, false);
    } else {
        if (fast)
            U3D_VARIANT(true, true);
        else
            U3D_VARIANT(true, false);
    }
#undef U3D_VARIANT
}

int launchRasterBatched(hipStream_t s, bool backward, const RenderItem *items, int numItems,
                        const uint32_t *order, int64_t totalBlocks, unsigned *counter,
                        int numWorkgroups, const int *tileOffsets, const int *tileList,
                        const ScreenBuffers &sb, const u3d_raster_settings &st, float *grad2d,
                        int ldsBatch, int fewLanes, const SegmentState &seg)
{
    if (totalBlocks <= 0)
        return U3D_OK;
    U3D_CHECK(ldsBatch == 64 || ldsBatch == 128 || ldsBatch == 256, U3D_ERR_INVALID,
              "LDS batch must be 64, 128 or 256");
    U3D_CHECK(fewLanes == 0 || fewLanes == 4 || fewLanes == 16, U3D_ERR_INVALID,
              "atomic lane threshold must be 0, 4 or 16");
#define U3D_BATCHED(B, L, F)                                                                     \
    launchBatchedVariant<B, L, F>(s, numWorkgroups, items, numItems, order, totalBlocks, counter, \
                                  tileOffsets, tileList, sb, st, grad2d, seg)
#define U3D_BATCHED_L(B, F)                                                                      \
    do {                                                                                         \
        if (ldsBatch == 64)                                                                      \
            U3D_BATCHED(B, 64, F);                                                               \
        else if (ldsBatch == 128)                                                                \
            U3D_BATCHED(B, 128, F);                                                              \
        else                                                                                     \
            U3D_BATCHED(B, 256, F);                                                              \
    } while (0)
    if (!backward)
        U3D_BATCHED_L(false, 0);
    else if (fewLanes == 0)
        U3D_BATCHED_L(true, 0);
    else if (fewLanes == 4)
        U3D_BATCHED_L(true, 4);
    else
        U3D_BATCHED_L(true, 16);
#undef U3D_BATCHED_L
#undef U3D_BATCHED
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int rasterBatchedOccupancy(bool backward, int ldsBatch)
{
    int blocks = 0;
    hipError_t e = hipErrorInvalidValue;
    if (backward)
        e = ldsBatch == 64    ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, rasterBatchedKernel<true, 64, 0, false, true>, kThreads, 0)
            : ldsBatch == 128 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, rasterBatchedKernel<true, 128, 0, false, true>, kThreads, 0)
                              : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, rasterBatchedKernel<true, 256, 0, false, true>, kThreads, 0);
    else
        e = ldsBatch == 64    ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, rasterBatchedKernel<false, 64, 0, false, true>, kThreads, 0)
            : ldsBatch == 128 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, rasterBatchedKernel<false, 128, 0, false, true>, kThreads, 0)
                              : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, rasterBatchedKernel<false, 256, 0, false, true>, kThreads, 0);
    return (e == hipSuccess && blocks > 0) ? blocks : 1;
}

int launchRasterForward(hipStream_t s, const RasterGeometry &g, const u3d_raster_settings &st,
                        const int *tileOffsets, const int *tileList, const ScreenBuffers &sb,
                        float *rgb, float *finalT, int *nContrib, const SegmentState &seg,
                        const int *blockOrder, unsigned long long *evalCount, uint32_t *walkCost)
{
    U3D_CHECK(!seg.ckpt || g.tileSize == 16, U3D_ERR_INVALID, "segment checkpoints need 16-pixel tiles");
    const int blocksX = (g.width + kEdge - 1) / kEdge, blocksY = (g.height + kEdge - 1) / kEdge;
    if (blocksX * blocksY == 0)
        return U3D_OK;
    if (g.tileSize % kEdge != 0) {
        hipLaunchKernelGGL(rasterForwardAnyKernel, dim3(blocksX * blocksY), dim3(kThreads), 0, s,
                           g.width, g.height, g.tileSize, g.tilesX, blocksX, tileOffsets, tileList,
                           sb.rec0, sb.rec1, sb.rec2, st.background[0], st.background[1],
                           st.background[2], st.min_alpha, st.min_transmittance, rgb, finalT,
                           nContrib);
        U3D_LAUNCH_CHECK();
        return U3D_OK;
    }
    const bool fast = st.min_alpha >= 1e-37f; // as noSpecialExp
    if (g.tileSize == kEdge) {
        auto quad = seg.ckpt ? (fast ? rasterForwardQuadKernel<true, true> : rasterForwardQuadKernel<true, false>)
                             : (fast ? rasterForwardQuadKernel<false, true> : rasterForwardQuadKernel<false, false>);
        const int numTiles = blocksX * blocksY;
        hipLaunchKernelGGL(quad, dim3(32 * ((numTiles + 7) / 8)), dim3(kThreads), 0, s, g.width,
                           g.height, g.tilesX, numTiles, tileOffsets, tileList, sb.rec0, sb.rec1, sb.rec2,
                           st.background[0], st.background[1], st.background[2], st.min_alpha,
                           st.min_transmittance, rgb, finalT, nContrib, seg, blockOrder, evalCount,
                           walkCost);
        U3D_LAUNCH_CHECK();
        return U3D_OK;
    }
    auto kernel = seg.ckpt ? (fast ? rasterForwardKernel<true, true> : rasterForwardKernel<true, false>)
                           : (fast ? rasterForwardKernel<false, true> : rasterForwardKernel<false, false>);
    hipLaunchKernelGGL(kernel, dim3(blocksX * blocksY), dim3(kThreads), 0, s, g.width, g.height,
                       g.tileSize, g.tilesX, blocksX, tileOffsets, tileList, sb.rec0, sb.rec1,
                       sb.rec2, st.background[0], st.background[1], st.background[2],
                       st.min_alpha, st.min_transmittance, rgb, finalT, nContrib, seg,
                       g.tileSize == kEdge ? blockOrder : nullptr);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchRasterBackward(hipStream_t s, const RasterGeometry &g, const u3d_raster_settings &st,
                         const int *tileOffsets, const int *tileList, const ScreenBuffers &sb,
                         const float *dlDrgb, const float *finalT, const int *nContrib,
                         float *grad2d)
{
    const int blocksX = (g.width + kEdge - 1) / kEdge, blocksY = (g.height + kEdge - 1) / kEdge;
    if (blocksX * blocksY == 0)
        return U3D_OK;
    if (g.tileSize % kEdge != 0) {
        hipLaunchKernelGGL(rasterBackwardAnyKernel, dim3(blocksX * blocksY), dim3(kThreads), 0, s,
                           g.width, g.height, g.tileSize, g.tilesX, blocksX, tileOffsets, tileList,
                           sb.rec0, sb.rec1, sb.rec2, st.background[0], st.background[1],
                           st.background[2], st.min_alpha, dlDrgb, finalT, nContrib, grad2d);
        U3D_LAUNCH_CHECK();
        return U3D_OK;
    }
    auto kernel = st.min_alpha >= 1e-37f ? rasterBackwardKernel<true> : rasterBackwardKernel<false>;
    hipLaunchKernelGGL(kernel, dim3(blocksX * blocksY), dim3(kThreads), 0, s,
                       g.width, g.height, g.tileSize, g.tilesX, blocksX, tileOffsets, tileList,
                       sb.rec0, sb.rec1, sb.rec2, st.background[0], st.background[1],
                       st.background[2], st.min_alpha, dlDrgb, finalT, nContrib, grad2d);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchRasterSegmented(hipStream_t s, const RasterGeometry &g, const u3d_raster_settings &st,
                          const uint32_t *segCount, uint32_t *segStart, uint32_t *totalItems,
                          unsigned *counter, int *itemTile, int numWorkgroups,
                          const int *tileOffsets, const int *tileList, const ScreenBuffers &sb,
                          const float *dl, const float *finalT, const int *nContrib,
                          const float4 *ckpt, float *grad2d, const LossFinalize *fin,
                          bool prepDone, const uint64_t *hitBits)
{
    const int tiles = g.tilesX * g.tilesY;
    if (tiles == 0)
        return fin ? launchLossFinalize(s, *fin) : U3D_OK;
    U3D_CHECK(g.tileSize == 16, U3D_ERR_INVALID, "segmented backward needs 16-pixel tiles");
    LossFinalize inMain; // the loss finalize rides on whichever launch comes first
    if (!prepDone) {
        SegmentPrep prep;
        prep.segCount = segCount;
        prep.perTile = kSegCountsPerTile;
        prep.tiles = tiles;
        prep.segStart = segStart;
        prep.totalItems = totalItems;
        prep.itemTile = itemTile;
        prep.counter = counter;
        hipLaunchKernelGGL(segmentPrepKernel, dim3(fin ? 2 : 1), dim3(kPrepThreads), 0, s, prep,
                           fin ? *fin : LossFinalize{});
        U3D_LAUNCH_CHECK();
    } else if (fin) {
        inMain = *fin;
    }
    auto kernel = st.min_alpha >= 1e-37f ? rasterSegmentedKernel<true> : rasterSegmentedKernel<false>;
    hipLaunchKernelGGL(kernel, dim3(numWorkgroups), dim3(kThreads), 0, s, g.width,
                       g.height, g.tilesX, itemTile, segStart, totalItems, counter, tileOffsets,
                       tileList, sb.rec0, sb.rec1, sb.rec2, st.background[0], st.background[1],
                       st.background[2], st.min_alpha, dl, finalT, nContrib, ckpt, grad2d, inMain,
                       hitBits);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int rasterSegmentedOccupancy()
{
    int blocks = 0;
    const hipError_t e =
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, rasterSegmentedKernel<true>, kThreads, 0);
    return (e == hipSuccess && blocks > 0) ? blocks : 1;
}

} // namespace u3d
