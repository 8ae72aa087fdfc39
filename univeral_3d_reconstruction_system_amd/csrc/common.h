// Shared internals of libu3d_hip: error handling, device buffers, the context, and the
// exact-arithmetic device helpers used by every kernel that must agree bit for bit with the
// reference trainer (engine/src/splat).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/u3d_hip.h"
#include "status.h"
#include "u3d_expf.h"

namespace u3d {

// ------------------------------------------------------------------------------ errors
int failHip(hipError_t e, const char *what, const char *file, int line);

#define U3D_HIP(call)                                                       \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess)                                               \
            return ::u3d::failHip(e_, #call, __FILE__, __LINE__);           \
    } while (0)

#define U3D_LAUNCH_CHECK() U3D_HIP(hipGetLastError())

// --------------------------------------------------------------------- device buffers
// Grow-only device allocation owned by a library object. ensure() never shrinks.
template <typename T> struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    int ensure(size_t n)
    {
        if (n <= cap)
            return U3D_OK;
        size_t want = n + n / 4 + 64;
        T *q = nullptr;
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&q), want * sizeof(T));
        if (e != hipSuccess)
            return failHip(e, "hipMalloc", __FILE__, __LINE__);
        release();
        p = q;
        cap = want;
        return U3D_OK;
    }
    // Like ensure() but keeps the first `keep` elements (stream-ordered copy on `s`).
    int grow(size_t n, size_t keep, hipStream_t s)
    {
        if (n <= cap)
            return U3D_OK;
        size_t want = n + n / 4 + 64;
        T *q = nullptr;
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&q), want * sizeof(T));
        if (e != hipSuccess)
            return failHip(e, "hipMalloc", __FILE__, __LINE__);
        if (keep && p) {
            e = hipMemcpyAsync(q, p, keep * sizeof(T), hipMemcpyDeviceToDevice, s);
            if (e == hipSuccess)
                e = hipStreamSynchronize(s);
            if (e != hipSuccess) {
                (void)hipFree(q);
                return failHip(e, "hipMemcpyAsync", __FILE__, __LINE__);
            }
        }
        release();
        p = q;
        cap = want;
        return U3D_OK;
    }
};

// Pinned host staging for small read-backs.
struct PinnedScratch {
    void *p = nullptr;
    size_t cap = 0;
    ~PinnedScratch()
    {
        if (p)
            (void)hipHostFree(p);
    }
    int ensure(size_t bytes)
    {
        if (bytes <= cap)
            return U3D_OK;
        if (p)
            (void)hipHostFree(p);
        p = nullptr;
        hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
        if (e != hipSuccess)
            return failHip(e, "hipHostMalloc", __FILE__, __LINE__);
        cap = bytes;
        return U3D_OK;
    }
};

// Waits until the pinned word *p carries the 31-bit seq in bits 32-62 (stored by a device-side
// system-scope release) and returns its low half, and bit 63 in *flag (nullable). Polls without
// any event on the stream; every few thousand polls it asks the stream for an error, and fails
// if the stream has drained or faulted without the word arriving.
int waitPinnedWord(const uint64_t *p, uint32_t seq, hipStream_t s, uint32_t *value,
                   bool *flag = nullptr);
// Waits (the same way) until the pinned word *p no longer holds `pending`.
int waitPinnedU32(const uint32_t *p, uint32_t pending, hipStream_t s);
// The psnr word of a step's statistics before the device has stored them (a NaN payload the
// loss reduction never produces).
constexpr uint32_t kStatsPending = 0x7fc0dead;

// Per-stage GPU time accumulated from HIP events recorded on the context's stream.
struct StageTimer {
    bool enabled = false;
    struct Pending {
        int stage;
        hipEvent_t start, end;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    double totalMs[U3D_NUM_STAGES] = {};
    int64_t count[U3D_NUM_STAGES] = {};
    ~StageTimer()
    {
        for (auto &p : pending) {
            (void)hipEventDestroy(p.start);
            (void)hipEventDestroy(p.end);
        }
        for (hipEvent_t e : pool)
            (void)hipEventDestroy(e);
    }
    hipEvent_t take()
    {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    // Resolves every recorded pair (waits for the events).
    void collect()
    {
        for (auto &p : pending) {
            float ms = 0.0f;
            if (hipEventSynchronize(p.end) == hipSuccess && hipEventElapsedTime(&ms, p.start, p.end) == hipSuccess) {
                totalMs[p.stage] += ms;
                count[p.stage] += 1;
            }
            pool.push_back(p.start);
            pool.push_back(p.end);
        }
        pending.clear();
    }
};

} // namespace u3d

struct u3d_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    u3d::PinnedScratch pinned;
    u3d::DevBuf<unsigned char> scratch; // shared short-lived device scratch (loss partials, ...)
    u3d::StageTimer timer;
    bool segmentedBackward = true; // U3D_OPT_SEGMENTED_BACKWARD
    int sortSpinLimit = 0;         // U3D_OPT_SORT_SPIN_LIMIT (0: the default bound)
    // this device's copy of the library's static device scratch (u3d_photometric_loss_noalloc)
    unsigned char *staticScratch = nullptr;
};

namespace u3d {

// Makes ctx's device current for the calling thread.
int useDevice(u3d_ctx *ctx);

// Records HIP events around one pipeline stage on the context's stream when timing is on.
class StageScope
{
public:
    StageScope(u3d_ctx *ctx, int stage) : m_ctx(ctx), m_stage(stage)
    {
        if (ctx->timer.enabled) {
            m_start = ctx->timer.take();
            (void)hipEventRecord(m_start, ctx->stream);
        }
    }
    ~StageScope()
    {
        if (m_start) {
            hipEvent_t end = m_ctx->timer.take();
            (void)hipEventRecord(end, m_ctx->stream);
            m_ctx->timer.pending.push_back({m_stage, m_start, end});
            if (m_ctx->timer.pending.size() > 4096)
                m_ctx->timer.collect();
        }
    }
    StageScope(const StageScope &) = delete;
    StageScope &operator=(const StageScope &) = delete;

private:
    u3d_ctx *m_ctx;
    int m_stage;
    hipEvent_t m_start = nullptr;
};

// ----------------------------------------------------------------- device-side helpers
constexpr float kShC0 = 0.28209479177387814f; // GaussianModel.h:10
constexpr float kMaxAlpha = 0.99f;            // Rasterizer.cpp:13
constexpr int kGrad2dStride = 9;              // Rasterizer.cpp:14

// float -> int32 with x86 cvttss2si semantics (the reference's int(float) on its platform):
// truncation toward zero; NaN and out-of-range values give INT32_MIN.
__host__ __device__ inline int cvtTruncI32(float f)
{
    if (!(f > -2147483904.0f && f < 2147483648.0f))
        return INT32_MIN;
    return int(f);
}

// One copy per translation unit (no relocatable device code).
static __constant__ uint64_t kExp2Table[32] = U3D_EXP2_TABLE;

// Host-identical expf reading the 2^(i/32) table from constant memory.
__device__ inline float expfExact(float x)
{
    bool special;
    const float s = u3dmath::expf_special(x, &special);
    if (special)
        return s;
    return u3dmath::expf_core(x, kExp2Table);
}

// Same, with the table staged in LDS by the caller (32 entries).
__device__ inline float expfExactLds(float x, const uint64_t *ldsTable)
{
    bool special;
    const float s = u3dmath::expf_special(x, &special);
    if (special)
        return s;
    return u3dmath::expf_core(x, ldsTable);
}

// Same, for callers that know (wave-uniformly, noSpecial) that x is NaN or lies in (-87, 0]:
// the core alone then gives the library result (NaN in, NaN out).
__device__ inline float expfExactLdsIf(float x, const uint64_t *ldsTable, bool noSpecial)
{
    if (!noSpecial) {
        bool special;
        const float s = u3dmath::expf_special(x, &special);
        if (special)
            return s;
    }
    return u3dmath::expf_core(x, ldsTable);
}

__device__ inline uint32_t floatBits(float f) { return __float_as_uint(f); }

// Orders LDS accesses of one wavefront: data one lane wrote is read by another lane after it.
__device__ __forceinline__ void waveSync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

} // namespace u3d
