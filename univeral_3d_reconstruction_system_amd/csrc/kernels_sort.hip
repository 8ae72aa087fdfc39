// Tile binning on the device (replaces the serial Rasterizer::binTiles, Rasterizer.cpp:206-261).
//
// The reference orders every tile list by (depth bits, Gaussian index): a stable radix sort of
// the visible Gaussians' depth bits followed by a stable per-tile scatter. Here the same order
// comes from an LSD radix sort on the composite (tile, depth) key, with the depth digits sorted
// before the (tile, Gaussian) pairs are duplicated, so those passes cost N elements, not K:
//   1. stable radix sort of (depth bits, index) over all N Gaussians (invisible keys 0xffffffff
//      sort last, ties keep index order);
//   2. exclusive scan of the tiles each Gaussian covers, in depth order -> pair slots, K;
//   3. emit (tile id, Gaussian) pairs in depth order;
//   4. stable radix sort of the pairs on the tile-id bits;
//   5. tile offsets from the sorted tile ids.
// All integer work: HBM-bound. Each radix pass is three launches: an upsweep (per-block digit
// histogram, digit-major), a row scan (one workgroup per digit scans that digit's counts over
// the blocks and records the digit total), and a downsweep that adds the exclusive scan of the
// 256 digit totals, ranks keys stably inside each wavefront with ballot-based digit matching and
// scatters them.
//
// The plain (single-view) path replaces steps 2-5 with a direct tile binning of the depth-ordered
// Gaussians (tileCountKernel / tileScanKernel / tileEmitKernel below): the pairs are generated
// twice from the rectangles, once to count them per (group of Gaussians, tile) and once to write
// each Gaussian id at its final slot, so the K pairs are written once (4 B each) and never sorted.
#include "kernels.h"

#include <utility>

namespace u3d {

namespace {

constexpr int kRadixThreads = 256;
constexpr int kWaves = kRadixThreads / 64;
constexpr int kMinItems = 4;                                // keys per lane per block (smallest)
constexpr int kScanThreads = 1024;

// Keys per lane: enough blocks to fill the chip for small arrays, more work per block for
// large ones (amortises the 256-entry histogram each block writes).
inline int itemsFor(int64_t count) { return count <= (int64_t(1) << 20) ? 4 : 8; }

// Element count: the host value, or min(*countDev, host value) when the count is on the device
// (the host value is then the capacity the grid was sized for).
__device__ __forceinline__ int64_t deviceCount(int64_t count, const uint32_t *countDev)
{
    if (!countDev)
        return count;
    const int64_t c = int64_t(*countDev);
    return c < count ? c : count;
}

__device__ __forceinline__ uint64_t laneMaskLt()
{
    const unsigned lane = __lane_id();
    return (uint64_t(1) << lane) - 1;
}

// Mask of the active lanes whose 8-bit digit equals this lane's (restricted to `valid`).
__device__ __forceinline__ uint64_t matchDigit(uint32_t digit, uint64_t valid, int bits)
{
    uint64_t m = valid;
    for (int b = 0; b < bits; ++b) {
        const uint64_t bal = __ballot((digit >> b) & 1u);
        m &= ((digit >> b) & 1u) ? bal : ~bal;
    }
    return m;
}

template <int kItems>
__global__ __launch_bounds__(kRadixThreads) void radixUpsweep(const uint32_t *__restrict__ keys,
                                                              int64_t count, int shift, int bits,
                                                              int numBlocks, uint32_t *hist,
                                                              const uint32_t *__restrict__ countDev)
{
    __shared__ uint32_t h[256];
    count = deviceCount(count, countDev);
    const int tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    const int wave = tid >> 6, lane = tid & 63;
    const uint32_t mask = (1u << bits) - 1u;
    const int64_t base = int64_t(blockIdx.x) * kRadixThreads * kItems + int64_t(wave) * 64 * kItems;
    for (int it = 0; it < kItems; ++it) {
        const int64_t idx = base + int64_t(it) * 64 + lane;
        const bool valid = idx < count;
        const uint32_t d = valid ? (keys[idx] >> shift) & mask : 0u;
        const uint64_t vm = __ballot(valid);
        if (vm == 0)
            break;
        const uint64_t m = matchDigit(d, vm, bits);
        if (valid && (m & laneMaskLt()) == 0) // lowest lane of its digit group
            atomicAdd(&h[d], uint32_t(__popcll(m)));
    }
    __syncthreads();
    hist[size_t(tid) * size_t(numBlocks) + blockIdx.x] = h[tid];
}

// Exclusive scan of `len` u32 in place with one block, 4 values per thread per chunk.
__global__ __launch_bounds__(kScanThreads) void scanSingleBlock(uint32_t *data, int64_t len,
                                                                 uint32_t *total)
{
    __shared__ uint32_t waveSums[kScanThreads / 64];
    __shared__ uint32_t carry;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0)
        carry = 0;
    __syncthreads();
    constexpr int kPer = 4;
    for (int64_t start = 0; start < len; start += int64_t(kScanThreads) * kPer) {
        uint32_t v[kPer];
        uint32_t local = 0;
        const int64_t myBase = start + int64_t(tid) * kPer;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            v[k] = (myBase + k < len) ? data[myBase + k] : 0u;
            local += v[k];
        }
        // inclusive wave scan of `local`
        uint32_t x = local;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= off)
                x += y;
        }
        if (lane == 63)
            waveSums[wave] = x;
        __syncthreads();
        if (wave == 0) {
            uint32_t s = lane < kScanThreads / 64 ? waveSums[lane] : 0u;
#pragma unroll
            for (int off = 1; off < kScanThreads / 64; off <<= 1) {
                const uint32_t y = __shfl_up(s, off, 64);
                if (lane >= off)
                    s += y;
            }
            if (lane < kScanThreads / 64)
                waveSums[lane] = s; // inclusive
        }
        __syncthreads();
        const uint32_t waveOff = wave > 0 ? waveSums[wave - 1] : 0u;
        uint32_t run = carry + waveOff + (x - local);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            if (myBase + k < len)
                data[myBase + k] = run;
            run += v[k];
        }
        __syncthreads();
        if (tid == 0)
            carry += waveSums[kScanThreads / 64 - 1];
        __syncthreads();
    }
    if (tid == 0 && total)
        *total = carry;
}

// Exclusive scan in place of one digit's row hist[d][0, numBlocks) (one workgroup per digit);
// the row total goes to digitTotal[d].
__global__ __launch_bounds__(kRadixThreads) void radixRowScan(uint32_t *hist, int numBlocks,
                                                              uint32_t *digitTotal)
{
    __shared__ uint32_t waveSums[kWaves];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t *row = hist + size_t(blockIdx.x) * size_t(numBlocks);
    constexpr int kPer = 4;
    uint32_t carry = 0;
    for (int start = 0; start < numBlocks; start += kRadixThreads * kPer) {
        const int base = start + tid * kPer;
        uint32_t v[kPer], local = 0;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            v[k] = base + k < numBlocks ? row[base + k] : 0u;
            local += v[k];
        }
        uint32_t x = local;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= off)
                x += y;
        }
        if (lane == 63)
            waveSums[wave] = x;
        __syncthreads();
        uint32_t waveOff = 0, chunk = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t ws = waveSums[w];
            waveOff += w < wave ? ws : 0u;
            chunk += ws;
        }
        uint32_t run = carry + waveOff + (x - local);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            if (base + k < numBlocks)
                row[base + k] = run;
            run += v[k];
        }
        carry += chunk;
        __syncthreads();
    }
    if (tid == 0)
        digitTotal[blockIdx.x] = carry;
}

// kIota: the values are the input positions (valsIn is not read).
template <int kItems, bool kIota>
__global__ __launch_bounds__(kRadixThreads) void radixDownsweep(
    const uint32_t *__restrict__ keysIn, const uint32_t *__restrict__ valsIn, uint32_t *keysOut,
    uint32_t *valsOut, int64_t count, int shift, int bits, int numBlocks,
    const uint32_t *__restrict__ hist, const uint32_t *__restrict__ digitTotal,
    const uint32_t *__restrict__ countDev)
{
    count = deviceCount(count, countDev);
    __shared__ uint32_t waveCount[kWaves][256];
    __shared__ uint32_t gbase[256];
    __shared__ uint32_t waveTot[kWaves];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
#pragma unroll
    for (int w = 0; w < kWaves; ++w)
        waveCount[w][tid] = 0;
    {
        // digit base = exclusive scan of the digit totals + this block's row prefix
        const uint32_t tot = digitTotal[tid];
        uint32_t x = tot;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= off)
                x += y;
        }
        if (lane == 63)
            waveTot[wave] = x;
        __syncthreads();
        uint32_t off = x - tot;
#pragma unroll
        for (int w = 0; w < kWaves; ++w)
            off += w < wave ? waveTot[w] : 0u;
        gbase[tid] = off + hist[size_t(tid) * size_t(numBlocks) + blockIdx.x];
    }
    __syncthreads();
    const uint32_t mask = (1u << bits) - 1u;
    const int64_t base = int64_t(blockIdx.x) * kRadixThreads * kItems + int64_t(wave) * 64 * kItems;
    uint32_t k[kItems], v[kItems], rank[kItems];
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const int64_t idx = base + int64_t(it) * 64 + lane;
        const bool valid = idx < count;
        k[it] = valid ? keysIn[idx] : 0u;
        v[it] = valid ? (kIota ? uint32_t(idx) : valsIn[idx]) : 0u;
        const uint32_t d = (k[it] >> shift) & mask;
        const uint64_t vm = __ballot(valid);
        rank[it] = 0;
        if (vm == 0)
            continue;
        const uint64_t m = matchDigit(d, vm, bits);
        if (valid) {
            const uint32_t prev = waveCount[wave][d];
            rank[it] = prev + uint32_t(__popcll(m & laneMaskLt()));
            if ((m & laneMaskLt()) == 0)
                waveCount[wave][d] = prev + uint32_t(__popcll(m));
        }
    }
    __syncthreads();
    {
        // exclusive prefix over waves, per digit
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = waveCount[w][tid];
            waveCount[w][tid] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const int64_t idx = base + int64_t(it) * 64 + lane;
        if (idx < count) {
            const uint32_t d = (k[it] >> shift) & mask;
            const uint32_t pos = gbase[d] + waveCount[wave][d] + rank[it];
            keysOut[pos] = k[it];
            valsOut[pos] = v[it];
        }
    }
}

// Sum on one workgroup (up to 1024 threads). hostOut (optional): one 64-bit word (seq << 32 | sum)
// stored with a system-scope release; seq has 31 bits, and bit 63 of the word is set when
// errWord (optional: the small sort's error word, set by an earlier call whose lookback wait
// gave up) was non-zero. The word is cleared here, so each abandoned wait is reported once.
__device__ __forceinline__ void sumBlock(const uint32_t *__restrict__ in, int64_t count,
                                         uint32_t *devOut, uint64_t *hostOut, uint32_t seq,
                                         uint32_t *errWord = nullptr)
{
    __shared__ uint32_t ws[16];
    const int tid = threadIdx.x, nt = int(blockDim.x);
    uint32_t v = 0;
    for (int64_t i = tid; i < count; i += nt)
        v += in[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    if ((tid & 63) == 0)
        ws[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
        uint32_t t = 0;
        for (int w = 0; w < nt / 64; ++w)
            t += ws[w];
        *devOut = t;
        uint64_t err = 0;
        if (errWord)
            err = __hip_atomic_exchange(errWord, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u
                      ? uint64_t(1) << 63
                      : 0u;
        if (hostOut)
            __hip_atomic_store(hostOut, err | (uint64_t(seq & 0x7fffffffu) << 32) | uint64_t(t),
                               __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// ------------------------------------------------------------------ small-sort path
// Up to kOsMaxBlocks * 4096 keys the sort is bound by launch and memory latency, not by bytes,
// so it runs as one histogram launch for every digit followed by one launch per 8-bit pass
// (instead of three per pass). A pass workgroup draws a ticket (= its block of 4096 keys in
// the current order), ranks its keys per wavefront, publishes its per-digit counts as 64-bit
// words tagged with the call's epoch and the pass, and adds up the counts of every block with a
// smaller ticket. Those blocks started earlier and wait for nobody before publishing, so every
// wait ends; a thread that has re-read spinLimit words in all (kOsSpinLimit unless the host
// lowers it) sets the error word and stops waiting, so a fault elsewhere cannot hang the
// stream. The tags make resetting the lookback
// words between calls unnecessary; the epoch comes from the host, one per call.
constexpr int kOsThreads = 1024;
constexpr int kOsWaves = kOsThreads / 64;
constexpr int kOsItems = 4;
constexpr int kOsBlockKeys = kOsThreads * kOsItems;
constexpr int kOsMaxBlocks = 256;
constexpr int kOsHistBlocks = 64;
constexpr int kOsSpinLimit = 1 << 20;
constexpr int kOsBatch = 24;
// Scratch (u64 words): the state (u32 tickets[4], u32 error), the lookback words
// [kOsMaxBlocks][256] and the histogram launch's per-block digit counts as u32
// [kOsHistBlocks][4][256].
constexpr size_t kOsStateWords = 8;
constexpr size_t kOsLookbackWords = size_t(kOsMaxBlocks) * 256;
constexpr size_t kOsPartialWords = size_t(kOsHistBlocks) * 4 * 256 / 2;
constexpr int kOsTicket = 0, kOsError = 4;

// Buffer resource word 3 for gfx9-family raw buffers; load cache policy sc0 | sc1 plus the
// compiler's volatile bit (every load is issued, none is merged or hoisted).
constexpr int kBufferRsrcWord3 = 0x00020000;
constexpr int kAuxCoherentVolatile = 1 | 16 | int(1u << 31);
typedef unsigned v2u32 __attribute__((ext_vector_type(2)));

// Adds 1 to h[d] for every active lane; one add when the whole wavefront shares the digit.
__device__ __forceinline__ void countDigit(uint32_t *h, uint32_t d, bool valid)
{
    const uint64_t vm = __ballot(valid);
    if (vm == 0)
        return;
    const uint32_t d0 = uint32_t(__shfl(int(d), __ffsll((long long)vm) - 1, 64));
    if (__ballot(valid && d == d0) == vm) {
        if ((vm & laneMaskLt() & __ballot(valid)) == 0 && valid)
            atomicAdd(&h[d0], uint32_t(__popcll(vm)));
    } else if (valid) {
        atomicAdd(&h[d], 1u);
    }
}

// Digit counts of every pass over the keys [blockIdx.x * per, (blockIdx.x + 1) * per); resets
// the pass tickets. Workgroup kOsHistBlocks (when launched) runs the side sum instead.
__global__ __launch_bounds__(kOsThreads) void onesweepHistKernel(
    const uint32_t *__restrict__ keys, int64_t count, int64_t per, int bits,
    uint32_t *__restrict__ partials, uint32_t *state, const uint32_t *__restrict__ sideIn,
    int64_t sideCount, uint32_t *sideDev, uint64_t *sideHost, uint32_t sideSeq)
{
    if (blockIdx.x == kOsHistBlocks) {
        // the passes of earlier calls have finished (stream order): report their error word
        sumBlock(sideIn, sideCount, sideDev, sideHost, sideSeq, state + kOsError);
        return;
    }
    __shared__ uint32_t h[4][256];
    const int tid = threadIdx.x;
    for (int k = tid; k < 4 * 256; k += kOsThreads)
        h[k >> 8][k & 255] = 0;
    if (blockIdx.x == 0 && tid < 4)
        state[kOsTicket + tid] = 0u;
    __syncthreads();
    const int passes = (bits + 7) >> 3;
    const int64_t lo = int64_t(blockIdx.x) * per;
    const int64_t hi = lo + per < count ? lo + per : count;
    for (int64_t i0 = lo; i0 < hi; i0 += int64_t(kOsThreads) * 4) {
        uint32_t key[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t i = i0 + int64_t(j) * kOsThreads + tid;
            key[j] = i < hi ? keys[i] : 0u;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool valid = i0 + int64_t(j) * kOsThreads + tid < hi;
            for (int p = 0; p < passes; ++p) {
                const int b = bits - 8 * p < 8 ? bits - 8 * p : 8;
                countDigit(h[p], (key[j] >> (8 * p)) & ((1u << b) - 1u), valid);
            }
        }
    }
    __syncthreads();
    for (int k = tid; k < passes * 256; k += kOsThreads)
        partials[size_t(blockIdx.x) * 1024 + size_t(k)] = h[k >> 8][k & 255];
}

// One stable counting pass on bits [shift, shift + bits) of the keys. kIota: the values are
// the input positions (valsIn is not read). The block's keys are first placed in LDS in digit
// order, so that each digit's keys leave as one contiguous run of addresses.
template <bool kIota>
__global__ __launch_bounds__(kOsThreads) void onesweepPassKernel(
    const uint32_t *__restrict__ keysIn, const uint32_t *__restrict__ valsIn,
    uint32_t *__restrict__ keysOut, uint32_t *__restrict__ valsOut, int64_t count, int shift,
    int bits, int pass, uint32_t epoch, const uint32_t *__restrict__ partials,
    uint64_t *lookback, uint32_t *state, int spinLimit)
{
    __shared__ uint32_t waveCount[kOsWaves][256];
    __shared__ uint32_t sKey[kOsBlockKeys], sVal[kOsBlockKeys];
    __shared__ uint32_t sTot[4][256];
    __shared__ uint32_t sPre[4][256];
    __shared__ uint32_t blockStart[256];
    __shared__ uint32_t waveTotA[4], waveTotT[4];
    __shared__ int sBlock;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    // a wavefront's 64 threads share their quarter: uniform, so per-block addresses are scalar
    const int d = tid & 255, quarter = __builtin_amdgcn_readfirstlane(tid >> 8);
    if (tid == 0)
        sBlock = int(atomicAdd(&state[kOsTicket + pass], 1u));
    for (int k = tid; k < kOsWaves * 256; k += kOsThreads)
        waveCount[k >> 8][k & 255] = 0;
    {
        // this digit's total over the whole input (a quarter of the histogram blocks each)
        uint32_t part[kOsHistBlocks / 4];
#pragma unroll
        for (int j = 0; j < kOsHistBlocks / 4; ++j)
            part[j] = partials[size_t(quarter + 4 * j) * 1024 + size_t(pass) * 256 + size_t(d)];
        uint32_t tot = 0;
#pragma unroll
        for (int j = 0; j < kOsHistBlocks / 4; ++j)
            tot += part[j];
        sTot[quarter][d] = tot;
    }
    __syncthreads();
    const int blk = __builtin_amdgcn_readfirstlane(sBlock);
    const uint32_t tag = (epoch << 2) | uint32_t(pass);
    const uint32_t mask = (1u << bits) - 1u;
    const int64_t blockBase = int64_t(blk) * kOsBlockKeys;
    const int blockKeys = int(count - blockBase < kOsBlockKeys ? count - blockBase : kOsBlockKeys);
    const int waveBase = wave * 64 * kOsItems;
    uint32_t k[kOsItems], v[kOsItems], rank[kOsItems];
#pragma unroll
    for (int it = 0; it < kOsItems; ++it) {
        const int li = waveBase + it * 64 + lane;
        const bool valid = li < blockKeys;
        k[it] = valid ? keysIn[blockBase + li] : 0u;
        v[it] = valid ? (kIota ? uint32_t(blockBase + li) : valsIn[blockBase + li]) : 0u;
    }
#pragma unroll
    for (int it = 0; it < kOsItems; ++it) {
        const bool valid = waveBase + it * 64 + lane < blockKeys;
        const uint32_t dg = (k[it] >> shift) & mask;
        const uint64_t vm = __ballot(valid);
        rank[it] = 0;
        if (vm == 0)
            continue;
        const uint64_t m = matchDigit(dg, vm, bits);
        if (valid) {
            const uint32_t prev = waveCount[wave][dg];
            rank[it] = prev + uint32_t(__popcll(m & laneMaskLt()));
            if ((m & laneMaskLt()) == 0)
                waveCount[wave][dg] = prev + uint32_t(__popcll(m));
        }
    }
    __syncthreads();
    // Per digit: exclusive prefix over the wavefronts, the block's count (published at once),
    // and inclusive wavefront scans over the digits of the block counts and of the totals.
    uint32_t agg = 0, xa = 0, tot = 0, xt = 0;
    if (tid < 256) {
#pragma unroll
        for (int w = 0; w < kOsWaves; ++w) {
            const uint32_t c = waveCount[w][tid];
            waveCount[w][tid] = agg;
            agg += c;
        }
        __hip_atomic_store(&lookback[size_t(blk) * 256 + size_t(tid)],
                           (uint64_t(tag) << 32) | uint64_t(agg), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        tot = sTot[0][tid] + sTot[1][tid] + sTot[2][tid] + sTot[3][tid];
        xa = agg;
        xt = tot;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t ya = __shfl_up(xa, off, 64);
            const uint32_t yt = __shfl_up(xt, off, 64);
            if (lane >= off) {
                xa += ya;
                xt += yt;
            }
        }
        if (lane == 63) {
            waveTotA[wave] = xa;
            waveTotT[wave] = xt;
        }
    }
    __syncthreads();
    uint32_t globalStart = 0;
    if (tid < 256) {
        uint32_t sa = xa - agg, st = xt - tot;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            sa += w < wave ? waveTotA[w] : 0u;
            st += w < wave ? waveTotT[w] : 0u;
        }
        blockStart[tid] = sa;
        globalStart = st;
    }
    __syncthreads();
    // the block's keys in digit order (stable) in LDS
#pragma unroll
    for (int it = 0; it < kOsItems; ++it) {
        if (waveBase + it * 64 + lane < blockKeys) {
            const uint32_t dg = (k[it] >> shift) & mask;
            const uint32_t lp = blockStart[dg] + waveCount[wave][dg] + rank[it];
            sKey[lp] = k[it];
            sVal[lp] = v[it];
        }
    }
    // counts of this digit in the blocks with smaller tickets (a quarter of them each); all
    // pending words of a batch are read again together. Buffer loads: one lane offset for all
    // blocks, the block's row offset in a scalar register; volatile and coherent (sc0 sc1).
    const __amdgpu_buffer_rsrc_t rows = __builtin_amdgcn_make_buffer_rsrc(
        lookback, 0, int(kOsLookbackWords * sizeof(uint64_t)), kBufferRsrcWord3);
    auto readWord = [&](int b) {
        const v2u32 x = __builtin_amdgcn_raw_buffer_load_b64(rows, d * 8, b * 256 * 8,
                                                             kAuxCoherentVolatile);
        return (uint64_t(x.y) << 32) | uint64_t(x.x);
    };
    uint32_t pre = 0;
    int spins = 0; // rounds of re-reads, for the whole wait
    // spinLimit < 0 (set by the host to exercise the error path): every block with a wait gives
    // it up at once, whether or not its predecessors have published
    const bool forced = spinLimit < 0 && blk > 0;
    if (forced && tid == 0)
        __hip_atomic_store(&state[kOsError], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int b0 = quarter; b0 < (forced ? 0 : blk); b0 += 4 * kOsBatch) {
        uint64_t w[kOsBatch];
        uint32_t pending = 0;
#pragma unroll
        for (int j = 0; j < kOsBatch; ++j) {
            const int b = b0 + 4 * j;
            w[j] = b < blk ? readWord(b) : uint64_t(tag) << 32;
        }
#pragma unroll
        for (int j = 0; j < kOsBatch; ++j)
            pending |= uint32_t(w[j] >> 32) != tag ? 1u << j : 0u;
        while (pending) {
            if (++spins > spinLimit) {
                __hip_atomic_store(&state[kOsError], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
#pragma unroll
            for (int j = 0; j < kOsBatch; ++j)
                if (pending & (1u << j))
                    w[j] = readWord(b0 + 4 * j);
#pragma unroll
            for (int j = 0; j < kOsBatch; ++j)
                if (uint32_t(w[j] >> 32) == tag)
                    pending &= ~(1u << j);
        }
#pragma unroll
        for (int j = 0; j < kOsBatch; ++j)
            pre += uint32_t(w[j]);
    }
    sPre[quarter][d] = pre;
    __syncthreads();
    // LDS slot i of digit dg goes to digitBase[dg] + i (unsigned arithmetic wraps back)
    if (tid < 256)
        sTot[0][tid] = globalStart + sPre[0][tid] + sPre[1][tid] + sPre[2][tid] + sPre[3][tid] -
                       blockStart[tid];
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kOsItems; ++it) {
        const int i = it * kOsThreads + tid;
        if (i < blockKeys) {
            const uint32_t key = sKey[i];
            const uint32_t pos = sTot[0][(key >> shift) & mask] + uint32_t(i);
            // always in range when every wait ended normally; after an abandoned wait (error
            // word set) the order is wrong but no write leaves the arrays
            if (int64_t(pos) < count) {
                keysOut[pos] = key;
                valsOut[pos] = sVal[i];
            }
        }
    }
}

__global__ void iotaKernel(uint32_t *vals, int64_t count)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count)
        vals[i] = uint32_t(i);
}

// out[i] = in[idx[i]]; an index at or beyond inCount (only after a sort whose lookback wait
// gave up) reads as 0.
__global__ void gatherKernel(const uint32_t *__restrict__ in, const uint32_t *__restrict__ idx,
                             uint32_t *out, int64_t count, int64_t inCount)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) {
        const uint32_t j = idx[i];
        out[i] = int64_t(j) < inCount ? in[j] : 0u;
    }
}

// Block-level partial sums for the multi-block scan.
constexpr int kScanBlockItems = 4096; // 256 threads x 16

__global__ __launch_bounds__(256) void scanReduce(const uint32_t *__restrict__ in, int64_t count,
                                                  uint32_t *blockSums)
{
    __shared__ uint32_t ws[4];
    const int tid = threadIdx.x;
    const int64_t base = int64_t(blockIdx.x) * kScanBlockItems;
    uint32_t s = 0;
    for (int k = 0; k < 16; ++k) {
        const int64_t i = base + int64_t(k) * 256 + tid;
        if (i < count)
            s += in[i];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        s += __shfl_xor(s, off, 64);
    if ((tid & 63) == 0)
        ws[tid >> 6] = s;
    __syncthreads();
    if (tid == 0)
        blockSums[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

__global__ __launch_bounds__(256) void scanDownsweep(const uint32_t *__restrict__ in,
                                                     uint32_t *out, int64_t count,
                                                     const uint32_t *__restrict__ blockOffsets)
{
    __shared__ uint32_t ws[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // each thread owns 16 consecutive items
    const int64_t base = int64_t(blockIdx.x) * kScanBlockItems + int64_t(tid) * 16;
    uint32_t v[16];
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        v[k] = (base + k < count) ? in[base + k] : 0u;
        local += v[k];
    }
    uint32_t x = local;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off)
            x += y;
    }
    if (lane == 63)
        ws[wave] = x;
    __syncthreads();
    uint32_t waveOff = 0;
    for (int w = 0; w < wave; ++w)
        waveOff += ws[w];
    uint32_t run = blockOffsets[blockIdx.x] + waveOff + (x - local);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        if (base + k < count)
            out[base + k] = run;
        run += v[k];
    }
}

__global__ __launch_bounds__(1024) void sumKernel(const uint32_t *__restrict__ in, int64_t count,
                                                  uint32_t *devOut, uint64_t *hostOut, uint32_t seq)
{
    sumBlock(in, count, devOut, hostOut, seq);
}

// Scan downsweep that sums the previous blocks' partial sums itself (for up to
// kScanFusedBlocks blocks), so the scan needs no single-block pass; the last block writes the
// total.
constexpr int kScanFusedBlocks = 4096;

__global__ __launch_bounds__(256) void scanDownsweepFused(const uint32_t *__restrict__ in,
                                                          uint32_t *out, int64_t count,
                                                          const uint32_t *__restrict__ blockSums,
                                                          uint32_t *total)
{
    __shared__ uint32_t ws[4];
    __shared__ uint32_t blockOff;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    {
        uint32_t p = 0;
        for (int j = tid; j < int(blockIdx.x); j += 256)
            p += blockSums[j];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            p += __shfl_xor(p, off, 64);
        if (lane == 0)
            ws[wave] = p;
        __syncthreads();
        if (tid == 0) {
            blockOff = ws[0] + ws[1] + ws[2] + ws[3];
            if (total && blockIdx.x == gridDim.x - 1)
                *total = blockOff + blockSums[blockIdx.x];
        }
        __syncthreads();
    }
    const int64_t base = int64_t(blockIdx.x) * kScanBlockItems + int64_t(tid) * 16;
    uint32_t v[16];
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        v[k] = (base + k < count) ? in[base + k] : 0u;
        local += v[k];
    }
    uint32_t x = local;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off)
            x += y;
    }
    if (lane == 63)
        ws[wave] = x;
    __syncthreads();
    uint32_t waveOff = 0;
    for (int w = 0; w < wave; ++w)
        waveOff += ws[w];
    uint32_t run = blockOff + waveOff + (x - local);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        if (base + k < count)
            out[base + k] = run;
        run += v[k];
    }
}

// Pairs past `capacity` (only when the depth order is not a permutation, after a sort whose
// lookback wait gave up) are not written.
__global__ void emitPairsKernel(int64_t n, const uint32_t *__restrict__ sorted,
                                const uint32_t *__restrict__ offsets,
                                const int4 *__restrict__ rect, int tilesX, uint32_t *pairTile,
                                uint32_t *pairGauss, const uint32_t *__restrict__ nTilesSorted,
                                int64_t capacity)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || nTilesSorted[i] == 0)
        return;
    const uint32_t g = sorted[i];
    const int4 r = rect[g];
    int64_t o = offsets[i];
    for (int ty = r.y; ty < r.w; ++ty)
        for (int tx = r.x; tx < r.z; ++tx) {
            if (o < capacity) {
                pairTile[o] = uint32_t(ty * tilesX + tx);
                pairGauss[o] = g;
            }
            ++o;
        }
}

__global__ void emitPairsBatchedKernel(int64_t n, const uint32_t *__restrict__ sorted,
                                       const uint32_t *__restrict__ offsets,
                                       const int4 *__restrict__ rect, const int *__restrict__ entryItem,
                                       const RenderItem *__restrict__ items, uint32_t *pairTile,
                                       uint32_t *pairGauss, const uint32_t *__restrict__ nTilesSorted,
                                       int64_t capacity, unsigned *overflow)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t nt = nTilesSorted[i];
    if (nt == 0)
        return;
    const uint32_t e = sorted[i];
    const RenderItem &it = items[entryItem[e]];
    const int4 r = rect[e];
    int64_t o = int64_t(offsets[i]);
    if (o + int64_t(nt) > capacity)
        atomicOr(overflow, 1u);
    for (int ty = r.y; ty < r.w; ++ty)
        for (int tx = r.x; tx < r.z; ++tx) {
            if (o < capacity) {
                pairTile[o] = uint32_t(it.tileBase + ty * it.tilesX + tx);
                pairGauss[o] = e;
            }
            ++o;
        }
}

// key = inverted log-scale tile list length of the block's tile (8 bits), value = block id.
__global__ void blockCostKernel(const RenderItem *__restrict__ items, int numItems,
                                int64_t totalBlocks, int tileSize, const int *__restrict__ tileOffsets,
                                uint32_t *keys, uint32_t *order)
{
    const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= totalBlocks)
        return;
    int lo = 0, hi = numItems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].blockBase <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    const RenderItem &it = items[lo];
    const int local = int(b - it.blockBase);
    const int bx = local % it.blocksX, by = local / it.blocksX;
    const int tile = it.tileBase + it.offsetShift + (by * 16 / tileSize) * it.tilesX + bx * 16 / tileSize;
    const uint32_t len = uint32_t(tileOffsets[tile + 1] - tileOffsets[tile]);
    // 8-bit log-scale cost (exponent, 3 mantissa bits), inverted so that an ascending one-pass
    // radix sort lists the heaviest blocks first.
    uint32_t bucket = 0;
    if (len > 0) {
        const int lg = 31 - __clz(len);
        const uint32_t frac = lg >= 3 ? (len >> (lg - 3)) & 7u : (len << (3 - lg)) & 7u;
        bucket = uint32_t(lg) * 8u + frac;
    }
    keys[b] = 255u - bucket;
    order[b] = uint32_t(b);
}

__global__ void tileOffsetsKernel(const uint32_t *__restrict__ pairTile, int64_t numPairs,
                                  int tiles, int *tileOffsets, const uint32_t *__restrict__ countDev)
{
    numPairs = deviceCount(numPairs, countDev);
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i > numPairs)
        return;
    const int t = i < numPairs ? int(pairTile[i]) : tiles;
    const int prev = i > 0 ? int(pairTile[i - 1]) : -1;
    for (int u = prev + 1; u <= t; ++u)
        tileOffsets[u] = int(i);
}

// ---- Tile binning (plain path) ----
// The depth-ordered Gaussians are cut into groups of `group` (a multiple of 512), one workgroup
// per group; each of its W wavefronts takes a contiguous slice of the group. Per-tile pair counts
// come from the rectangles directly: +1 / -1 at the four corners of each rectangle in a 2-D
// difference grid of (tilesY + 1) x (tilesX + 1) cells, then 2-D prefix sums. The emission
// enumerates each slice's pairs in the order the per-tile lists need (Gaussian by Gaussian in
// depth order, then the rectangle's tiles row by row), 64 pairs at a time: lane p finds its
// Gaussian by a binary search over the 64 exclusive pair prefixes of the current 64 Gaussians.

constexpr int kBinWaves = 8;           // wavefronts per binning workgroup
constexpr int kBinLdsCells = 12288;    // u32 grid cells in LDS (W wavefronts x cells each)
constexpr int kScanTiles = 32;         // tiles per scan workgroup (a lane each in both wave halves)
constexpr int kScanWaves = 16;         // wavefronts per scan workgroup
constexpr int kScanSubs = kScanWaves * (64 / kScanTiles); // ranges of groups per scan workgroup
constexpr int kMaxScanBlocks = (kBinLdsCells + kScanTiles - 1) / kScanTiles;

// Per-wavefront record of 64 Gaussians of the emission.
struct BinLds {
    uint32_t nt[64];     // tiles of the rectangle
    uint32_t origin[64]; // x0 | y0 << 16 of the rectangle
    uint32_t size[64];   // width | height << 16 in tiles
    uint32_t magic[64];  // 2^32 / width rounded up (width >= 2): k / width = mulhi(k, magic)
    uint32_t gauss[64];
};

// Corner updates of the difference grid D (row stride `stride`) for the rectangles of the
// Gaussians at depth positions [first, first + count) (one wavefront). kGather: the tile counts
// and rectangles are gathered through sorted (nTiles, rect by Gaussian) and written in depth
// order to ntS / rectS for every position (rectS only where the count is non-zero); otherwise
// they are read from ntS / rectS.
// failed (kGather only): the depth sort gave up a lookback wait, so its order may not be a
// permutation; every position then counts as invisible (empty tile lists, nothing out of bounds).
// Ids at or above idLimit (the length of nTiles / rect) count as invisible too.
// indexOrder (kGather false): ntS / rectS are the tile counts and rectangles by Gaussian index
// (positions are Gaussians, invisible ones anywhere) instead of a depth order.
template <bool kGather>
__device__ __forceinline__ void addRects(int64_t n, const uint32_t *__restrict__ sorted,
                                         const uint32_t *__restrict__ nTiles,
                                         const int4 *__restrict__ rect, uint32_t *ntS,
                                         int4 *rectS, int64_t first, int count, uint32_t *D,
                                         int stride, bool failed = false, int64_t idLimit = 0,
                                         bool indexOrder = false)
{
    const int lane = threadIdx.x & 63;
    for (int sub = 0; sub < count; sub += 64) {
        const int64_t i = first + sub + lane;
        uint32_t nt = 0;
        int4 r = make_int4(0, 0, 0, 0);
        if (i < n) {
            if (kGather) {
                const uint32_t g = sorted[i];
                nt = (!failed && int64_t(g) < idLimit) ? nTiles[g] : 0u;
                if (nt)
                    r = rect[g];
                ntS[i] = nt;
                if (nt)
                    rectS[i] = r;
            } else {
                nt = ntS[i];
                if (nt)
                    r = rectS[i];
            }
        }
        if (__ballot(nt != 0) == 0) { // invisible Gaussians sort last in a depth order
            if (kGather || indexOrder)
                continue; // every position is still written / more may follow
            break;
        }
        if (nt) {
            atomicAdd(&D[r.y * stride + r.x], 1u);
            atomicAdd(&D[r.y * stride + r.z], ~0u); // -1 modulo 2^32
            atomicAdd(&D[r.w * stride + r.x], ~0u);
            atomicAdd(&D[r.w * stride + r.z], 1u);
        }
    }
}

// Inclusive 2-D prefix sums, in place, of `grids` grids of rows x cols cells stored back to back
// (all threads of the workgroup; ends with a barrier).
__device__ __forceinline__ void prefix2d(uint32_t *D, int grids, int rows, int cols)
{
    const int cells = rows * cols;
    for (int task = threadIdx.x; task < grids * rows; task += int(blockDim.x)) {
        uint32_t *row = D + (task / rows) * cells + (task % rows) * cols;
        uint32_t run = 0;
#pragma unroll 4
        for (int x = 0; x < cols; ++x) {
            run += row[x];
            row[x] = run;
        }
    }
    __syncthreads();
    for (int task = threadIdx.x; task < grids * cols; task += int(blockDim.x)) {
        uint32_t *col = D + (task / cols) * cells + task % cols;
        uint32_t run = 0;
#pragma unroll 4
        for (int y = 0; y < rows; ++y) {
            run += col[y * cols];
            col[y * cols] = run;
        }
    }
    __syncthreads();
}

// Workgroups are dealt to the eight XCDs in turn (workgroup b runs on XCD b % 8). Groups are
// given to workgroups in chunks of kChunk consecutive groups, chunk k to XCD k % 8: the short
// runs that neighbouring groups write next to each other (their counters, their slices of a
// tile's list) meet in one L2 and leave it as whole lines, while the first chunks (the nearest
// Gaussians, the largest footprints) start first and spread over all XCDs. The grid has
// xcdGridSize(numGroups, kChunk) workgroups; those without a group (-1) exit.
constexpr int kXcds = 8;
inline int xcdGridSize(int numGroups, int chunk)
{
    const int chunks = (numGroups + chunk - 1) / chunk;
    return (chunks + kXcds - 1) / kXcds * kXcds * chunk;
}
template <int kChunk> __device__ __forceinline__ int xcdGroup(int b, int numGroups)
{
    const int i = b / kXcds;
    const int grp = ((i / kChunk) * kXcds + b % kXcds) * kChunk + i % kChunk;
    return grp < numGroups ? grp : -1;
}
constexpr int kCountChunk = 32, kEmitChunk = 8;

// counts[grp * tiles + t] = pairs of group grp in tile t (group-major: each workgroup writes one
// contiguous row, and the emission reads it back as one).
// Also writes the tile counts and rectangles in depth order (ntS, rectS) for the emission.
// sorted == nullptr: the groups are Gaussians in index order, read from nTiles / rect directly.
// The counts of the groups that workgroup blk takes (its chunk of groups; workgroups without a
// group return).
__device__ __forceinline__ void tileCountBody(
    int blk, int64_t n, const uint32_t *__restrict__ sorted, const uint32_t *__restrict__ nTiles,
    const int4 *__restrict__ rect, uint32_t *ntS, int4 *rectS, int tilesX, int tilesY,
    int group, int numGroups, uint32_t *counts, const uint32_t *__restrict__ sortError,
    int64_t idLimit)
{
    extern __shared__ uint32_t D[];
    const int tid = threadIdx.x, wave = tid >> 6;
    const int stride = tilesX + 1, cells = stride * (tilesY + 1);
    for (int t = tid; t < cells; t += 64 * kBinWaves)
        D[t] = 0;
    __syncthreads();
    const int slice = group / kBinWaves;
    const int grp = xcdGroup<kCountChunk>(blk, numGroups);
    if (grp < 0)
        return;
    const int64_t first = int64_t(grp) * group + int64_t(wave) * slice;
    if (sorted) {
        // the sort's passes have finished (stream order), so every workgroup reads the same word
        const bool failed = sortError && *sortError != 0u;
        addRects<true>(n, sorted, nTiles, rect, ntS, rectS, first, slice, D, stride, failed,
                       idLimit);
    } else {
        addRects<false>(n, nullptr, nullptr, nullptr, const_cast<uint32_t *>(nTiles),
                        const_cast<int4 *>(rect), first, slice, D, stride, false, 0, true);
    }
    __syncthreads();
    prefix2d(D, 1, tilesY + 1, stride);
    const int tiles = tilesX * tilesY;
    uint32_t *row = counts + size_t(grp) * size_t(tiles);
    for (int t = tid; t < tiles; t += 64 * kBinWaves) {
        const int ty = t / tilesX, tx = t - ty * tilesX;
        row[t] = D[ty * stride + tx];
    }
}

// side.in: workgroup 0 runs that sum instead (the groups start at workgroup 1; the shift keeps
// each chunk of groups on one XCD, a rotation of the XCDs).
__global__ __launch_bounds__(64 * kBinWaves) void tileCountKernel(
    int64_t n, const uint32_t *__restrict__ sorted, const uint32_t *__restrict__ nTiles,
    const int4 *__restrict__ rect, uint32_t *ntS, int4 *rectS, int tilesX, int tilesY,
    int group, int numGroups, uint32_t *counts, const uint32_t *__restrict__ sortError,
    int64_t idLimit, SideSum side)
{
    if (side.in && blockIdx.x == 0) {
        sumBlock(side.in, side.count, side.dev, side.host, side.seq);
        return;
    }
    tileCountBody(int(blockIdx.x) - (side.in ? 1 : 0), n, sorted, nTiles, rect, ntS, rectS,
                  tilesX, tilesY, group, numGroups, counts, sortError, idLimit);
}

// The item of workgroup blk of a launch over several images: the last one whose first
// workgroup (field) is at or before blk.
template <int BinItem::*kFirst>
__device__ __forceinline__ int binItemOf(const BinItem *__restrict__ items, int numItems, int blk)
{
    int v = 0;
    while (v + 1 < numItems && items[v + 1].*kFirst <= blk)
        ++v;
    return v;
}

// Counts of every image of a batch, one launch (items[v].countBlock0 first): in index order,
// or (sorted) in the depth order sorted + idBase of the image's entries (global ids below
// idLimit), gathering their tile counts and rectangles to ntS / rectS + idBase.
__global__ __launch_bounds__(64 * kBinWaves) void tileCountItemsKernel(
    const BinItem *__restrict__ items, int numItems, const uint32_t *__restrict__ sorted,
    const uint32_t *__restrict__ nTiles, const int4 *__restrict__ rect, uint32_t *ntS,
    int4 *rectS, int64_t idLimit, uint32_t *counts)
{
    const int blk = int(blockIdx.x);
    const BinItem it = items[binItemOf<&BinItem::countBlock0>(items, numItems, blk)];
    if (sorted)
        tileCountBody(blk - it.countBlock0, it.n, sorted + it.idBase, nTiles, rect,
                      ntS + it.idBase, rectS + it.idBase, it.tilesX, it.tilesY, it.group,
                      it.numGroups, counts + it.countsBase, nullptr, idLimit);
    else
        tileCountBody(blk - it.countBlock0, it.n, nullptr, nTiles + it.idBase, rect + it.idBase,
                      nullptr, nullptr, it.tilesX, it.tilesY, it.group, it.numGroups,
                      counts + it.countsBase, nullptr, it.n);
}

// Scan of the group counts, one workgroup per block of kScanTiles consecutive tiles. Lane l of
// wavefront w takes tile l % kScanTiles and the (2w + l / kScanTiles)-th of kScanSubs consecutive
// ranges of groups, so each load and store of a wavefront is two contiguous runs of counts rows.
// Each count becomes, in place, the pairs of the earlier groups in its tile plus the pairs of
// the earlier tiles of its block (the offset of the tile's list inside the block's lists); adding
// the block's own offset (the exclusive scan of the block totals, which the emission does) gives
// the group's first slot in the tile's list. tileTotal[t] receives tile t's pair count,
// tileTotal[tiles + b] block b's.
__device__ __forceinline__ void tileScanBody(int blk, uint32_t *counts, int numGroups, int tiles,
                                             uint32_t *tileTotal)
{
    static_assert(kScanTiles == 32, "two ranges of groups per wavefront");
    __shared__ uint32_t waveSum[kScanWaves][kScanTiles];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & (kScanTiles - 1), half = lane >> 5;
    const int t = int(blk) * kScanTiles + col;
    const bool valid = t < tiles;
    const int per = (numGroups + kScanSubs - 1) / kScanSubs;
    const int sub = wave * 2 + half;
    const int g0 = min(sub * per, numGroups), g1 = min(g0 + per, numGroups);
    uint32_t *cnt = counts + (valid ? t : 0);
    const size_t step = size_t(tiles);
    uint32_t sum = 0;
    if (valid) {
#pragma unroll 16
        for (int g = g0; g < g1; ++g)
            sum += cnt[size_t(g) * step];
    }
    // the range of the lower half comes first within the wavefront
    const uint32_t low = uint32_t(__shfl(int(sum), col, 64));
    if (half == 1)
        waveSum[wave][col] = low + sum;
    __syncthreads();
    uint32_t before = half ? low : 0u, total = 0;
#pragma unroll
    for (int w = 0; w < kScanWaves; ++w) {
        const uint32_t c = waveSum[w][col];
        before += w < wave ? c : 0u;
        total += c;
    }
    uint32_t x = total; // inclusive scan of the tile totals over the block
#pragma unroll
    for (int off = 1; off < kScanTiles; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, kScanTiles);
        if (col >= off)
            x += y;
    }
    if (valid) {
        uint32_t run = before + (x - total);
#pragma unroll 16
        for (int g = g0; g < g1; ++g) {
            const uint32_t c = cnt[size_t(g) * step];
            cnt[size_t(g) * step] = run;
            run += c;
        }
    }
    if (wave == 0 && half == 0) {
        if (valid)
            tileTotal[t] = total;
        if (col == kScanTiles - 1)
            tileTotal[size_t(tiles) + blk] = x;
    }
}

__global__ __launch_bounds__(64 * kScanWaves) void tileScanKernel(uint32_t *counts, int numGroups,
                                                                 int tiles, uint32_t *tileTotal)
{
    tileScanBody(int(blockIdx.x), counts, numGroups, tiles, tileTotal);
}

__global__ __launch_bounds__(64 * kScanWaves) void tileScanItemsKernel(
    const BinItem *__restrict__ items, int numItems, uint32_t *counts, uint32_t *tileTotal)
{
    const int blk = int(blockIdx.x);
    const BinItem it = items[binItemOf<&BinItem::scanBlock0>(items, numItems, blk)];
    tileScanBody(blk - it.scanBlock0, counts + it.countsBase, it.numGroups,
                 it.tilesX * it.tilesY, tileTotal + it.totalsBase);
}

// The tiles ordered by decreasing cost (64 log-scale classes, any order inside a class), so
// that a launch that maps workgroup i to tile heavyFirst[i] starts the longest walks first
// (the whole workgroup; LDS counters cls[64]). The cost of a tile is walkCost[t] when given and
// non-zero -- the largest contributor count of its pixels the last time this view was
// composited, which bounds how far the walk goes -- else its list length. walkCost is then
// cleared for the forward that follows to record the new maxima.
__device__ __forceinline__ void writeHeavyFirst(const uint32_t *__restrict__ tileTotal, int tiles,
                                                int *heavyFirst, uint32_t *cls,
                                                uint32_t *walkCost)
{
    const int tid = threadIdx.x, nt = int(blockDim.x);
    auto classOf = [](uint32_t len) {
        const uint32_t v = len + 1u;
        const int lg = 31 - __clz(v);
        const int q = lg >= 2 ? int((v >> (lg - 2)) & 3u) : int((v << (2 - lg)) & 3u);
        const int c = lg * 4 + q;
        return 63 - (c < 63 ? c : 63); // heavier first
    };
    for (int k = tid; k < 64; k += nt)
        cls[k] = 0;
    __syncthreads();
    auto costOf = [&](int t) {
        const uint32_t w = walkCost ? walkCost[t] : 0u;
        return w ? w : tileTotal[t];
    };
    for (int t = tid; t < tiles; t += nt)
        atomicAdd(&cls[classOf(costOf(t))], 1u);
    __syncthreads();
    if (tid == 0) {
        uint32_t run = 0;
        for (int k = 0; k < 64; ++k) {
            const uint32_t c = cls[k];
            cls[k] = run;
            run += c;
        }
    }
    __syncthreads();
    for (int t = tid; t < tiles; t += nt)
        heavyFirst[atomicAdd(&cls[classOf(costOf(t))], 1u)] = t;
    if (walkCost) {
        __syncthreads();
        for (int t = tid; t < tiles; t += nt)
            walkCost[t] = 0u;
    }
}

// Writes the ids of the Gaussians at positions [first, first + count) of the depth order sorted
// (or of the index order, sorted == nullptr; one wavefront) at their tiles' cursors in mine,
// advancing them. A step gives each of up to four Gaussians a run of
// lanes, one lane per tile of its rectangle (16 lanes each when all four have at most 16 tiles,
// else 32 lanes for two with at most 32, else 64 lanes for one, in as many rounds as it needs).
// Inside one Gaussian the tiles are distinct, so its lanes take their slots with one LDS atomic
// add; the Gaussians of a step take theirs in separate rounds in list order, so equal tiles keep
// the order of the positions.
// (idBase: added to the index-order positions, the ids of a batch's image.)
__device__ __forceinline__ void emitSlice(int64_t n, const uint32_t *__restrict__ sorted,
                                          const uint32_t *__restrict__ ntS,
                                          const int4 *__restrict__ rectS, int64_t first,
                                          int count, int stride, BinLds &L, uint32_t *mine,
                                          uint32_t *tileList, uint32_t capacity, uint32_t idBase)
{
    const int lane = threadIdx.x & 63;
    for (int sub = 0; sub < count; sub += 64) {
        const int64_t i = first + sub + lane;
        uint32_t g = 0, nt = 0;
        int4 r = make_int4(0, 0, 1, 1);
        if (i < n) {
            g = sorted ? sorted[i] : idBase + uint32_t(i);
            nt = ntS[i];
            if (nt)
                r = rectS[i];
        }
        const uint64_t vis = __ballot(nt != 0);
        if (vis == 0) { // invisible Gaussians sort last in a depth order
            if (sorted)
                break;
            continue;
        }
        // the visible Gaussians in lane order, packed to the front of the records (in a depth
        // order they already are a prefix of the lanes)
        const int cnt = __popcll(vis);
        const int rk = __popcll(vis & laneMaskLt());
        const uint32_t w = uint32_t(r.z - r.x), h = uint32_t(r.w - r.y);
        waveSync();
        if (nt) {
            L.nt[rk] = nt;
            L.origin[rk] = uint32_t(r.x) | (uint32_t(r.y) << 16);
            L.size[rk] = w | (h << 16);
            // q = mulhi(k, magic) = k / w exactly for k * w < 2^32 (w >= 2); w == 1 is q = k
            L.magic[rk] = w > 1u ? 0xffffffffu / w + 1u : 0u;
            L.gauss[rk] = g;
        }
        waveSync();
        const uint32_t ntc = L.nt[lane]; // lanes >= cnt: not read
        for (int j = 0; j < cnt;) {
            const uint32_t n0 = uint32_t(__builtin_amdgcn_readlane(int(ntc), j));
            const uint32_t n1 = j + 1 < cnt ? uint32_t(__builtin_amdgcn_readlane(int(ntc), j + 1)) : 0u;
            const uint32_t n2 = j + 2 < cnt ? uint32_t(__builtin_amdgcn_readlane(int(ntc), j + 2)) : 0u;
            const uint32_t n3 = j + 3 < cnt ? uint32_t(__builtin_amdgcn_readlane(int(ntc), j + 3)) : 0u;
            const uint32_t m4 = max(max(n0, n1), max(n2, n3));
            const int shift = m4 <= 16u ? 4 : (max(n0, n1) <= 32u ? 5 : 6); // lanes per Gaussian
            const int slots = 64 >> shift;
            const int q = lane >> shift, k0 = lane & ((1 << shift) - 1);
            const int gi = (j + q) & 63;
            const uint32_t ntg = j + q < cnt ? L.nt[gi] : 0u;
            const uint32_t org = L.origin[gi], wh = L.size[gi];
            const uint32_t magic = L.magic[gi], gg = L.gauss[gi];
            const uint32_t gw = wh & 0xffffu;
            for (uint32_t c0 = 0;; c0 += 64u) {
                const uint32_t k = c0 + uint32_t(k0);
                const bool act = k < ntg;
                const uint32_t qy = gw > 1u ? __umulhi(k, magic) : k;
                const uint32_t qx = k - qy * gw;
                const int cell = int((org >> 16) + qy) * stride + int((org & 0xffffu) + qx);
                // One LDS atomic per round for every lane (the round's lanes add 1, the others 0;
                // lanes without a tile add 0 to a word of their own): a wavefront's LDS
                // instructions execute in order, so a tile shared by two Gaussians of the step
                // gives the earlier one the lower slot.
                uint32_t *dst = act ? &mine[cell] : &L.gauss[lane];
                uint32_t slot = 0;
                for (int sl = 0; sl < slots; ++sl) {
                    const uint32_t v = atomicAdd(dst, (act && q == sl) ? 1u : 0u);
                    slot = q == sl ? v : slot;
                }
                if (act && slot < capacity) // see tileEmitKernel
                    tileList[slot] = gg;
                if (shift < 6 || c0 + 64u >= n0)
                    break;
            }
            j += slots;
        }
    }
}

// Writes every Gaussian id at its final slot of its tile's list. W = blockDim.x / 64 wavefronts
// split the group; each counts its slice per tile in its own difference grid, the counts become
// per-wavefront cursors (tile offset + earlier groups + earlier slices) in place, then each
// emits its slice in order (emitSlice).
// The offset of a block of kScanTiles tiles is the exclusive scan of the block totals, which
// every workgroup computes (at most kMaxScanBlocks values, in one wavefront); with the scanned counts (earlier
// groups and earlier tiles of the block) it gives the group's first slot in each tile's list.
// Group 0's row holds the in-block tile offsets alone, so group 0 writes the tile offsets
// (tileOffsets[0, tiles]) and, if asked, the heavy-first tile order and the tiles by decreasing
// list length (longFirst). The totals equal the pair count the list was sized for whenever the
// depth order is a permutation; after a sort whose lookback wait gave up they may not, and then
// slots and offsets past `capacity` are clamped.
// sorted == nullptr: positions are Gaussians in index order (ntS / rectS by Gaussian), and each
// tile's list comes out in index order.
__device__ __forceinline__ void tileEmitBody(
    int blk, int64_t n, const uint32_t *__restrict__ sorted, const uint32_t *__restrict__ ntS,
    const int4 *__restrict__ rectS, int tilesX, int tilesY, int group,
    int numGroups, const uint32_t *__restrict__ counts, const uint32_t *__restrict__ tileTotal,
    int *tileOffsets, int *heavyFirst, uint32_t *tileList, uint32_t capacity, uint32_t *walkCost,
    uint32_t listBase, uint32_t *listSum, uint32_t *listMax, uint32_t *overflow, int *longFirst,
    uint32_t idBase)
{
    extern __shared__ uint32_t cur[]; // [W][cells]
    __shared__ BinLds L[kBinWaves];
    __shared__ uint32_t blockBase[kMaxScanBlocks];
    __shared__ uint32_t listTotal;
    __shared__ uint32_t cls[64];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int W = int(blockDim.x) >> 6;
    const int stride = tilesX + 1, cells = stride * (tilesY + 1);
    const int tiles = tilesX * tilesY;
    const int slice = group / W;
    const int grp = xcdGroup<kEmitChunk>(blk, numGroups);
    if (grp < 0)
        return;
    const int64_t first = int64_t(grp) * group + int64_t(wave) * slice;
    uint32_t *mine = cur + size_t(wave) * size_t(cells);
    // the group's row of scanned counts; each thread's first kPre tiles are loaded here, so the
    // loads overlap the difference grids
    const uint32_t *row = counts + size_t(grp) * size_t(tiles);
    constexpr int kPre = 4;
    uint32_t pre[kPre];
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
        const int t = tid + k * int(blockDim.x);
        pre[k] = t < tiles ? row[t] : 0u;
    }
    if (wave == 0) { // exclusive scan of the block totals, kPerLane per lane
        constexpr int kPerLane = (kMaxScanBlocks + 63) / 64;
        const int nb = (tiles + kScanTiles - 1) / kScanTiles;
        uint32_t v[kPerLane], local = 0;
#pragma unroll
        for (int k = 0; k < kPerLane; ++k) {
            const int b = lane * kPerLane + k;
            v[k] = b < nb ? tileTotal[size_t(tiles) + size_t(b)] : 0u;
            local += v[k];
        }
        uint32_t x = local;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= off)
                x += y;
        }
        uint32_t run = x - local;
#pragma unroll
        for (int k = 0; k < kPerLane; ++k) {
            if (lane * kPerLane + k < nb)
                blockBase[lane * kPerLane + k] = run;
            run += v[k];
        }
        if (lane == 63)
            listTotal = x;
    }
    for (int t = tid; t < W * cells; t += int(blockDim.x))
        cur[t] = 0;
    __syncthreads();
    if (W > 1) {
        addRects<false>(n, sorted, nullptr, nullptr, const_cast<uint32_t *>(ntS),
                        const_cast<int4 *>(rectS), first, slice, mine, stride, false, 0,
                        sorted == nullptr);
        __syncthreads();
        prefix2d(cur, W, tilesY + 1, stride);
    }
    // the cursors of tile t: its slots for this group, wavefront by wavefront
    auto setCursors = [&](int t, uint32_t scanned) {
        const uint32_t off = blockBase[t / kScanTiles] + scanned;
        if (grp == 0)
            tileOffsets[t] = int(listBase + (off < capacity ? off : capacity));
        const int ty = t / tilesX, tx = t - ty * tilesX;
        const int cell = ty * stride + tx;
        uint32_t base = off;
        for (int w = 0; w < W; ++w) {
            const uint32_t c = cur[w * cells + cell];
            cur[w * cells + cell] = base;
            base += c;
        }
    };
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
        const int t = tid + k * int(blockDim.x);
        if (t < tiles)
            setCursors(t, pre[k]);
    }
    for (int t = tid + kPre * int(blockDim.x); t < tiles; t += int(blockDim.x))
        setCursors(t, row[t]);
    if (grp == 0) {
        if (tid == 0) {
            const uint32_t total = listTotal;
            tileOffsets[tiles] = int(listBase + (total < capacity ? total : capacity));
            if (listSum) { // the batched engine's step status: pairs, longest list set, overflow
                atomicAdd(listSum, total);
                atomicMax(listMax, total);
                if (total > capacity)
                    atomicOr(overflow, 1u);
            }
        }
        if (longFirst) {
            writeHeavyFirst(tileTotal, tiles, longFirst, cls, nullptr);
            __syncthreads(); // cls is reused
        }
        if (heavyFirst)
            writeHeavyFirst(tileTotal, tiles, heavyFirst, cls, walkCost);
    }
    __syncthreads(); // every cursor is set before the wavefronts emit
    emitSlice(n, sorted, ntS, rectS, first, slice, stride, L[wave], mine, tileList + listBase,
              capacity, idBase);
}

__global__ __launch_bounds__(64 * kBinWaves) void tileEmitKernel(
    int64_t n, const uint32_t *__restrict__ sorted, const uint32_t *__restrict__ ntS,
    const int4 *__restrict__ rectS, int tilesX, int tilesY, int group,
    int numGroups, const uint32_t *__restrict__ counts, const uint32_t *__restrict__ tileTotal,
    int *tileOffsets, int *heavyFirst, uint32_t *tileList, uint32_t capacity, uint32_t *walkCost,
    uint32_t listBase, uint32_t *listSum, uint32_t *listMax, uint32_t *overflow, int *longFirst)
{
    tileEmitBody(int(blockIdx.x), n, sorted, ntS, rectS, tilesX, tilesY, group, numGroups, counts,
                 tileTotal, tileOffsets, heavyFirst, tileList, capacity, walkCost, listBase,
                 listSum, listMax, overflow, longFirst, 0u);
}

// Emission of every image of a batch, one launch: image v's ids (its entries plus
// items[v].idBase in index order, or sorted[idBase + i] in the depth order with ntS / rectS as
// gathered by the count launch) go to its slice of the lists (items[v].listBase, capacity
// slots), its offsets and longest-first order (nullable) to their places; listSum / listMax /
// overflow as for one image.
__global__ __launch_bounds__(64 * kBinWaves) void tileEmitItemsKernel(
    const BinItem *__restrict__ items, int numItems, const uint32_t *__restrict__ sorted,
    const uint32_t *__restrict__ ntS, const int4 *__restrict__ rectS,
    const uint32_t *__restrict__ counts, const uint32_t *__restrict__ tileTotal, int *tileOffsets,
    uint32_t *tileList, uint32_t capacity, uint32_t *listSum, uint32_t *listMax,
    uint32_t *overflow, int *longFirst)
{
    const int blk = int(blockIdx.x);
    const BinItem it = items[binItemOf<&BinItem::emitBlock0>(items, numItems, blk)];
    tileEmitBody(blk - it.emitBlock0, it.n, sorted ? sorted + it.idBase : nullptr,
                 ntS + it.idBase, rectS + it.idBase, it.tilesX, it.tilesY, it.group,
                 it.numGroups, counts + it.countsBase, tileTotal + it.totalsBase,
                 tileOffsets + it.offsetsBase, nullptr, tileList, capacity, nullptr, it.listBase,
                 listSum, listMax, overflow, longFirst ? longFirst + it.orderBase : nullptr,
                 it.idBase);
}

// ------------------------------------------------------------------ per-tile depth order
// The plain path emits each tile's list in Gaussian index order and then sorts every list
// stably on the Gaussians' depth keys (the float bits of the camera-space depth: positive, so
// their unsigned order is the numeric one). The result is the reference's (depth bits, index)
// order of each tile (Rasterizer.cpp:215-260) with no sort over all Gaussians, whose passes
// are global round trips. One workgroup per tile, the longest lists first. Keys are taken
// relative to the smallest key of the list, and only the 8-bit digits below the list's key
// range are sorted (three passes when its depths lie within a factor of two).
// A list of at most kTsCap entries stays in registers: each pass ranks the keys per wavefront
// (ballot matching, stable in position order), scatters keys and positions to LDS and reads
// them back in the new order. A longer list runs the same passes in chunks of kTsLongChunk through
// global scratch (its histogram first, then the chunks in order, each advancing the digits'
// output cursors).
constexpr int kTsThreads = 512;
constexpr int kTsWaves = kTsThreads / 64;
constexpr int kTsItems = 22;
constexpr int kTsCap = kTsThreads * kTsItems; // 11264: two workgroups' LDS fit in one CU
constexpr int kTsLongItems = 8;                // keys per lane of a chunk of a longer list
constexpr int kTsLongChunk = kTsThreads * kTsLongItems;

struct TileSortLds {
    uint32_t key[kTsCap];
    uint16_t pos[kTsCap];
    uint32_t waveCount[kTsWaves][256];
    uint32_t digitRun[256];
    uint32_t part[kTsWaves];
    uint32_t lo[kTsWaves], hi[kTsWaves];
};

__device__ __forceinline__ int tsPosition(int wave, int ipw, int it, int lane)
{
    return (wave * ipw + it) * 64 + lane;
}

// Per-wavefront stable ranks of one chunk, whose position p = tsPosition(wave, ipw, it, lane)
// (p < count) holds k[it]: L.waveCount[w][d] = keys of digit d in wavefront w, rank[it] = the
// key's place among the keys of its digit in its wavefront. The lanes of an item that share a
// digit find each other through a 64-bit mask per (wavefront, digit) in LDS: each lane ORs in its
// bit, then every lane reads the mask back (a wavefront's LDS instructions complete in order),
// and the lowest lane of the digit advances the count and clears the mask. The masks live in
// L.key, which holds nothing while keys are ranked.
template <int kItems>
__device__ __forceinline__ void tsCount(const uint32_t (&k)[kItems], uint32_t (&rank)[kItems],
                                        int ipw, int count, uint32_t base, int shift,
                                        TileSortLds &L)
{
    static_assert(sizeof(TileSortLds::key) >= kTsWaves * 256 * sizeof(uint64_t), "digit masks");
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    uint64_t *masks = reinterpret_cast<uint64_t *>(L.key);
    uint64_t *mask = masks + wave * 256;
    __syncthreads(); // earlier readers of waveCount and of key are done
    for (int j = tid; j < kTsWaves * 256; j += kTsThreads) {
        (&L.waveCount[0][0])[j] = 0;
        masks[j] = 0;
    }
    __syncthreads();
    const uint64_t bit = uint64_t(1) << lane, below = bit - 1;
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        rank[it] = 0;
        const int p = tsPosition(wave, ipw, it, lane);
        if (it < ipw && p < count) {
            const uint32_t d = ((k[it] - base) >> shift) & 255u;
            atomicOr(reinterpret_cast<unsigned long long *>(&mask[d]), (unsigned long long)bit);
            const uint64_t m = mask[d];
            const uint32_t prev = L.waveCount[wave][d];
            rank[it] = prev + uint32_t(__popcll(m & below));
            if ((m & below) == 0) {
                L.waveCount[wave][d] = prev + uint32_t(__popcll(m));
                mask[d] = 0;
            }
        }
        // one item at a time (each item's LDS instructions after the previous item's)
        __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
}

// L.digitRun[d] = sum of v over the digits below d, where thread d < 256 holds v (all threads).
__device__ __forceinline__ void tsScan(uint32_t v, TileSortLds &L)
{
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off)
            x += y;
    }
    if (tid < 256 && lane == 63)
        L.part[wave] = x;
    __syncthreads();
    if (tid < 256) {
        uint32_t before = x - v;
        for (int w = 0; w < wave; ++w)
            before += L.part[w];
        L.digitRun[tid] = before;
    }
    __syncthreads();
}

// Turns the chunk's counts into output slots: L.waveCount[w][d] = the slot of wavefront w's
// first key of digit d, from the digit's cursor L.digitRun[d], which moves past the chunk.
__device__ __forceinline__ void tsBases(TileSortLds &L)
{
    const int tid = threadIdx.x;
    if (tid < 256) {
        uint32_t run = L.digitRun[tid];
#pragma unroll
        for (int w = 0; w < kTsWaves; ++w) {
            const uint32_t c = L.waveCount[w][tid];
            L.waveCount[w][tid] = run;
            run += c;
        }
        L.digitRun[tid] = run;
    }
    __syncthreads();
}

// The smallest and largest of the workgroup's values, and the 8-bit passes that sort keys
// relative to the smallest.
__device__ __forceinline__ int tsRange(uint32_t lo, uint32_t hi, TileSortLds &L, uint32_t *base)
{
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, uint32_t(__shfl_xor(int(lo), off, 64)));
        hi = max(hi, uint32_t(__shfl_xor(int(hi), off, 64)));
    }
    if (lane == 0) {
        L.lo[wave] = lo;
        L.hi[wave] = hi;
    }
    __syncthreads();
    lo = L.lo[0];
    hi = L.hi[0];
#pragma unroll
    for (int w = 1; w < kTsWaves; ++w) {
        lo = min(lo, L.lo[w]);
        hi = max(hi, L.hi[w]);
    }
    *base = lo;
    const uint32_t range = hi > lo ? hi - lo : 0u;
    const int bits = range ? 32 - __clz(int(range)) : 0;
    return (bits + 7) >> 3;
}

// Sorts a list of count <= kTsCap entries at listIn + off into listOut + off.
__device__ __forceinline__ void tsSortShort(int off, int count, const uint32_t *__restrict__ listIn,
                                            uint32_t *__restrict__ listOut,
                                            const uint32_t *__restrict__ depthKey, TileSortLds &L)
{
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int ipw = (count + kTsThreads - 1) / kTsThreads;
    uint32_t k[kTsItems], v[kTsItems];
    uint32_t lo = 0xffffffffu, hi = 0u;
#pragma unroll
    for (int it = 0; it < kTsItems; ++it) {
        k[it] = 0u;
        v[it] = 0u;
        const int p = tsPosition(wave, ipw, it, lane);
        if (it < ipw && p < count) {
            k[it] = depthKey[listIn[off + p]];
            v[it] = uint32_t(p);
            lo = min(lo, k[it]);
            hi = max(hi, k[it]);
        }
    }
    uint32_t base;
    const int passes = tsRange(lo, hi, L, &base);
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = 8 * pass;
        uint32_t rank[kTsItems];
        tsCount<kTsItems>(k, rank, ipw, count, base, shift, L);
        uint32_t tot = 0;
        if (tid < 256) {
#pragma unroll
            for (int w = 0; w < kTsWaves; ++w)
                tot += L.waveCount[w][tid];
        }
        tsScan(tot, L);
        tsBases(L);
#pragma unroll
        for (int it = 0; it < kTsItems; ++it) {
            const int p = tsPosition(wave, ipw, it, lane);
            if (it < ipw && p < count) {
                const uint32_t d = ((k[it] - base) >> shift) & 255u;
                const uint32_t slot = L.waveCount[wave][d] + rank[it];
                L.key[slot] = k[it];
                L.pos[slot] = uint16_t(v[it]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < kTsItems; ++it) {
            const int p = tsPosition(wave, ipw, it, lane);
            if (it < ipw && p < count) {
                k[it] = L.key[p];
                v[it] = L.pos[p];
            }
        }
    }
#pragma unroll
    for (int it = 0; it < kTsItems; ++it) {
        const int p = tsPosition(wave, ipw, it, lane);
        if (it < ipw && p < count)
            listOut[off + p] = listIn[off + int(v[it])];
    }
}

// Sorts a list of any length at listIn + off into listOut + off, passing through (gk0, gv0)
// and (gk1, gv1) at the same offsets.
__device__ __forceinline__ void tsSortLong(int off, int count, const uint32_t *__restrict__ listIn,
                                           uint32_t *__restrict__ listOut,
                                           const uint32_t *__restrict__ depthKey, uint32_t *gk0,
                                           uint32_t *gv0, uint32_t *gk1, uint32_t *gv1,
                                           TileSortLds &L)
{
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    uint32_t lo = 0xffffffffu, hi = 0u;
    for (int q = tid; q < count; q += kTsThreads) {
        const uint32_t key = depthKey[listIn[off + q]];
        lo = min(lo, key);
        hi = max(hi, key);
    }
    uint32_t base;
    const int passes = tsRange(lo, hi, L, &base);
    if (passes == 0) {
        for (int q = tid; q < count; q += kTsThreads)
            listOut[off + q] = listIn[off + q];
        return;
    }
    const uint32_t *kin = nullptr, *vin = nullptr;
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = 8 * pass;
        const bool last = pass == passes - 1;
        uint32_t *kout = (pass & 1) ? gk1 : gk0, *vout = (pass & 1) ? gv1 : gv0;
        // the digit histogram of the whole list (waveCount[0] as counters), then its scan
        __syncthreads();
        if (tid < 256)
            L.waveCount[0][tid] = 0u;
        __syncthreads();
        for (int q = tid; q < count; q += kTsThreads) {
            const uint32_t key = pass == 0 ? depthKey[listIn[off + q]] : kin[off + q];
            atomicAdd(&L.waveCount[0][((key - base) >> shift) & 255u], 1u);
        }
        __syncthreads();
        tsScan(tid < 256 ? L.waveCount[0][tid] : 0u, L);
        for (int c0 = 0; c0 < count; c0 += kTsLongChunk) {
            const int cnt = count - c0 < kTsLongChunk ? count - c0 : kTsLongChunk;
            uint32_t k[kTsLongItems], v[kTsLongItems], rank[kTsLongItems];
#pragma unroll
            for (int it = 0; it < kTsLongItems; ++it) {
                k[it] = 0u;
                v[it] = 0u;
                const int p = tsPosition(wave, kTsLongItems, it, lane);
                if (p < cnt) {
                    const int q = off + c0 + p;
                    if (pass == 0) {
                        v[it] = listIn[q];
                        k[it] = depthKey[v[it]];
                    } else {
                        k[it] = kin[q];
                        v[it] = vin[q];
                    }
                }
            }
            tsCount<kTsLongItems>(k, rank, kTsLongItems, cnt, base, shift, L);
            tsBases(L);
#pragma unroll
            for (int it = 0; it < kTsLongItems; ++it) {
                const int p = tsPosition(wave, kTsLongItems, it, lane);
                if (p < cnt) {
                    const uint32_t d = ((k[it] - base) >> shift) & 255u;
                    const int slot = off + int(L.waveCount[wave][d] + rank[it]);
                    if (last) {
                        listOut[slot] = v[it];
                    } else {
                        kout[slot] = k[it];
                        vout[slot] = v[it];
                    }
                }
            }
        }
        kin = kout;
        vin = vout;
    }
}

// The workgroup sorts the list of tile order[blk] (tile blk without an order).
__device__ __forceinline__ void tileSortBody(int blk, const int *__restrict__ tileOffsets,
                                             const int *__restrict__ order,
                                             const uint32_t *__restrict__ listIn,
                                             uint32_t *__restrict__ listOut,
                                             const uint32_t *__restrict__ depthKey, uint32_t *gk0,
                                             uint32_t *gv0, uint32_t *gk1, uint32_t *gv1)
{
    __shared__ TileSortLds L;
    const int t = order ? order[blk] : blk;
    const int off = tileOffsets[t], count = tileOffsets[t + 1] - off;
    if (count <= 0)
        return;
    if (count <= 64) {
        // one wavefront, no barrier: lane i's key is ranked by counting the smaller keys (and
        // the equal keys of lower lanes) among the list's
        if (threadIdx.x >= 64)
            return;
        const int lane = int(threadIdx.x);
        const bool valid = lane < count;
        const uint32_t id = listIn[off + (valid ? lane : 0)];
        const uint32_t key = depthKey[id];
        uint32_t rank = 0;
        for (int j = 0; j < count; ++j) {
            const uint32_t kj = uint32_t(__builtin_amdgcn_readlane(int(key), j));
            rank += (kj < key) | ((kj == key) & (j < lane)) ? 1u : 0u;
        }
        if (valid)
            listOut[off + int(rank)] = id;
        return;
    }
    if (count <= kTsThreads) {
        // one key per thread, the keys in LDS (one barrier), each ranked by counting as above
        const int tid = int(threadIdx.x);
        const bool valid = tid < count;
        const uint32_t id = listIn[off + (valid ? tid : 0)];
        const uint32_t key = depthKey[id];
        if (valid)
            L.key[tid] = key;
        __syncthreads();
        if (!valid)
            return;
        uint32_t rank = 0;
#pragma unroll 8
        for (int j = 0; j < count; ++j) {
            const uint32_t kj = L.key[j];
            rank += (kj < key) | ((kj == key) & (j < tid)) ? 1u : 0u;
        }
        listOut[off + int(rank)] = id;
        return;
    }
    if (count <= kTsCap)
        tsSortShort(off, count, listIn, listOut, depthKey, L);
    else
        tsSortLong(off, count, listIn, listOut, depthKey, gk0, gv0, gk1, gv1, L);
}

// Workgroup b sorts the list of tile order[b] (every tile when order lists them all).
__global__ __launch_bounds__(kTsThreads, 4) void tileSortKernel(
    const int *__restrict__ tileOffsets, const int *__restrict__ order,
    const uint32_t *__restrict__ listIn, uint32_t *__restrict__ listOut,
    const uint32_t *__restrict__ depthKey, uint32_t *gk0, uint32_t *gv0, uint32_t *gk1,
    uint32_t *gv1)
{
    tileSortBody(int(blockIdx.x), tileOffsets, order, listIn, listOut, depthKey, gk0, gv0, gk1,
                 gv1);
}

// The tiles of every image of a batch: workgroup b sorts rank sortMap[b] & 0xffffff of image
// sortMap[b] >> 24 (its longest-first order).
__global__ __launch_bounds__(kTsThreads, 4) void tileSortItemsKernel(
    const BinItem *__restrict__ items, const uint32_t *__restrict__ sortMap,
    const int *__restrict__ tileOffsets, const int *__restrict__ longFirst,
    const uint32_t *__restrict__ listIn, uint32_t *__restrict__ listOut,
    const uint32_t *__restrict__ depthKey, uint32_t *gk0, uint32_t *gv0, uint32_t *gk1,
    uint32_t *gv1)
{
    const uint32_t m = sortMap[blockIdx.x];
    const BinItem &it = items[m >> 24];
    tileSortBody(int(m & 0xffffffu), tileOffsets + it.offsetsBase, longFirst + it.orderBase,
                 listIn, listOut, depthKey, gk0, gv0, gk1, gv1);
}

inline int blocksFor(int64_t n, int per) { return int((n + per - 1) / per); }

} // namespace

size_t radixHistSize(int64_t count)
{
    return size_t(256) * size_t(blocksFor(count, kRadixThreads * kMinItems) + 1);
}

size_t radixScanScratchSize(int64_t) { return 256; } // the digit totals

namespace {
template <int kItems>
int radixPasses(hipStream_t s, uint32_t *keys, uint32_t *vals, uint32_t *keysAlt, uint32_t *valsAlt,
                int64_t count, int bits, RadixScratch &scratch, bool *resultInAlt,
                const uint32_t *countDev)
{
    const int numBlocks = blocksFor(count, kRadixThreads * kItems);
    U3D_CHECK(size_t(numBlocks) * 256 <= scratch.histCap, U3D_ERR_CAPACITY, "radix histogram scratch too small");
    U3D_CHECK(scratch.scanTmpCap >= 256, U3D_ERR_CAPACITY, "radix digit-total scratch too small");
    uint32_t *kin = keys, *vin = vals, *kout = keysAlt, *vout = valsAlt;
    const int digit = scratch.radixBits >= 1 && scratch.radixBits <= 8 ? scratch.radixBits : 8;
    for (int shift = 0; shift < bits; shift += digit) {
        const int b = bits - shift < digit ? bits - shift : digit;
        hipLaunchKernelGGL(radixUpsweep<kItems>, dim3(numBlocks), dim3(kRadixThreads), 0, s, kin,
                           count, shift, b, numBlocks, scratch.hist, countDev);
        U3D_LAUNCH_CHECK();
        hipLaunchKernelGGL(radixRowScan, dim3(256), dim3(kRadixThreads), 0, s, scratch.hist,
                           numBlocks, scratch.scanTmp);
        U3D_LAUNCH_CHECK();
        if (shift == 0 && scratch.valsAreIndex)
            hipLaunchKernelGGL((radixDownsweep<kItems, true>), dim3(numBlocks), dim3(kRadixThreads),
                               0, s, kin, vin, kout, vout, count, shift, b, numBlocks, scratch.hist,
                               scratch.scanTmp, countDev);
        else
            hipLaunchKernelGGL((radixDownsweep<kItems, false>), dim3(numBlocks), dim3(kRadixThreads),
                               0, s, kin, vin, kout, vout, count, shift, b, numBlocks, scratch.hist,
                               scratch.scanTmp, countDev);
        U3D_LAUNCH_CHECK();
        uint32_t *tk = kin, *tv = vin;
        kin = kout;
        vin = vout;
        kout = tk;
        vout = tv;
        *resultInAlt = !*resultInAlt;
    }
    return U3D_OK;
}
} // namespace

size_t onesweepScratchWords() { return kOsStateWords + kOsLookbackWords + kOsPartialWords; }

const uint32_t *onesweepErrorPtr(const uint64_t *scratch)
{
    return reinterpret_cast<const uint32_t *>(scratch) + kOsError;
}

int onesweepErrorWord(uint64_t *scratch, hipStream_t s, uint32_t *value)
{
    uint32_t *state = reinterpret_cast<uint32_t *>(scratch);
    U3D_HIP(hipMemcpyAsync(value, state + kOsError, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    U3D_HIP(hipStreamSynchronize(s));
    if (*value)
        U3D_TRY(launchFillI32(s, reinterpret_cast<int *>(state + kOsError), 0, 1));
    return U3D_OK;
}

namespace {
int onesweepSort(hipStream_t s, uint32_t *keys, uint32_t *vals, uint32_t *keysAlt,
                 uint32_t *valsAlt, int64_t count, int bits, RadixScratch &scratch,
                 bool *resultInAlt)
{
    const int numBlocks = blocksFor(count, kOsBlockKeys);
    uint32_t *state = reinterpret_cast<uint32_t *>(scratch.onesweep);
    uint64_t *lookback = scratch.onesweep + kOsStateWords;
    uint32_t *partials = reinterpret_cast<uint32_t *>(lookback + kOsLookbackWords);
    // a new epoch per call (never 0, the value of a zero-filled word); 30 bits with the pass
    uint32_t &epoch = *scratch.onesweepEpoch;
    epoch = (epoch + 1u) & 0x3fffffffu;
    if (epoch == 0)
        epoch = 1;
    const int64_t per = (count + kOsHistBlocks - 1) / kOsHistBlocks;
    const int spinLimit = scratch.spinLimit != 0 ? scratch.spinLimit : kOsSpinLimit;
    hipLaunchKernelGGL(onesweepHistKernel, dim3(kOsHistBlocks + (scratch.sideIn ? 1 : 0)),
                       dim3(kOsThreads), 0, s, keys, count, per, bits, partials, state,
                       scratch.sideIn, scratch.sideCount, scratch.sideDev, scratch.sideHost,
                       scratch.sideSeq);
    U3D_LAUNCH_CHECK();
    uint32_t *kin = keys, *vin = vals, *kout = keysAlt, *vout = valsAlt;
    for (int pass = 0; pass * 8 < bits; ++pass) {
        const int shift = pass * 8;
        const int b = bits - shift < 8 ? bits - shift : 8;
        if (pass == 0 && scratch.valsAreIndex)
            hipLaunchKernelGGL(onesweepPassKernel<true>, dim3(numBlocks), dim3(kOsThreads), 0, s,
                               kin, vin, kout, vout, count, shift, b, pass, epoch, partials,
                               lookback, state, spinLimit);
        else
            hipLaunchKernelGGL(onesweepPassKernel<false>, dim3(numBlocks), dim3(kOsThreads), 0, s,
                               kin, vin, kout, vout, count, shift, b, pass, epoch, partials,
                               lookback, state, spinLimit);
        U3D_LAUNCH_CHECK();
        std::swap(kin, kout);
        std::swap(vin, vout);
        *resultInAlt = !*resultInAlt;
    }
    return U3D_OK;
}
} // namespace

int radixSortPairs(hipStream_t s, uint32_t *keys, uint32_t *vals, uint32_t *keysAlt,
                   uint32_t *valsAlt, int64_t count, int bits, RadixScratch &scratch,
                   bool *resultInAlt, const uint32_t *countDev)
{
    *resultInAlt = false;
    const bool small = scratch.onesweep && scratch.onesweepEpoch && !countDev && bits <= 32 &&
                       count > 0 && count <= int64_t(kOsMaxBlocks) * kOsBlockKeys;
    if (scratch.sideIn && !(small && bits > 0))
        U3D_TRY(launchSumU32(s, scratch.sideIn, scratch.sideCount, scratch.sideDev,
                             scratch.sideHost, scratch.sideSeq));
    if (count <= 0 || bits <= 0)
        return U3D_OK;
    if (small) {
        U3D_CHECK(scratch.onesweepCap >= onesweepScratchWords(), U3D_ERR_CAPACITY,
                  "small-sort scratch too small");
        return onesweepSort(s, keys, vals, keysAlt, valsAlt, count, bits, scratch, resultInAlt);
    }
    const int items = scratch.items == 4 || scratch.items == 8 ? scratch.items : itemsFor(count);
    if (items == 4)
        return radixPasses<4>(s, keys, vals, keysAlt, valsAlt, count, bits, scratch, resultInAlt,
                              countDev);
    return radixPasses<8>(s, keys, vals, keysAlt, valsAlt, count, bits, scratch, resultInAlt,
                          countDev);
}

int launchIota(hipStream_t s, uint32_t *vals, int64_t count)
{
    if (count <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(iotaKernel, dim3(blocksFor(count, 256)), dim3(256), 0, s, vals, count);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchGatherU32(hipStream_t s, const uint32_t *in, const uint32_t *idx, uint32_t *out,
                    int64_t count, int64_t inCount)
{
    if (count <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(gatherKernel, dim3(blocksFor(count, 256)), dim3(256), 0, s, in, idx, out,
                       count, inCount);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

size_t scanScratchSize(int64_t count) { return size_t(blocksFor(count, kScanBlockItems)) + 2; }

int launchSumU32(hipStream_t s, const uint32_t *in, int64_t count, uint32_t *devOut,
                 uint64_t *hostOut, uint32_t seq)
{
    hipLaunchKernelGGL(sumKernel, dim3(1), dim3(1024), 0, s, in, count, devOut, hostOut, seq);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int exclusiveScanU32(hipStream_t s, const uint32_t *in, uint32_t *out, int64_t count,
                     uint32_t *total, uint32_t *blockScratch, size_t blockScratchCap)
{
    if (count <= 0) {
        if (total)
            U3D_TRY(launchFillI32(s, reinterpret_cast<int *>(total), 0, 1));
        return U3D_OK;
    }
    const int nb = blocksFor(count, kScanBlockItems);
    U3D_CHECK(size_t(nb) + 1 <= blockScratchCap, U3D_ERR_CAPACITY, "scan scratch too small");
    hipLaunchKernelGGL(scanReduce, dim3(nb), dim3(256), 0, s, in, count, blockScratch);
    U3D_LAUNCH_CHECK();
    if (nb <= kScanFusedBlocks) {
        hipLaunchKernelGGL(scanDownsweepFused, dim3(nb), dim3(256), 0, s, in, out, count,
                           blockScratch, total);
        U3D_LAUNCH_CHECK();
        return U3D_OK;
    }
    hipLaunchKernelGGL(scanSingleBlock, dim3(1), dim3(kScanThreads), 0, s, blockScratch,
                       int64_t(nb), total);
    U3D_LAUNCH_CHECK();
    hipLaunchKernelGGL(scanDownsweep, dim3(nb), dim3(256), 0, s, in, out, count, blockScratch);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int tileBinningGroup(int64_t n, int tilesX, int tilesY)
{
    const int64_t tiles = int64_t(tilesX) * tilesY;
    if (tiles <= 0 || int64_t(tilesX + 1) * (tilesY + 1) > kBinLdsCells)
        return 0;
    // about 2^20 (group, tile) counters at most; 64 Gaussians per wavefront slice at least
    constexpr int64_t kUnit = 64 * kBinWaves;
    const int64_t want = (n * tiles + (int64_t(1) << 20) - 1) >> 20;
    int64_t g = want < kUnit ? kUnit : want;
    g = (g + kUnit - 1) / kUnit * kUnit;
    return int(g < (int64_t(1) << 22) ? g : (int64_t(1) << 22));
}

size_t tileBinningTotalsSize(int tilesX, int tilesY)
{
    const size_t tiles = size_t(tilesX) * size_t(tilesY);
    return tiles + (tiles + kScanTiles - 1) / kScanTiles;
}

size_t tileBinningCountsSize(int64_t n, int tilesX, int tilesY)
{
    const int g = tileBinningGroup(n, tilesX, tilesY);
    return g > 0 ? size_t(blocksFor(n > 0 ? n : 1, g)) * size_t(tilesX) * size_t(tilesY) : 0;
}

int launchTileCounts(hipStream_t s, int64_t n, const uint32_t *sorted, const uint32_t *nTiles,
                     const int4 *rect, uint32_t *ntS, int4 *rectS, int tilesX, int tilesY,
                     uint32_t *counts, size_t countsCap, uint32_t *tileTotal,
                     const uint32_t *sortError, int64_t idLimit, const SideSum *side)
{
    const int group = tileBinningGroup(n, tilesX, tilesY);
    U3D_CHECK(group > 0, U3D_ERR_UNSUPPORTED, "tile binning needs (tilesX + 1) (tilesY + 1) <= " + std::to_string(kBinLdsCells));
    if (n <= 0)
        return U3D_OK;
    const int tiles = tilesX * tilesY;
    const int numGroups = blocksFor(n, group);
    U3D_CHECK(size_t(numGroups) * size_t(tiles) <= countsCap, U3D_ERR_CAPACITY, "tile binning counters too small");
    const int cells = (tilesX + 1) * (tilesY + 1);
    const SideSum sd = side ? *side : SideSum{};
    hipLaunchKernelGGL(tileCountKernel,
                       dim3(xcdGridSize(numGroups, kCountChunk) + (sd.in ? 1 : 0)),
                       dim3(64 * kBinWaves), size_t(cells) * sizeof(uint32_t), s, n, sorted,
                       nTiles, rect, ntS, rectS, tilesX, tilesY, group, numGroups, counts,
                       sortError, sorted ? idLimit : n, sd);
    U3D_LAUNCH_CHECK();
    hipLaunchKernelGGL(tileScanKernel, dim3((tiles + kScanTiles - 1) / kScanTiles),
                       dim3(64 * kScanWaves), 0, s, counts, numGroups, tiles, tileTotal);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchTileEmit(hipStream_t s, int64_t n, const uint32_t *sorted, const uint32_t *ntS,
                   const int4 *rectS, int tilesX, int tilesY, const uint32_t *counts,
                   const uint32_t *tileTotal, int *tileOffsets, uint32_t *tileList,
                   size_t listCapacity, int *heavyFirst, uint32_t *walkCost,
                   const BinningList *list, int *longFirst)
{
    const BinningList plain{n, 0, nullptr, nullptr, nullptr};
    const BinningList &bl = list ? *list : plain;
    const int group = tileBinningGroup(n, tilesX, tilesY);
    U3D_CHECK(group > 0, U3D_ERR_UNSUPPORTED, "tile binning needs (tilesX + 1) (tilesY + 1) <= " + std::to_string(kBinLdsCells));
    const int tiles = tilesX * tilesY;
    if (n <= 0)
        return launchFillI32(s, tileOffsets, int(bl.listBase), int64_t(tiles) + 1);
    const int numGroups = blocksFor(n, group);
    const int cells = (tilesX + 1) * (tilesY + 1);
    // wavefronts per emission workgroup: as many slices as their grids fit in LDS
    int waves = kBinWaves;
    while (waves > 1 && waves * cells > kBinLdsCells)
        waves >>= 1;
    hipLaunchKernelGGL(tileEmitKernel, dim3(xcdGridSize(numGroups, kEmitChunk)), dim3(64 * waves),
                       size_t(waves) * size_t(cells) * sizeof(uint32_t), s, n, sorted, ntS, rectS,
                       tilesX, tilesY, group, numGroups, counts, tileTotal, tileOffsets,
                       heavyFirst, tileList,
                       uint32_t(listCapacity < size_t(INT32_MAX) ? listCapacity : size_t(INT32_MAX)),
                       walkCost, bl.listBase, bl.listSum, bl.listMax, bl.overflow, longFirst);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchTileBinning(hipStream_t s, int64_t n, const uint32_t *sorted, const uint32_t *nTiles,
                      const int4 *rect, uint32_t *ntS, int4 *rectS, int tilesX, int tilesY,
                      uint32_t *counts, size_t countsCap, uint32_t *tileTotal, int *tileOffsets,
                      uint32_t *tileList, size_t listCapacity, int *heavyFirst,
                      const uint32_t *sortError, uint32_t *walkCost, const BinningList *list)
{
    const int64_t idLimit = list ? list->idLimit : n;
    U3D_TRY(launchTileCounts(s, n, sorted, nTiles, rect, ntS, rectS, tilesX, tilesY, counts,
                             countsCap, tileTotal, sortError, idLimit));
    return launchTileEmit(s, n, sorted, sorted ? ntS : nTiles, sorted ? rectS : rect, tilesX,
                          tilesY, counts, tileTotal, tileOffsets, tileList, listCapacity,
                          heavyFirst, walkCost, list, nullptr);
}

size_t tileSortScratchSize(size_t listCapacity) { return 4 * listCapacity; }

int launchTileSort(hipStream_t s, int tiles, const int *tileOffsets, const int *order,
                   const uint32_t *listIn, uint32_t *listOut, const uint32_t *depthKey,
                   uint32_t *scratch, size_t scratchCap, size_t listCapacity)
{
    if (tiles <= 0)
        return U3D_OK;
    U3D_CHECK(scratchCap >= tileSortScratchSize(listCapacity), U3D_ERR_CAPACITY,
              "tile sort scratch too small");
    uint32_t *gk0 = scratch, *gv0 = scratch + listCapacity;
    uint32_t *gk1 = scratch + 2 * listCapacity, *gv1 = scratch + 3 * listCapacity;
    hipLaunchKernelGGL(tileSortKernel, dim3(tiles), dim3(kTsThreads), 0, s, tileOffsets, order,
                       listIn, listOut, depthKey, gk0, gv0, gk1, gv1);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int binItemsLayout(std::vector<BinItem> &items, BinItemsGrid *grid, std::vector<uint32_t> *sortMap)
{
    U3D_CHECK(items.size() <= 256, U3D_ERR_CAPACITY, "at most 256 images per binning launch");
    BinItemsGrid g;
    g.emitWaves = kBinWaves;
    int mostTiles = 0;
    for (BinItem &it : items) {
        const int tiles = it.tilesX * it.tilesY;
        it.group = tileBinningGroup(it.n > 0 ? it.n : 1, it.tilesX, it.tilesY);
        U3D_CHECK(it.group > 0, U3D_ERR_UNSUPPORTED, "tile binning needs (tilesX + 1) (tilesY + 1) <= " + std::to_string(kBinLdsCells));
        it.numGroups = blocksFor(it.n > 0 ? it.n : 1, it.group); // group 0 writes the offsets
        it.countsBase = uint32_t(g.counts);
        it.totalsBase = uint32_t(g.totals);
        it.orderBase = uint32_t(g.orders);
        it.countBlock0 = g.countBlocks;
        it.scanBlock0 = g.scanBlocks;
        it.emitBlock0 = g.emitBlocks;
        it.sortBlock0 = g.sortBlocks;
        g.counts += size_t(it.numGroups) * size_t(tiles);
        g.totals += tileBinningTotalsSize(it.tilesX, it.tilesY);
        g.orders += size_t(tiles);
        g.countBlocks += xcdGridSize(it.numGroups, kCountChunk);
        g.scanBlocks += (tiles + kScanTiles - 1) / kScanTiles;
        g.emitBlocks += xcdGridSize(it.numGroups, kEmitChunk);
        g.sortBlocks += tiles;
        const int cells = (it.tilesX + 1) * (it.tilesY + 1);
        g.maxCells = std::max(g.maxCells, cells);
        int waves = kBinWaves;
        while (waves > 1 && waves * cells > kBinLdsCells)
            waves >>= 1;
        g.emitWaves = std::min(g.emitWaves, waves);
        mostTiles = std::max(mostTiles, tiles);
    }
    U3D_CHECK(mostTiles < (1 << 24), U3D_ERR_CAPACITY, "too many tiles in one image");
    sortMap->clear();
    sortMap->reserve(size_t(g.sortBlocks));
    for (int k = 0; k < mostTiles; ++k)
        for (size_t v = 0; v < items.size(); ++v)
            if (k < items[v].tilesX * items[v].tilesY)
                sortMap->push_back(uint32_t(v << 24) | uint32_t(k));
    U3D_CHECK(g.counts < (size_t(1) << 32) && g.totals < (size_t(1) << 32), U3D_ERR_CAPACITY,
              "binning counters of the batch exceed 2^32 entries");
    *grid = g;
    return U3D_OK;
}

int launchTileBinningItems(hipStream_t s, const BinItem *items, int numItems,
                           const BinItemsGrid &grid, const uint32_t *nTiles, const int4 *rect,
                           uint32_t *counts, uint32_t *tileTotal, int *tileOffsets,
                           uint32_t *tileList, size_t capacity, uint32_t *listSum,
                           uint32_t *listMax, uint32_t *overflow, int *longFirst,
                           const uint32_t *sorted, uint32_t *ntS, int4 *rectS, int64_t idLimit)
{
    if (numItems <= 0)
        return U3D_OK;
    U3D_CHECK(!sorted || (ntS && rectS), U3D_ERR_INVALID, "depth-order binning needs ntS / rectS");
    if (grid.countBlocks > 0) {
        hipLaunchKernelGGL(tileCountItemsKernel, dim3(grid.countBlocks), dim3(64 * kBinWaves),
                           size_t(grid.maxCells) * sizeof(uint32_t), s, items, numItems, sorted,
                           nTiles, rect, ntS, rectS, idLimit, counts);
        U3D_LAUNCH_CHECK();
    }
    if (grid.scanBlocks > 0) {
        hipLaunchKernelGGL(tileScanItemsKernel, dim3(grid.scanBlocks), dim3(64 * kScanWaves), 0, s,
                           items, numItems, counts, tileTotal);
        U3D_LAUNCH_CHECK();
    }
    if (grid.emitBlocks > 0) {
        hipLaunchKernelGGL(tileEmitItemsKernel, dim3(grid.emitBlocks), dim3(64 * grid.emitWaves),
                           size_t(grid.emitWaves) * size_t(grid.maxCells) * sizeof(uint32_t), s,
                           items, numItems, sorted, sorted ? ntS : nTiles,
                           sorted ? rectS : rect, counts, tileTotal, tileOffsets, tileList,
                           uint32_t(capacity < size_t(INT32_MAX) ? capacity : size_t(INT32_MAX)),
                           listSum, listMax, overflow, longFirst);
        U3D_LAUNCH_CHECK();
    }
    return U3D_OK;
}

int launchTileSortItems(hipStream_t s, const BinItem *items, const uint32_t *sortMap,
                        const BinItemsGrid &grid, const int *tileOffsets, const int *longFirst,
                        const uint32_t *listIn, uint32_t *listOut, const uint32_t *depthKey,
                        uint32_t *scratch, size_t scratchCap, size_t listCapacity)
{
    if (grid.sortBlocks <= 0)
        return U3D_OK;
    U3D_CHECK(scratchCap >= tileSortScratchSize(listCapacity), U3D_ERR_CAPACITY,
              "tile sort scratch too small");
    uint32_t *gk0 = scratch, *gv0 = scratch + listCapacity;
    uint32_t *gk1 = scratch + 2 * listCapacity, *gv1 = scratch + 3 * listCapacity;
    hipLaunchKernelGGL(tileSortItemsKernel, dim3(grid.sortBlocks), dim3(kTsThreads), 0, s, items,
                       sortMap, tileOffsets, longFirst, listIn, listOut, depthKey, gk0, gv0, gk1,
                       gv1);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchEmitPairs(hipStream_t s, int64_t n, const uint32_t *sorted, const uint32_t *offsets,
                    const int4 *rect, int tilesX, uint32_t *pairTile, uint32_t *pairGauss,
                    const uint32_t *nTilesSorted, int64_t capacity)
{
    if (n <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(emitPairsKernel, dim3(blocksFor(n, 256)), dim3(256), 0, s, n, sorted,
                       offsets, rect, tilesX, pairTile, pairGauss, nTilesSorted, capacity);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchEmitPairsBatched(hipStream_t s, int64_t n, const uint32_t *sorted,
                           const uint32_t *offsets, const int4 *rect, const int *entryItem,
                           const RenderItem *items, uint32_t *pairTile, uint32_t *pairGauss,
                           const uint32_t *nTilesSorted, int64_t capacity, unsigned *overflow)
{
    if (n <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(emitPairsBatchedKernel, dim3(blocksFor(n, 256)), dim3(256), 0, s, n, sorted,
                       offsets, rect, entryItem, items, pairTile, pairGauss, nTilesSorted, capacity,
                       overflow);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchBlockCosts(hipStream_t s, const RenderItem *items, int numItems, int64_t totalBlocks,
                     int tileSize, const int *tileOffsets, uint32_t *keys, uint32_t *order)
{
    if (totalBlocks <= 0)
        return U3D_OK;
    hipLaunchKernelGGL(blockCostKernel, dim3(blocksFor(totalBlocks, 256)), dim3(256), 0, s, items,
                       numItems, totalBlocks, tileSize, tileOffsets, keys, order);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

int launchTileOffsets(hipStream_t s, const uint32_t *pairTile, int64_t numPairs, int tiles,
                      int *tileOffsets, const uint32_t *countDev)
{
    hipLaunchKernelGGL(tileOffsetsKernel, dim3(blocksFor(numPairs + 1, 256)), dim3(256), 0, s,
                       pairTile, numPairs, tiles, tileOffsets, countDev);
    U3D_LAUNCH_CHECK();
    return U3D_OK;
}

} // namespace u3d
