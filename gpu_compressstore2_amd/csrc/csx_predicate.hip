// csx_predicate.hip -- a column and a comparison give the packed mask.
//
// The device counterpart of gen_predicate (src/utils.cuh:102-122), which packs
// predicate_function(value) (src/main.cu:24-28, `value > threshold`) MSB-first,
// eight elements per byte.  Every comparison is one uniform range test:
//   selected = (no lower bound || v >(=) lo) && (no upper bound || v <(=) hi),
//   xor negate
// so GT/GE/LT/LE/EQ/NE/BETWEEN share one kernel per element type, and IEEE
// comparisons give NaN its usual answers (false, true for NE).
//
// Layout: a wave takes two chunks of 1024 elements per step (grid-stride over
// chunk pairs).  A chunk is W rows of 64 lanes x 16 bytes; both chunks' loads
// (non-temporal: the column is read once) are issued before any use.
// A lane's 16/W elements give 16/W bits; OR-shuffles across the 2W lanes of a
// 32-element word assemble the word, and lane j < 32 collects the chunk's word
// j, so the chunk's 128 mask bytes leave in one coalesced store.  The chunk's
// popcount is the per-chunk count of the three-phase pipeline (chunk length
// 1024), so csx_compact_predicate needs no separate popcount pass.
#include "csx_internal.hpp"
#include "csx_device.hpp"
#include <csx.h>

namespace csx {

constexpr uint32_t kPredChunk = 1024;
constexpr uint32_t kPredPair = 2;  // chunks per wave step
constexpr int kPredBlock = 256;

template <typename T> __device__ __forceinline__ T from_bits(uint64_t b)
{
    T v;
    __builtin_memcpy(&v, &b, sizeof(T));  // low bytes: little-endian
    return v;
}

template <typename T> __device__ __forceinline__ bool pred_eval(T v, T lo, T hi, uint32_t f)
{
    bool ok = true;
    if (f & kPredLo) ok = (f & kPredLoIncl) ? (v >= lo) : (v > lo);
    if (f & kPredHi) ok = ok && ((f & kPredHiIncl) ? (v <= hi) : (v < hi));
    return ok != ((f & kPredNegate) != 0);
}

template <typename T, int ORDER, bool ALIGNED>
__global__ __launch_bounds__(kPredBlock) void k_predicate(const T* __restrict__ in, uint64_t n, uint64_t lo_bits, uint64_t hi_bits, uint32_t flags,
                                                          uint8_t* __restrict__ mask, uint32_t* __restrict__ counts, uint64_t* total, WsHeader* hdr,
                                                          uint64_t* status, uint64_t init_tiles)
{
    constexpr uint32_t W = sizeof(T);
    constexpr uint32_t V = 16 / W;    // elements per lane per step
    constexpr uint32_t L = 32 / V;    // lanes per 32-bit mask word
    constexpr uint32_t WPS = 64 / L;  // words per step
    constexpr uint32_t S = W;         // steps per chunk: 64 * V * S = 1024
    static_assert(64 * V * S == kPredChunk, "chunk shape");

    if (hdr) {  // look-back state of the scan that follows
        const uint64_t tid = blockIdx.x * (uint64_t)kPredBlock + threadIdx.x;
        if (tid == 0) {
            hdr->ticket = 0;
            hdr->error = 0;
        }
        for (uint64_t i = tid; i < init_tiles; i += (uint64_t)gridDim.x * kPredBlock) status[i] = 0;
    }
    const T lo = from_bits<T>(lo_bits), hi = from_bits<T>(hi_bits);
    const uint64_t nbytes = (n + 7) / 8;
    const uint64_t chunks = (n + kPredChunk - 1) / kPredChunk;
    const uint32_t lane = lane_id();
    // blocks in XCD runs (see xcd_runs): measured 4 % faster at 2^28 float32
    const uint64_t wave0 = (xcd_runs(blockIdx.x, gridDim.x) * (uint64_t)kPredBlock + threadIdx.x) / kWave;
    const uint64_t waves = (uint64_t)gridDim.x * (kPredBlock / kWave);
    uint64_t selected = 0;

    for (uint64_t c0 = wave0 * kPredPair; c0 < chunks; c0 += waves * kPredPair) {
        union {
            T e[V];
            u32x4 v;
        } x[kPredPair][S];
#pragma unroll
        for (uint32_t q = 0; q < kPredPair; q++) {
#pragma unroll
            for (uint32_t s = 0; s < S; s++) {
                const uint64_t e = (c0 + q) * kPredChunk + s * (64 * V) + lane * V;
                if (ALIGNED && e + V <= n) {
                    x[q][s].v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(in + e));
                }
                else {
#pragma unroll
                    for (uint32_t j = 0; j < V; j++) x[q][s].e[j] = e + j < n ? in[e + j] : T(0);
                }
            }
        }
#pragma unroll
        for (uint32_t q = 0; q < kPredPair; q++) {
            const uint64_t c = c0 + q;
            if (c >= chunks) break;
            uint32_t word = 0;
#pragma unroll
            for (uint32_t s = 0; s < S; s++) {
                const uint64_t e = c * kPredChunk + s * (64 * V) + lane * V;
                uint32_t bits = 0;
#pragma unroll
                for (uint32_t j = 0; j < V; j++) bits |= (uint32_t)(e + j < n && pred_eval<T>(x[q][s].e[j], lo, hi, flags)) << j;
                // word s*WPS + lane/L of the chunk: lanes of one group hold disjoint bit ranges
                uint32_t w = bits << ((lane % L) * V);
#pragma unroll
                for (uint32_t o = 1; o < L; o <<= 1) w |= __shfl_xor(w, o, kWave);
                // lane j < 32 takes chunk word j when step s makes it (lanes >= 32 never match)
                const uint32_t got = __shfl(w, (lane % WPS) * L, kWave);
                if (lane / WPS == s) word = got;
            }
            if (lane < 32) store_mask_word<ORDER>(mask, c * 32 + lane, word, nbytes);  // bits at index >= n are 0
            const uint32_t cnt = wave_sum<uint32_t>(lane < 32 ? popc32(word) : 0u);
            if (counts && lane == 0) counts[c] = cnt;
            selected += cnt;
        }
    }
    if (total) {
        __shared__ unsigned long long acc;
        if (threadIdx.x == 0) acc = 0;
        __syncthreads();
        if (lane == 0 && selected) atomicAdd(&acc, (unsigned long long)selected);
        __syncthreads();
        if (threadIdx.x == 0 && acc) atomicAdd(reinterpret_cast<unsigned long long*>(total), acc);
    }
}

template <typename T>
static hipError_t launch_pred_t(const void* in, uint64_t n, const PredSpec& p, uint8_t* mask, int order, uint32_t* counts, uint64_t* total, WsHeader* hdr,
                                uint64_t* status, uint64_t init_tiles, hipStream_t s)
{
    const uint64_t chunks = (n + kPredChunk - 1) / kPredChunk;
    const uint64_t per_block = kPredPair * (kPredBlock / kWave);
    uint64_t g = (chunks + per_block - 1) / per_block;
    if (g < 1) g = 1;
    // one step per wave up to 2^28 elements (one-pass grids stream fastest on
    // this HBM); 1-byte columns are ALU-bound and do best with 4 chunks per wave.
    // A running total costs one same-address atomic per block, which the L2
    // serialises (~16 ns each), so counting grids stay at 4096 blocks.
    const uint64_t cap = total ? 4096 : sizeof(T) == 1 ? 8192 : 32768;
    if (g > cap) g = cap;
    const bool aligned = ((uintptr_t)in & 15) == 0;
    auto x = static_cast<const T*>(in);
#define CSX_PRED_LAUNCH(ORD, AL) \
    k_predicate<T, ORD, AL><<<(unsigned)g, kPredBlock, 0, s>>>(x, n, p.lo, p.hi, p.flags, mask, counts, total, hdr, status, init_tiles)
    if (order == kBitMsb) {
        if (aligned)
            CSX_PRED_LAUNCH(kBitMsb, true);
        else
            CSX_PRED_LAUNCH(kBitMsb, false);
    }
    else {
        if (aligned)
            CSX_PRED_LAUNCH(kBitLsb, true);
        else
            CSX_PRED_LAUNCH(kBitLsb, false);
    }
#undef CSX_PRED_LAUNCH
    return hipGetLastError();
}

hipError_t launch_predicate(const void* in, uint64_t n, const PredSpec& p, uint8_t* mask, int order, uint32_t* counts, uint64_t* total, WsHeader* hdr,
                            uint64_t* status, uint64_t init_tiles, hipStream_t s)
{
    switch (p.type) {
        case CSX_TYPE_U8: return launch_pred_t<uint8_t>(in, n, p, mask, order, counts, total, hdr, status, init_tiles, s);
        case CSX_TYPE_U16: return launch_pred_t<uint16_t>(in, n, p, mask, order, counts, total, hdr, status, init_tiles, s);
        case CSX_TYPE_U32: return launch_pred_t<uint32_t>(in, n, p, mask, order, counts, total, hdr, status, init_tiles, s);
        case CSX_TYPE_U64: return launch_pred_t<uint64_t>(in, n, p, mask, order, counts, total, hdr, status, init_tiles, s);
        case CSX_TYPE_I8: return launch_pred_t<int8_t>(in, n, p, mask, order, counts, total, hdr, status, init_tiles, s);
        case CSX_TYPE_I16: return launch_pred_t<int16_t>(in, n, p, mask, order, counts, total, hdr, status, init_tiles, s);
        case CSX_TYPE_I32: return launch_pred_t<int32_t>(in, n, p, mask, order, counts, total, hdr, status, init_tiles, s);
        case CSX_TYPE_I64: return launch_pred_t<int64_t>(in, n, p, mask, order, counts, total, hdr, status, init_tiles, s);
        case CSX_TYPE_F32: return launch_pred_t<float>(in, n, p, mask, order, counts, total, hdr, status, init_tiles, s);
        case CSX_TYPE_F64: return launch_pred_t<double>(in, n, p, mask, order, counts, total, hdr, status, init_tiles, s);
        default: return hipErrorInvalidValue;
    }
}

} // namespace csx
