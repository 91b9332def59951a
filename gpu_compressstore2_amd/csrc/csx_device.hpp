// csx_device.hpp -- device-side building blocks shared by the gfx950 kernels.
//
// Mask convention inside every kernel: a mask word is 32 consecutive
// elements with element k at bit k ("LSB-first" word).  Loads convert from
// the caller's byte bit order once, right after the load:
//   MSB-first (the reference's order, src/cub_wraps.cuh:20-24): element i is
//   bit 7-(i%8) of byte i/8, so each byte is bit-reversed in place.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "csx_types.hpp"

namespace csx {

constexpr int kWave = 64;

// native vector types (the HIP uint4 struct is not accepted by the
// non-temporal builtins)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
constexpr int kBitMsb = 0;
constexpr int kBitLsb = 1;

// Status word of the single-pass (decoupled look-back) scans:
//   bits 62..63 flag (0 = not ready, 1 = tile aggregate, 2 = inclusive prefix)
//   bits  0..61 value
constexpr uint64_t kFlagShift = 62;
constexpr uint64_t kFlagAggregate = 1ull << kFlagShift;
constexpr uint64_t kFlagPrefix = 2ull << kFlagShift;
constexpr uint64_t kValueMask = (1ull << kFlagShift) - 1;
// A look-back that has waited this many polls sets the error word and gives
// up; predecessors always make progress, so only a broken launch reaches it.
constexpr uint32_t kSpinLimit = 1u << 26;

template <int ORDER> __device__ __forceinline__ uint32_t to_lsb_word(uint32_t w)
{
    if constexpr (ORDER == kBitLsb) {
        return w;
    }
    else {
        // reverse all 32 bits, then the bytes: each byte ends up bit-reversed in place
        return __builtin_bswap32(__builtin_bitreverse32(w));
    }
}

__device__ __forceinline__ uint32_t popc32(uint32_t w) { return (uint32_t)__builtin_popcount(w); }

// keep the low `bits` bits (0..32)
__device__ __forceinline__ uint32_t low_bits(uint32_t w, uint32_t bits)
{
    return bits >= 32 ? w : (w & ((1u << bits) - 1u));
}

// ---- mask loads --------------------------------------------------------

// 4 consecutive mask words (16 bytes) starting at byte `off`; bytes at or
// past `nbytes` read as zero.  Fast path: one 16-byte load.
template <bool ALIGNED>
__device__ __forceinline__ u32x4 load_mask16(const uint8_t* __restrict__ mask, uint64_t off, uint64_t nbytes)
{
    if (ALIGNED && off + 16 <= nbytes) {
        return __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(mask + off));
    }
    uint32_t w[4] = {0, 0, 0, 0};
    if (off + 16 <= nbytes && (off & 3) == 0 && ((uintptr_t)mask & 3) == 0) {
#pragma unroll
        for (int k = 0; k < 4; k++) w[k] = *reinterpret_cast<const uint32_t*>(mask + off + 4 * k);
    }
    else {
#pragma unroll
        for (int b = 0; b < 16; b++) {
            if (off + b < nbytes) w[b >> 2] |= (uint32_t)mask[off + b] << (8 * (b & 3));
        }
    }
    u32x4 r;
    r.x = w[0];
    r.y = w[1];
    r.z = w[2];
    r.w = w[3];
    return r;
}

// one mask word at byte `off` (bytes past nbytes read as zero)
__device__ __forceinline__ uint32_t load_mask4(const uint8_t* __restrict__ mask, uint64_t off, uint64_t nbytes)
{
    if (off + 4 <= nbytes && ((uintptr_t)(mask + off) & 3) == 0) {
        return *reinterpret_cast<const uint32_t*>(mask + off);
    }
    uint32_t w = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
        if (off + b < nbytes) w |= (uint32_t)mask[off + b] << (8 * b);
    }
    return w;
}

// ---- wave / block scans ------------------------------------------------

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// inclusive prefix sum across the 64 lanes of a wave
template <typename V> __device__ __forceinline__ V wave_inclusive_sum(V v)
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        V o = __shfl_up(v, d, kWave);
        if (lane >= (uint32_t)d) v += o;
    }
    return v;
}

// 32-bit inclusive prefix sum across the 64 lanes of a wave from DPP row
// operations (no LDS traffic): a scan within each row of 16 lanes by
// row_shr 1, 2, 4, 8, then row_bcast:15 adds lane 15 of rows 0 and 2 to rows
// 1 and 3, and row_bcast:31 adds lane 31 to rows 2 and 3.  A lane whose DPP
// source is outside its row, or whose row is masked off, adds the `old`
// operand, 0.  All 64 lanes must be active.
__device__ __forceinline__ uint32_t wave_inclusive_sum_u32(uint32_t v)
{
    constexpr int kRowShr = 0x110, kRowBcast15 = 0x142, kRowBcast31 = 0x143;
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr + 1, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr + 2, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr + 4, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr + 8, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowBcast15, 0xa, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowBcast31, 0xc, 0xf, false);
    return v;
}

// the wave's total, in every lane (lane 63 of the inclusive scan)
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)wave_inclusive_sum_u32(v), kWave - 1);
}

template <typename V> __device__ __forceinline__ V wave_sum(V v)
{
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

// Exclusive prefix sum over a block of BLOCK threads; returns the exclusive
// value for this thread and the block total in *total.  `scratch` holds
// BLOCK/64 entries.  Contains __syncthreads.
template <int BLOCK, typename V>
__device__ __forceinline__ V block_exclusive_sum(V v, V* scratch, V* total)
{
    constexpr int NW = BLOCK / kWave;
    const uint32_t lane = lane_id();
    const uint32_t wid = threadIdx.x / kWave;
    V inc = wave_inclusive_sum(v);
    if (lane == kWave - 1) scratch[wid] = inc;
    __syncthreads();
    V before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        V s = scratch[w];
        if ((uint32_t)w < wid) before += s;
        all += s;
    }
    *total = all;
    return before + inc - v;
}

// ---- decoupled look-back -----------------------------------------------

__device__ __forceinline__ void publish_status(uint64_t* status, uint64_t word)
{
    __hip_atomic_store(status, word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t read_status(const uint64_t* status)
{
    return __hip_atomic_load(const_cast<uint64_t*>(status), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Called by all 64 lanes of ONE wave for tile `tile` > 0 whose aggregate has
// already been published.  Returns the exclusive prefix of the tile (sum of
// all earlier tiles).  Each round inspects the 64 closest unresolved
// predecessors; lane i looks at tile (base - i).
__device__ __forceinline__ uint64_t lookback_exclusive(const uint64_t* status, uint64_t tile, uint32_t* err)
{
    const uint32_t lane = lane_id();
    uint64_t prefix = 0;
    int64_t base = (int64_t)tile - 1;
    uint32_t spins = 0;
    while (true) {
        const int64_t me = base - (int64_t)lane;
        uint64_t s = me >= 0 ? read_status(status + me) : kFlagPrefix;  // before tile 0: prefix 0
        // wait until every inspected predecessor has at least its aggregate
        while (__any((s >> kFlagShift) == 0)) {
            if (++spins > kSpinLimit) {
                if (lane == 0) atomicOr(err, 1u);
                return prefix;
            }
            __builtin_amdgcn_s_sleep(1);
            if ((s >> kFlagShift) == 0) s = read_status(status + me);
        }
        const uint64_t is_p = __ballot((s >> kFlagShift) == 2);
        // closest predecessor carrying an inclusive prefix: lowest lane with P
        const uint32_t stop = is_p ? (uint32_t)__builtin_ctzll(is_p) : (uint32_t)kWave;
        const uint64_t contrib = lane <= stop ? (s & kValueMask) : 0;
        prefix += wave_sum(contrib);
        if (is_p) return prefix;
        base -= kWave;
    }
}

// The workgroups of a launch go to the chip's 8 XCDs round-robin (block b to
// XCD b mod 8), and each XCD has its own L2.  Blocks are taken in groups of
// 8 * kXcdRun: within a group, XCD x gets the run of kXcdRun consecutive tiles
// starting at x * kXcdRun, so the output line two neighbouring tiles share is
// mostly completed in one L2 instead of being written back half by each of
// two, while the chip as a whole still sweeps the column front to back.  A
// ragged last group keeps the plain order, so the map is a bijection on [0, n).
constexpr uint64_t kXcds = 8;
constexpr uint64_t kXcdRun = 16;
__device__ __forceinline__ uint64_t xcd_runs(uint64_t b, uint64_t n)
{
    constexpr uint64_t group = kXcds * kXcdRun;
    const uint64_t g = b / group;
    if ((g + 1) * group > n) return b;
    return g * group + (b % kXcds) * kXcdRun + (b / kXcds) % kXcdRun;
}

// ---- misc ---------------------------------------------------------------

// A value that no workgroup of the running launch writes, read through the
// constant address space: at a wave-uniform address it is a scalar load
// (into SGPRs), whatever the kernel does before it.
template <typename T> __device__ __forceinline__ T load_invariant(const T* p)
{
    return *(const __attribute__((address_space(4))) T*)p;
}

// store 32 mask bits (LSB-first word of elements 32*wi ..) in the caller's order,
// without touching bytes at or past nbytes
template <int ORDER> __device__ __forceinline__ void store_mask_word(uint8_t* mask, uint64_t wi, uint32_t word, uint64_t nbytes)
{
    const uint32_t v = to_lsb_word<ORDER>(word);  // the byte bit-reversal is its own inverse
    const uint64_t off = wi * 4;
    if (off + 4 <= nbytes && ((uintptr_t)(mask + off) & 3) == 0) {
        *reinterpret_cast<uint32_t*>(mask + off) = v;
    }
    else {
        for (int b = 0; b < 4; b++)
            if (off + b < nbytes) mask[off + b] = (uint8_t)(v >> (8 * b));
    }
}

template <int W> struct word_t;
template <> struct word_t<1> { using type = uint8_t; };
template <> struct word_t<2> { using type = uint16_t; };
template <> struct word_t<4> { using type = uint32_t; };
template <> struct word_t<8> { using type = uint64_t; };

} // namespace csx
