// csx_gen.hip -- device-side seeded input generators.
//
// All generators draw from one PCG32 XSH-RR stream (increment 1, state seeded
// by one splitmix64 step; the generator of src/fast_prng.cuh:7-36) and give
// element k the draws at fixed positions of that stream, so every thread
// jumps straight to its position (LCG jump-ahead, F. Brown, "Random Number
// Generation with Arbitrary Stride", 1994) and the device output equals a
// sequential generator's bit for bit:
//   bernoulli   element k <- draw k; selected iff draw < (uint64)(p*UINT32_MAX)
//               (the rule of generate_mask_uniform, src/data_generator.cuh:14-30;
//               seed 42 and MSB-first give the reference's mask exactly)
//   values      element k <- draw k (w <= 4) or draws 2k, 2k+1 (w = 8)
//   zipf blocks block b <- draws b*(B+1) (rank) and b*(B+1)+1+j (element j)
//   runs        element k flips the running bit iff draw k < 2^32/mean_run
//   pattern     no draws: element k <- bit 31-(k % L) of the pattern
//               (generate_mask_pattern, src/data_generator.cuh:113-128)
#include "csx_internal.hpp"
#include "csx_device.hpp"
#include <csx.h>

namespace csx {

constexpr uint64_t kPcgMult = 6364136223846793005ull;

__host__ __device__ inline uint64_t splitmix_seed(uint64_t seed)
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint32_t pcg_next(uint64_t& st)
{
    const uint64_t old = st;
    st = old * kPcgMult + 1ull;
    const uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
}

// The LCG step s -> M*s + 1 taken `delta` times is again affine,
// s -> mult*s + plus; maps of one LCG commute, so a jump applies the maps of
// 2^i steps for the set bits of delta in any order.
struct Affine {
    uint64_t mult, plus;
};

__host__ __device__ constexpr Affine lcg_steps(uint64_t delta)
{
    uint64_t cur_mult = kPcgMult, cur_plus = 1ull, acc_mult = 1ull, acc_plus = 0ull;
    while (delta) {
        if (delta & 1ull) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1ull) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1;
    }
    return Affine{acc_mult, acc_plus};
}

// maps of 2^i steps, i < 64, fixed at compile time
struct JumpTable {
    uint64_t mult[64], plus[64];
    constexpr JumpTable() : mult(), plus()
    {
        uint64_t m = kPcgMult, p = 1ull;
        for (int i = 0; i < 64; i++) {
            mult[i] = m;
            plus[i] = p;
            p = (m + 1ull) * p;
            m = m * m;
        }
    }
};
__constant__ const JumpTable kJump{};

// state after `delta` steps from `st`: one multiply-add per set bit
__device__ inline uint64_t pcg_advance(uint64_t st, uint64_t delta)
{
    for (int i = 0; delta >> i; i++)
        if ((delta >> i) & 1ull) st = kJump.mult[i] * st + kJump.plus[i];
    return st;
}

__device__ __forceinline__ uint32_t valid_bits(uint64_t wi, uint64_t n)
{
    const uint64_t b = wi * 32;
    return b >= n ? 0u : (n - b >= 32 ? 0xFFFFFFFFu : ((1u << (n - b)) - 1u));
}

// ---------------------------------------------------------------- bernoulli
// Lane t of S lanes makes words t, t+S, t+2S, ...: one jump to draw 32t, 32
// sequential steps per word, then `skip` (the map of 32*(S-1) steps) to the
// next word of the lane.  Consecutive lanes store consecutive words.
template <int ORDER>
__global__ __launch_bounds__(256) void k_gen_bernoulli(uint8_t* mask, uint64_t n, uint64_t thr, uint64_t st0, Affine skip)
{
    const uint64_t nwords = (n + 31) / 32, nbytes = (n + 7) / 8;
    const uint64_t wi0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (wi0 >= nwords) return;
    uint64_t st = pcg_advance(st0, wi0 * 32);
    for (uint64_t wi = wi0; wi < nwords; wi += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t word = 0;
        for (int k = 0; k < 32; k++) word |= (uint32_t)((uint64_t)pcg_next(st) < thr) << k;
        store_mask_word<ORDER>(mask, wi, word & valid_bits(wi, n), nbytes);
        st = skip.mult * st + skip.plus;
    }
}

// ---------------------------------------------------------------- pattern
template <int ORDER> __global__ __launch_bounds__(256) void k_gen_pattern(uint8_t* mask, uint64_t n, uint32_t pattern, uint32_t L)
{
    const uint64_t nwords = (n + 31) / 32, nbytes = (n + 7) / 8;
    for (uint64_t wi = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; wi < nwords; wi += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t word = 0;
        uint32_t pos = (uint32_t)((wi * 32) % L);
        for (int k = 0; k < 32; k++) {
            word |= ((pattern >> (31 - pos)) & 1u) << k;
            if (++pos == L) pos = 0;
        }
        store_mask_word<ORDER>(mask, wi, word & valid_bits(wi, n), nbytes);
    }
}

// ---------------------------------------------------------------- zipf blocks
constexpr uint32_t kMaxZipfLevels = 128;
struct ZipfTable {
    uint32_t cdf[kMaxZipfLevels];  // rank r (0-based) <- first r with draw < cdf[r]
    uint64_t thr[kMaxZipfLevels];  // Bernoulli threshold of rank r
    uint32_t levels;
};

template <int ORDER>
__global__ __launch_bounds__(256) void k_gen_zipf(uint8_t* mask, uint64_t n, uint32_t block_len, uint64_t st0, ZipfTable z)
{
    const uint64_t nwords = (n + 31) / 32, nbytes = (n + 7) / 8;
    for (uint64_t wi = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; wi < nwords; wi += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t e0 = wi * 32;
        const uint64_t b = e0 / block_len;
        const uint64_t d0 = b * ((uint64_t)block_len + 1);
        uint64_t st = pcg_advance(st0, d0);
        const uint32_t u = pcg_next(st);
        uint32_t r = 0;
        while (r + 1 < z.levels && u >= z.cdf[r]) r++;
        const uint64_t thr = z.thr[r];
        st = pcg_advance(st, e0 - b * block_len);
        uint32_t word = 0;
        for (int k = 0; k < 32; k++) word |= (uint32_t)((uint64_t)pcg_next(st) < thr) << k;
        store_mask_word<ORDER>(mask, wi, word & valid_bits(wi, n), nbytes);
    }
}

// ---------------------------------------------------------------- runs
// pass 1: flips per word -> popcount
__global__ __launch_bounds__(256) void k_gen_runs_count(uint32_t* counts, uint64_t n, uint32_t thr, uint64_t st0, Affine skip)
{
    const uint64_t nwords = (n + 31) / 32;
    const uint64_t wi0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (wi0 >= nwords) return;
    uint64_t st = pcg_advance(st0, wi0 * 32);
    for (uint64_t wi = wi0; wi < nwords; wi += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t f = 0;
        for (int k = 0; k < 32; k++) f |= (uint32_t)(pcg_next(st) < thr) << k;
        counts[wi] = popc32(f & valid_bits(wi, n));
        st = skip.mult * st + skip.plus;
    }
}

// pass 2: value = parity of all flips up to and including the element
template <int ORDER>
__global__ __launch_bounds__(256) void k_gen_runs_write(uint8_t* mask, const uint64_t* before, uint64_t n, uint32_t thr, uint64_t st0, Affine skip)
{
    const uint64_t nwords = (n + 31) / 32, nbytes = (n + 7) / 8;
    const uint64_t wi0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (wi0 >= nwords) return;
    uint64_t st = pcg_advance(st0, wi0 * 32);
    for (uint64_t wi = wi0; wi < nwords; wi += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t f = 0;
        for (int k = 0; k < 32; k++) f |= (uint32_t)(pcg_next(st) < thr) << k;
        uint32_t x = f;  // inclusive prefix xor along the word
        x ^= x << 1;
        x ^= x << 2;
        x ^= x << 4;
        x ^= x << 8;
        x ^= x << 16;
        if (before[wi] & 1ull) x = ~x;
        store_mask_word<ORDER>(mask, wi, x & valid_bits(wi, n), nbytes);
        st = skip.mult * st + skip.plus;
    }
}

// ---------------------------------------------------------------- values
// A group is 16 bytes of output: kGroup<W> consecutive elements.  Lane t of
// S lanes makes groups t, t+S, ...: one jump, the group's draws in sequence,
// then `skip` (the map of the draws of S-1 groups) to its next group, so each
// draw costs one LCG step and consecutive lanes store consecutive 16 bytes.
template <int W> constexpr uint32_t kGroup = 16 / W;
template <int W> constexpr uint32_t kDraws = W == 8 ? 2 : 1;  // draws per element

template <int W> __device__ __forceinline__ uint64_t make_value(uint64_t& st, int kind)
{
    if constexpr (W == 8) {
        const uint64_t hi = pcg_next(st), lo = pcg_next(st);
        uint64_t v = (hi << 32) | lo;
        if (kind == CSX_VALUES_FLOAT) v = __double_as_longlong((double)(v >> 11) * (1.0 / 9007199254740992.0));
        return v;
    }
    else {
        uint32_t v = pcg_next(st);
        if (W == 4 && kind == CSX_VALUES_FLOAT) v = __float_as_uint((float)(v >> 8) * (1.0f / 16777216.0f));
        return v;
    }
}

template <int W>
__global__ __launch_bounds__(256) void k_gen_values(uint8_t* out, uint64_t n, int kind, uint64_t st0, Affine skip, int aligned)
{
    using T = typename word_t<W>::type;
    constexpr uint32_t G = kGroup<W>;
    const uint64_t groups = (n + G - 1) / G;
    const uint64_t g0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (g0 >= groups) return;
    uint64_t st = pcg_advance(st0, g0 * G * kDraws<W>);
    for (uint64_t g = g0; g < groups; g += (uint64_t)gridDim.x * blockDim.x) {
        union {
            T e[G];
            u32x4 v;
        } pack;
#pragma unroll
        for (uint32_t j = 0; j < G; j++) pack.e[j] = (T)make_value<W>(st, kind);
        const uint64_t e0 = g * G;
        if (aligned && e0 + G <= n) {
            *reinterpret_cast<u32x4*>(out + e0 * W) = pack.v;
        }
        else {
            for (uint32_t j = 0; j < G; j++)
                if (e0 + j < n) reinterpret_cast<T*>(out)[e0 + j] = pack.e[j];
        }
        st = skip.mult * st + skip.plus;
    }
}

// ---------------------------------------------------------------- launchers
// 8 blocks of 256 per CU on 256 CUs fill the chip; lanes beyond that loop
static unsigned gen_grid(uint64_t items)
{
    uint64_t g = (items + 255) / 256;
    if (g < 1) g = 1;
    if (g > 2048) g = 2048;
    return (unsigned)g;
}

// map from the end of one item of `draws` draws to the lane's next item
static Affine lane_skip(unsigned grid, uint64_t draws)
{
    return lcg_steps(((uint64_t)grid * 256 - 1) * draws);
}

hipError_t launch_gen_bernoulli(uint8_t* mask, uint64_t n, uint64_t thr, uint64_t seed, int order, hipStream_t s)
{
    const uint64_t st0 = splitmix_seed(seed);
    const unsigned g = gen_grid((n + 31) / 32);
    const Affine skip = lane_skip(g, 32);
    if (order == kBitMsb)
        k_gen_bernoulli<kBitMsb><<<g, 256, 0, s>>>(mask, n, thr, st0, skip);
    else
        k_gen_bernoulli<kBitLsb><<<g, 256, 0, s>>>(mask, n, thr, st0, skip);
    return hipGetLastError();
}

hipError_t launch_gen_pattern(uint8_t* mask, uint64_t n, uint32_t pattern, uint32_t L, int order, hipStream_t s)
{
    const unsigned g = gen_grid((n + 31) / 32);
    if (order == kBitMsb)
        k_gen_pattern<kBitMsb><<<g, 256, 0, s>>>(mask, n, pattern, L);
    else
        k_gen_pattern<kBitLsb><<<g, 256, 0, s>>>(mask, n, pattern, L);
    return hipGetLastError();
}

hipError_t launch_gen_zipf(uint8_t* mask, uint64_t n, uint32_t block_len, uint64_t seed, const uint32_t* cdf, const uint64_t* thr, uint32_t levels,
                           int order, hipStream_t s)
{
    ZipfTable z{};
    z.levels = levels;
    for (uint32_t i = 0; i < levels; i++) {
        z.cdf[i] = cdf[i];
        z.thr[i] = thr[i];
    }
    const uint64_t st0 = splitmix_seed(seed);
    const unsigned g = gen_grid((n + 31) / 32);
    if (order == kBitMsb)
        k_gen_zipf<kBitMsb><<<g, 256, 0, s>>>(mask, n, block_len, st0, z);
    else
        k_gen_zipf<kBitLsb><<<g, 256, 0, s>>>(mask, n, block_len, st0, z);
    return hipGetLastError();
}

hipError_t launch_gen_runs(uint8_t* mask, uint64_t n, uint32_t thr, uint64_t seed, int order, uint32_t* counts, uint64_t* before, WsHeader* hdr,
                           uint64_t* status, hipStream_t s)
{
    const uint64_t st0 = splitmix_seed(seed);
    const uint64_t nwords = (n + 31) / 32;
    const unsigned g = gen_grid(nwords);
    const Affine skip = lane_skip(g, 32);
    k_gen_runs_count<<<g, 256, 0, s>>>(counts, n, thr, st0, skip);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = launch_scan(counts, before, nwords, nullptr, hdr, status, s);
    if (e != hipSuccess) return e;
    if (order == kBitMsb)
        k_gen_runs_write<kBitMsb><<<g, 256, 0, s>>>(mask, before, n, thr, st0, skip);
    else
        k_gen_runs_write<kBitLsb><<<g, 256, 0, s>>>(mask, before, n, thr, st0, skip);
    return hipGetLastError();
}

hipError_t launch_gen_values(void* out, uint64_t n, uint32_t w, int kind, uint64_t seed, hipStream_t s)
{
    const uint64_t st0 = splitmix_seed(seed);
    auto o = static_cast<uint8_t*>(out);
    const int aligned = ((uintptr_t)out & 15) == 0;
    const uint64_t groups = (n * w + 15) / 16;
    const unsigned g = gen_grid(groups);
    const Affine skip = lane_skip(g, (16 / w) * (w == 8 ? 2 : 1));  // draws per group: 16/w elements, 2 draws each at w = 8
    switch (w) {
        case 1: k_gen_values<1><<<g, 256, 0, s>>>(o, n, kind, st0, skip, aligned); break;
        case 2: k_gen_values<2><<<g, 256, 0, s>>>(o, n, kind, st0, skip, aligned); break;
        case 4: k_gen_values<4><<<g, 256, 0, s>>>(o, n, kind, st0, skip, aligned); break;
        default: k_gen_values<8><<<g, 256, 0, s>>>(o, n, kind, st0, skip, aligned); break;
    }
    return hipGetLastError();
}

} // namespace csx
