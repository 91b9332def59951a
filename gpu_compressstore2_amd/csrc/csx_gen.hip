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

// state after `delta` steps from `st`
__device__ inline uint64_t pcg_advance(uint64_t st, uint64_t delta)
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
    return acc_mult * st + acc_plus;
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

__device__ __forceinline__ uint32_t valid_bits(uint64_t wi, uint64_t n)
{
    const uint64_t b = wi * 32;
    return b >= n ? 0u : (n - b >= 32 ? 0xFFFFFFFFu : ((1u << (n - b)) - 1u));
}

// ---------------------------------------------------------------- bernoulli
template <int ORDER> __global__ __launch_bounds__(256) void k_gen_bernoulli(uint8_t* mask, uint64_t n, uint64_t thr, uint64_t st0)
{
    const uint64_t nwords = (n + 31) / 32, nbytes = (n + 7) / 8;
    for (uint64_t wi = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; wi < nwords; wi += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t st = pcg_advance(st0, wi * 32);
        uint32_t word = 0;
        for (int k = 0; k < 32; k++) word |= (uint32_t)((uint64_t)pcg_next(st) < thr) << k;
        store_mask_word<ORDER>(mask, wi, word & valid_bits(wi, n), nbytes);
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
__global__ __launch_bounds__(256) void k_gen_runs_count(uint32_t* counts, uint64_t n, uint32_t thr, uint64_t st0)
{
    const uint64_t nwords = (n + 31) / 32;
    for (uint64_t wi = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; wi < nwords; wi += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t st = pcg_advance(st0, wi * 32);
        uint32_t f = 0;
        for (int k = 0; k < 32; k++) f |= (uint32_t)(pcg_next(st) < thr) << k;
        counts[wi] = popc32(f & valid_bits(wi, n));
    }
}

// pass 2: value = parity of all flips up to and including the element
template <int ORDER>
__global__ __launch_bounds__(256) void k_gen_runs_write(uint8_t* mask, const uint64_t* before, uint64_t n, uint32_t thr, uint64_t st0)
{
    const uint64_t nwords = (n + 31) / 32, nbytes = (n + 7) / 8;
    for (uint64_t wi = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; wi < nwords; wi += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t st = pcg_advance(st0, wi * 32);
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
    }
}

// ---------------------------------------------------------------- values
// each thread produces a run of kValueRun consecutive elements after one jump
constexpr uint64_t kValueRun = 16;

template <int W>
__global__ __launch_bounds__(256) void k_gen_values(uint8_t* out, uint64_t n, int kind, uint64_t st0)
{
    const uint64_t runs = (n + kValueRun - 1) / kValueRun;
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < runs; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t e0 = r * kValueRun;
        const uint64_t e1 = e0 + kValueRun < n ? e0 + kValueRun : n;
        uint64_t st = pcg_advance(st0, (W == 8 ? 2 : 1) * e0);
        for (uint64_t e = e0; e < e1; e++) {
            if constexpr (W == 8) {
                const uint64_t hi = pcg_next(st), lo = pcg_next(st);
                uint64_t v = (hi << 32) | lo;
                if (kind == 1) {
                    const double d = (double)(v >> 11) * (1.0 / 9007199254740992.0);
                    v = __double_as_longlong(d);
                }
                reinterpret_cast<uint64_t*>(out)[e] = v;
            }
            else {
                uint32_t v = pcg_next(st);
                if (W == 4 && kind == 1) v = __float_as_uint((float)(v >> 8) * (1.0f / 16777216.0f));
                if constexpr (W == 4) reinterpret_cast<uint32_t*>(out)[e] = v;
                if constexpr (W == 2) reinterpret_cast<uint16_t*>(out)[e] = (uint16_t)v;
                if constexpr (W == 1) out[e] = (uint8_t)v;
            }
        }
    }
}

// ---------------------------------------------------------------- launchers
static unsigned gen_grid(uint64_t items)
{
    uint64_t g = (items + 255) / 256;
    if (g < 1) g = 1;
    if (g > 16384) g = 16384;
    return (unsigned)g;
}

hipError_t launch_gen_bernoulli(uint8_t* mask, uint64_t n, uint64_t thr, uint64_t seed, int order, hipStream_t s)
{
    const uint64_t st0 = splitmix_seed(seed);
    const unsigned g = gen_grid((n + 31) / 32);
    if (order == kBitMsb)
        k_gen_bernoulli<kBitMsb><<<g, 256, 0, s>>>(mask, n, thr, st0);
    else
        k_gen_bernoulli<kBitLsb><<<g, 256, 0, s>>>(mask, n, thr, st0);
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
    k_gen_runs_count<<<g, 256, 0, s>>>(counts, n, thr, st0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = launch_scan(counts, before, nwords, nullptr, hdr, status, s);
    if (e != hipSuccess) return e;
    if (order == kBitMsb)
        k_gen_runs_write<kBitMsb><<<g, 256, 0, s>>>(mask, before, n, thr, st0);
    else
        k_gen_runs_write<kBitLsb><<<g, 256, 0, s>>>(mask, before, n, thr, st0);
    return hipGetLastError();
}

hipError_t launch_gen_values(void* out, uint64_t n, uint32_t w, int kind, uint64_t seed, hipStream_t s)
{
    const uint64_t st0 = splitmix_seed(seed);
    const unsigned g = gen_grid((n + kValueRun - 1) / kValueRun);
    auto o = static_cast<uint8_t*>(out);
    switch (w) {
        case 1: k_gen_values<1><<<g, 256, 0, s>>>(o, n, kind, st0); break;
        case 2: k_gen_values<2><<<g, 256, 0, s>>>(o, n, kind, st0); break;
        case 4: k_gen_values<4><<<g, 256, 0, s>>>(o, n, kind, st0); break;
        default: k_gen_values<8><<<g, 256, 0, s>>>(o, n, kind, st0); break;
    }
    return hipGetLastError();
}

} // namespace csx
