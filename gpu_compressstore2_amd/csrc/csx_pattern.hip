// csx_pattern.hip -- compaction by a periodic selection pattern, without a
// mask in memory (the reference's bench9, src/kernels/kernel_pattern.cuh:13-131).
//
// Element i is selected iff bit 31 - (i mod L) of `pattern` is set
// (src/data_generator.cuh:113-128).  With P set bits at positions
// pos[0] < ... < pos[P-1] inside one period, output o comes from input
//   (o / P) * L + pos[o mod P]
// so each lane computes its source index from its output index: stores are
// fully coalesced, loads gather from a window the wave reads densely.
#include "csx_internal.hpp"
#include "csx_device.hpp"

namespace csx {

struct PatternArgs {
    const uint8_t* in;
    uint8_t* out;
    uint64_t k;  // selected count
    uint32_t P;  // set bits per period
    uint32_t L;  // period
    uint8_t pos[32];
};

constexpr int kPatternBlock = 256;
constexpr int kPatternPerThread = 8;

template <int W> __global__ __launch_bounds__(kPatternBlock) void k_pattern(PatternArgs a)
{
    using T = typename word_t<W>::type;
    __shared__ uint32_t s_pos[32];
    if (threadIdx.x < 32) s_pos[threadIdx.x] = a.pos[threadIdx.x];
    __syncthreads();
    const T* __restrict__ in = reinterpret_cast<const T*>(a.in);
    T* __restrict__ out = reinterpret_cast<T*>(a.out);
    const uint64_t span = (uint64_t)kPatternBlock * kPatternPerThread;
    for (uint64_t o0 = blockIdx.x * span; o0 < a.k; o0 += (uint64_t)gridDim.x * span) {
        uint64_t o = o0 + threadIdx.x;
        // period and in-period rank of the first output; later outputs step by 256
        uint64_t q = o / a.P;
        uint32_t r = (uint32_t)(o - q * a.P);
        const uint32_t dq = kPatternBlock / a.P, dr = kPatternBlock % a.P;
        T v[kPatternPerThread];
#pragma unroll
        for (int j = 0; j < kPatternPerThread; j++) {
            if (o + (uint64_t)j * kPatternBlock < a.k) v[j] = in[q * a.L + s_pos[r]];
            q += dq;
            r += dr;
            if (r >= a.P) {
                r -= a.P;
                q++;
            }
        }
#pragma unroll
        for (int j = 0; j < kPatternPerThread; j++) {
            const uint64_t oj = o + (uint64_t)j * kPatternBlock;
            if (oj < a.k) out[oj] = v[j];
        }
    }
}

hipError_t launch_pattern(uint32_t w, const void* in, void* out, uint64_t k, uint32_t pattern, uint32_t L, hipStream_t s)
{
    PatternArgs a{};
    a.in = static_cast<const uint8_t*>(in);
    a.out = static_cast<uint8_t*>(out);
    a.k = k;
    a.L = L;
    a.P = 0;
    for (uint32_t p = 0; p < L; p++)
        if ((pattern >> (31 - p)) & 1u) a.pos[a.P++] = (uint8_t)p;
    if (k == 0 || a.P == 0) return hipSuccess;
    const uint64_t span = (uint64_t)kPatternBlock * kPatternPerThread;
    uint64_t g = (k + span - 1) / span;
    if (g > 65536) g = 65536;
    switch (w) {
        case 1: k_pattern<1><<<(unsigned)g, kPatternBlock, 0, s>>>(a); break;
        case 2: k_pattern<2><<<(unsigned)g, kPatternBlock, 0, s>>>(a); break;
        case 4: k_pattern<4><<<(unsigned)g, kPatternBlock, 0, s>>>(a); break;
        default: k_pattern<8><<<(unsigned)g, kPatternBlock, 0, s>>>(a); break;
    }
    return hipGetLastError();
}

} // namespace csx
