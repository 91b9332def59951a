// csx_hbm_probe -- what this GPU's HBM sustains for plain streaming kernels.
//
// Four access mixes over 1 GiB, each a grid-stride loop of U independent
// 16-byte accesses per thread (wave loads 1 KiB contiguous) or one access per
// thread with a grid covering the whole buffer, with and without
// non-temporal hints, at several grid sizes:
//   read    1 GiB read, one word written per kernel
//   write   1 GiB written
//   copy    1 GiB read + 1 GiB written (reported as 2 GiB moved)
//   mix2:1  1 GiB read + 0.5 GiB written: the byte mix of a 50 % compaction
// The best "mix2:1" figure is the practical ceiling of the compaction's
// writeout; the published 8 TB/s is the roofline's peak.  Median of 9 runs.
// The one-pass grids also run with blocks taken in XCD runs: the workgroups
// of a launch go to the 8 XCDs round-robin, and block b -> block
// (b / 8R) * 8R + (b mod 8) * R + (b / 8) mod R gives each XCD runs of R
// consecutive blocks (the order of the compaction's writeout).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>
#include <cstdint>
#include <vector>
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("err %s line %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)

// block order: plain (xr = 0) or runs of xr consecutive blocks per XCD
__device__ __forceinline__ uint64_t blk_order(uint32_t xr)
{
    const uint64_t b = blockIdx.x, n = gridDim.x;
    if (!xr) return b;
    const uint64_t grp = 8ull * xr, g = b / grp;
    if ((g + 1) * grp > n) return b;
    return g * grp + (b % 8) * xr + (b / 8) % xr;
}

template <bool NT> __device__ __forceinline__ u32x4 ld(const u32x4* p) { if (NT) return __builtin_nontemporal_load(p); return *p; }
template <bool NT> __device__ __forceinline__ void st(u32x4* p, u32x4 v) { if (NT) __builtin_nontemporal_store(v, p); else *p = v; }

template <int U, bool NT> __global__ __launch_bounds__(256) void k_read(const u32x4* in, uint64_t nvec, uint32_t* sink, uint32_t xr)
{
    u32x4 acc = {0, 0, 0, 0};
    for (uint64_t b = blk_order(xr) * 256 * U + threadIdx.x; b < nvec; b += (uint64_t)gridDim.x * 256 * U) {
        u32x4 v[U];
#pragma unroll
        for (int j = 0; j < U; j++) { const uint64_t i = b + (uint64_t)j * 256; v[j] = i < nvec ? ld<NT>(in + i) : u32x4{0, 0, 0, 0}; }
#pragma unroll
        for (int j = 0; j < U; j++) acc ^= v[j];
    }
    if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x12345678u) sink[0] = 1;
}
template <int U, bool NT> __global__ __launch_bounds__(256) void k_write(u32x4* out, uint64_t nvec, uint32_t xr)
{
    for (uint64_t b = blk_order(xr) * 256 * U + threadIdx.x; b < nvec; b += (uint64_t)gridDim.x * 256 * U)
#pragma unroll
        for (int j = 0; j < U; j++) { const uint64_t i = b + (uint64_t)j * 256; if (i < nvec) st<NT>(out + i, u32x4{(uint32_t)i, 1, 2, 3}); }
}
template <int U, bool NT> __global__ __launch_bounds__(256) void k_copy(const u32x4* in, u32x4* out, uint64_t nvec, uint32_t xr)
{
    for (uint64_t b = blk_order(xr) * 256 * U + threadIdx.x; b < nvec; b += (uint64_t)gridDim.x * 256 * U) {
        u32x4 v[U];
#pragma unroll
        for (int j = 0; j < U; j++) { const uint64_t i = b + (uint64_t)j * 256; v[j] = i < nvec ? ld<NT>(in + i) : u32x4{0, 0, 0, 0}; }
#pragma unroll
        for (int j = 0; j < U; j++) { const uint64_t i = b + (uint64_t)j * 256; if (i < nvec) st<NT>(out + i, v[j]); }
    }
}
// out vector o (o < nvec/2) <- in[2o] ^ in[2o+1]; two reads per write, reads of a wave contiguous
template <int U, bool NT> __global__ __launch_bounds__(256) void k_mix(const u32x4* in, u32x4* out, uint64_t nvec, uint32_t xr)
{
    const uint64_t nout = nvec / 2;
    for (uint64_t b = blk_order(xr) * 256 * U + threadIdx.x; b < nout; b += (uint64_t)gridDim.x * 256 * U) {
        u32x4 a[U], c[U];
#pragma unroll
        for (int j = 0; j < U; j++) {
            const uint64_t o = b + (uint64_t)j * 256;
            // block-contiguous: reads in[2*blk + t] and in[2*blk + 256 + t] for out block blk
            const uint64_t blk = o - threadIdx.x, t = threadIdx.x;
            a[j] = o < nout ? ld<NT>(in + 2 * blk + t) : u32x4{0, 0, 0, 0};
            c[j] = o < nout ? ld<NT>(in + 2 * blk + 256 + t) : u32x4{0, 0, 0, 0};
        }
#pragma unroll
        for (int j = 0; j < U; j++) { const uint64_t o = b + (uint64_t)j * 256; if (o < nout) st<NT>(out + o, a[j] ^ c[j]); }
    }
}

template <typename F> float timeit(F f, hipEvent_t e0, hipEvent_t e1)
{
    f();
    std::vector<float> t;
    for (int r = 0; r < 9; r++) {
        (void)hipEventRecord(e0, 0);
        f();
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
        float ms; (void)hipEventElapsedTime(&ms, e0, e1); t.push_back(ms);
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

template <int U, bool NT> int run(u32x4* in, u32x4* out, uint32_t* sink, uint64_t nvec, hipEvent_t e0, hipEvent_t e1)
{
    const uint64_t bytes = nvec * 16;
    for (unsigned g : {1024u, 2048u, 4096u, 8192u, 16384u, 0u}) {
        unsigned gr = g ? g : (unsigned)((nvec + 256ull * U - 1) / (256ull * U));
        unsigned grm = g ? g : (unsigned)((nvec / 2 + 256ull * U - 1) / (256ull * U));
        // XCD runs only where one block covers one fixed range (the one-pass grids)
        for (uint32_t xr : {0u, 16u, 32u}) {
            if (xr && g) continue;
            float r = timeit([&] { k_read<U, NT><<<gr, 256>>>(in, nvec, sink, xr); }, e0, e1);
            float w = timeit([&] { k_write<U, NT><<<gr, 256>>>(out, nvec, xr); }, e0, e1);
            float c = timeit([&] { k_copy<U, NT><<<gr, 256>>>(in, out, nvec, xr); }, e0, e1);
            float m = timeit([&] { k_mix<U, NT><<<grm, 256>>>(in, out, nvec, xr); }, e0, e1);
            CK(hipGetLastError());
            printf("U=%d NT=%d grid=%6u xcd-runs=%2u  read %7.1f  write %7.1f  copy %7.1f  mix2:1 %7.1f GB/s\n", U, (int)NT, gr, xr, bytes / (r * 1e6),
                   bytes / (w * 1e6), 2 * bytes / (c * 1e6), 1.5 * bytes / (m * 1e6));
            fflush(stdout);
        }
    }
    return 0;
}

int main()
{
    const uint64_t nvec = 1ull << 26;  // 1 GiB
    u32x4 *in, *out;
    uint32_t* sink;
    CK(hipMalloc(&in, nvec * 16));
    CK(hipMalloc(&out, nvec * 16));
    CK(hipMalloc(&sink, 4));
    CK(hipMemset(in, 1, nvec * 16));
    CK(hipMemset(out, 0, nvec * 16));
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    if (run<1, true>(in, out, sink, nvec, e0, e1)) return 1;
    if (run<4, true>(in, out, sink, nvec, e0, e1)) return 1;
    if (run<8, true>(in, out, sink, nvec, e0, e1)) return 1;
    if (run<4, false>(in, out, sink, nvec, e0, e1)) return 1;
    if (run<8, false>(in, out, sink, nvec, e0, e1)) return 1;
    return 0;
}
