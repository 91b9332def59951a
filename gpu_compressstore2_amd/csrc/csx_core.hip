// csx_core.hip -- launchers of the width-independent kernels: popcount,
// look-back init, chunk-count scan, up-sweep tree, validation.
#include "csx_internal.hpp"
#include "csx_kernels.hpp"

namespace csx {

// ======================================================================
// single-pass scans: shared init kernel
// ======================================================================

// Zero the ticket, the error word and `tiles` status words.  A plain kernel,
// not a memset, so that streams captured into graphs replay it faithfully.
__global__ void k_lookback_init(WsHeader* hdr, uint64_t* status, uint64_t tiles)
{
    const uint64_t tid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (tid == 0) {
        hdr->ticket = 0;
        hdr->error = 0;
    }
    for (uint64_t i = tid; i < tiles; i += (uint64_t)gridDim.x * blockDim.x) status[i] = 0;
}

// Write a 64-bit value from the device (graph-safe replacement for a memset
// or a host copy of a single word).
__global__ void k_store_u64(uint64_t* dst, uint64_t v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = v;
}

// ======================================================================
// the reference's up-sweep tree and its tree walk
// ======================================================================

// In place: pss[j] (a count on entry) becomes the sum of counts over
// (j - lowbit(j+1), j], given the exclusive offsets `ex` of those counts.
// Equals the state left by the log2(chunks) up-sweep launches of
// src/kernels/kernel_3pass.cuh:69-153.
__global__ void k_tree_from_offsets(uint32_t* __restrict__ pss, const uint64_t* __restrict__ ex, uint64_t m)
{
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < m; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t low = (j + 1) & (~(j + 1) + 1);
        const uint64_t inclusive = ex[j] + pss[j];
        pss[j] = (uint32_t)(inclusive - ex[j + 1 - low]);
    }
}

// Exclusive offset of every chunk from the tree: the descent of
// d_3pass_pproc_pssidx (src/kernels/kernel_3pass.cuh:155-169).
__global__ void k_tree_offsets(const uint32_t* __restrict__ tree, uint32_t* __restrict__ out, uint64_t m, uint64_t p2)
{
    for (uint64_t idx = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; idx < m; idx += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t half = p2 / 2, consumed = 0;
        uint32_t acc = 0;
        while (half >= 1) {
            if (idx >= consumed + half) {
                acc += tree[consumed + half - 1];
                consumed += half;
            }
            half /= 2;
        }
        out[idx] = acc;
    }
}

static unsigned grid_for(uint64_t items, uint64_t per_block, uint64_t cap)
{
    uint64_t g = (items + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

hipError_t launch_popc(const uint8_t* mask, uint32_t* counts, uint64_t n, uint32_t log2c, int order, hipStream_t s, WsHeader* hdr, uint64_t* status,
                       uint64_t init_tiles)
{
    const uint64_t nbytes = (n + 7) / 8;
    const uint64_t nvec = (nbytes + 15) / 16;
    const uint64_t chunks = (n + (1ull << log2c) - 1) >> log2c;
    const bool aligned = ((uintptr_t)mask & 15) == 0;
    unsigned grid;
    if (log2c <= 13) {
        grid = grid_for(nvec, kPopcBlock, 1u << 20);
    }
    else {
        grid = grid_for(chunks, kPopcBlock / kWave, 1u << 20);
    }
    if (order == kBitMsb) {
        if (aligned)
            k_popc<kBitMsb, true><<<grid, kPopcBlock, 0, s>>>(mask, counts, n, log2c, hdr, status, init_tiles);
        else
            k_popc<kBitMsb, false><<<grid, kPopcBlock, 0, s>>>(mask, counts, n, log2c, hdr, status, init_tiles);
    }
    else {
        if (aligned)
            k_popc<kBitLsb, true><<<grid, kPopcBlock, 0, s>>>(mask, counts, n, log2c, hdr, status, init_tiles);
        else
            k_popc<kBitLsb, false><<<grid, kPopcBlock, 0, s>>>(mask, counts, n, log2c, hdr, status, init_tiles);
    }
    return hipGetLastError();
}

hipError_t launch_popc_blocks(const uint8_t* mask, uint64_t n, uint32_t log2b, uint32_t log2u, uint32_t* uoffs, uint32_t* btotals, int order,
                              hipStream_t s, WsHeader* hdr)
{
    const uint64_t blocks = (n + (1ull << log2b) - 1) >> log2b;
    if (blocks == 0) return hipSuccess;
    const bool aligned = ((uintptr_t)mask & 15) == 0;
    const unsigned g = (unsigned)blocks;
    // workgroups of 256, 512 or 1024 threads: a block of 2^17, 2^18 or at
    // least 2^19 elements is one round
    const uint32_t grow = log2b - kPopbMinLog2 < 2 ? log2b - kPopbMinLog2 : 2;
    const unsigned threads = (unsigned)kPopbThreads << grow;
    using K = void (*)(const uint8_t*, uint64_t, uint32_t, uint32_t, uint32_t*, uint32_t*, WsHeader*);
    static const K kernels[2][2][3] = {
        {{k_popc_blocks<kBitLsb, false, 256>, k_popc_blocks<kBitLsb, false, 512>, k_popc_blocks<kBitLsb, false, 1024>},
         {k_popc_blocks<kBitLsb, true, 256>, k_popc_blocks<kBitLsb, true, 512>, k_popc_blocks<kBitLsb, true, 1024>}},
        {{k_popc_blocks<kBitMsb, false, 256>, k_popc_blocks<kBitMsb, false, 512>, k_popc_blocks<kBitMsb, false, 1024>},
         {k_popc_blocks<kBitMsb, true, 256>, k_popc_blocks<kBitMsb, true, 512>, k_popc_blocks<kBitMsb, true, 1024>}}};
    kernels[order == kBitMsb][aligned][grow]<<<g, threads, 0, s>>>(mask, n, log2b, log2u, uoffs, btotals, hdr);
    return hipGetLastError();
}

// K2: bprefix[b] = sum of btotals[0..b) for b < nb, *count = the sum of all
// (nullable).  One workgroup; thread t owns the contiguous run of `per`
// totals from t * per, all its loads issued together.  The column's total is
// below 2^32 unless WIDE, so the scan runs on 32-bit DPP wave scans with the
// 16 wave sums through LDS; WIDE scans 64-bit sums.
constexpr int kBlockPrefixThreads = 1024;
constexpr int kBlockPrefixMaxPer = (int)(kPopbMaxBlocks / kBlockPrefixThreads);
template <bool WIDE>
__global__ __launch_bounds__(kBlockPrefixThreads) void k_block_prefix(const uint32_t* __restrict__ btotals, uint32_t nb,
                                                                      uint64_t* __restrict__ bprefix, uint64_t* __restrict__ count)
{
    const uint32_t t = threadIdx.x;
    const uint32_t per = (nb + kBlockPrefixThreads - 1) / kBlockPrefixThreads;
    const uint32_t b0 = t * per;
    uint32_t v[kBlockPrefixMaxPer];
#pragma unroll
    for (int k = 0; k < kBlockPrefixMaxPer; k++) v[k] = 0;
#pragma unroll
    for (int k = 0; k < kBlockPrefixMaxPer; k++)
        if ((uint32_t)k < per && b0 + k < nb) v[k] = btotals[b0 + k];
    uint64_t run, total;
    if constexpr (WIDE) {
        __shared__ uint64_t scr[kBlockPrefixThreads / kWave];
        uint64_t s = 0;
#pragma unroll
        for (int k = 0; k < kBlockPrefixMaxPer; k++) s += v[k];
        run = block_exclusive_sum<kBlockPrefixThreads>(s, scr, &total);
    }
    else {
        __shared__ uint32_t wsum[kBlockPrefixThreads / kWave];
        uint32_t s = 0;
#pragma unroll
        for (int k = 0; k < kBlockPrefixMaxPer; k++) s += v[k];
        const uint32_t inc = wave_inclusive_sum_u32(s);
        if (lane_id() == kWave - 1) wsum[t / kWave] = inc;
        __syncthreads();
        uint32_t before = 0, all = 0;
#pragma unroll
        for (int k = 0; k < kBlockPrefixThreads / kWave; k++) {
            const uint32_t x = wsum[k];
            before += (uint32_t)k < t / kWave ? x : 0u;
            all += x;
        }
        run = before + inc - s;
        total = all;
    }
#pragma unroll
    for (int k = 0; k < kBlockPrefixMaxPer; k++) {
        if ((uint32_t)k < per && b0 + k < nb) bprefix[b0 + k] = run;
        run += v[k];
    }
    if (t == 0 && count) *count = total;
}

hipError_t launch_block_prefix(const uint32_t* btotals, uint64_t blocks, uint64_t* bprefix, uint64_t* count, bool wide, hipStream_t s)
{
    if (blocks == 0 || blocks > kPopbMaxBlocks) return hipErrorInvalidValue;
    if (wide)
        k_block_prefix<true><<<1, kBlockPrefixThreads, 0, s>>>(btotals, (uint32_t)blocks, bprefix, count);
    else
        k_block_prefix<false><<<1, kBlockPrefixThreads, 0, s>>>(btotals, (uint32_t)blocks, bprefix, count);
    return hipGetLastError();
}

hipError_t launch_lookback_init(WsHeader* hdr, uint64_t* status, uint64_t tiles, hipStream_t s)
{
    k_lookback_init<<<grid_for(tiles, 256, 1024), 256, 0, s>>>(hdr, status, tiles);
    return hipGetLastError();
}

hipError_t launch_scan(const uint32_t* counts, uint64_t* offsets, uint64_t m, uint64_t* total, WsHeader* hdr, uint64_t* status, hipStream_t s,
                       bool initialized)
{
    const uint64_t tiles = (m + kScanTile - 1) / kScanTile;
    if (!initialized) {
        hipError_t e = launch_lookback_init(hdr, status, tiles, s);
        if (e != hipSuccess) return e;
    }
    k_scan<uint64_t><<<(unsigned)tiles, kScanBlock, 0, s>>>(counts, offsets, m, total, hdr, status);
    return hipGetLastError();
}

hipError_t launch_scan32(const uint32_t* counts, uint32_t* offsets, uint64_t m, uint64_t* total, WsHeader* hdr, uint64_t* status, hipStream_t s)
{
    const uint64_t tiles = (m + kScanTile - 1) / kScanTile;
    hipError_t e = launch_lookback_init(hdr, status, tiles, s);
    if (e != hipSuccess) return e;
    k_scan<uint32_t><<<(unsigned)tiles, kScanBlock, 0, s>>>(counts, offsets, m, total, hdr, status);
    return hipGetLastError();
}

hipError_t launch_tree_from_offsets(uint32_t* pss, const uint64_t* ex, uint64_t m, hipStream_t s)
{
    k_tree_from_offsets<<<grid_for(m, 256, 1u << 16), 256, 0, s>>>(pss, ex, m);
    return hipGetLastError();
}

hipError_t launch_tree_offsets(const uint32_t* tree, uint32_t* out, uint64_t m, hipStream_t s)
{
    uint64_t p2 = 1;
    while (p2 < m) p2 *= 2;
    k_tree_offsets<<<grid_for(m, 256, 1u << 16), 256, 0, s>>>(tree, out, m, p2);
    return hipGetLastError();
}

hipError_t launch_store_u64(uint64_t* dst, uint64_t v, hipStream_t s)
{
    k_store_u64<<<1, 64, 0, s>>>(dst, v);
    return hipGetLastError();
}

hipError_t launch_validate(const void* a, const void* b, uint64_t len, uint32_t w, uint64_t* failures, hipStream_t s)
{
    const unsigned grid = grid_for(len, 256, 4096);
    auto pa = static_cast<const uint8_t*>(a);
    auto pb = static_cast<const uint8_t*>(b);
    switch (w) {
        case 1: k_validate<1><<<grid, 256, 0, s>>>(pa, pb, len, failures); break;
        case 2: k_validate<2><<<grid, 256, 0, s>>>(pa, pb, len, failures); break;
        case 4: k_validate<4><<<grid, 256, 0, s>>>(pa, pb, len, failures); break;
        default: k_validate<8><<<grid, 256, 0, s>>>(pa, pb, len, failures); break;
    }
    return hipGetLastError();
}

} // namespace csx

namespace csx {

// Plain device copy, 16 bytes per lane, grid-stride: the calibration kernel
// for traffic counters and the bandwidth reference of the benchmarks.
static __global__ __launch_bounds__(256) void k_copy(const u32x4* __restrict__ src, u32x4* __restrict__ dst, uint64_t nv, const uint8_t* __restrict__ src_tail,
                                                     uint8_t* __restrict__ dst_tail, uint32_t tail)
{
    const uint64_t tid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    for (uint64_t i = tid; i < nv; i += (uint64_t)gridDim.x * blockDim.x) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
    if (tid < tail) dst_tail[tid] = src_tail[tid];
}

hipError_t launch_copy(void* dst, const void* src, uint64_t bytes, hipStream_t s)
{
    const uint64_t nv = bytes / 16;
    const uint32_t tail = (uint32_t)(bytes % 16);
    // one vector per thread over a grid that covers the copy once (the
    // fastest plain-copy shape on this HBM); past 2^30 vectors threads loop
    launch_k(k_copy, dim3(grid_for(nv, 256, 1u << 22)), dim3(256), s, static_cast<const u32x4*>(src), static_cast<u32x4*>(dst), nv,
             static_cast<const uint8_t*>(src) + nv * 16, static_cast<uint8_t*>(dst) + nv * 16, tail);
    return hipGetLastError();
}

// Streaming probes: what HBM sustains for plain access mixes with the
// writeout's own shape -- a grid of one block per 256 * U 16-byte vectors
// (no loop), non-temporal hints, blocks optionally in XCD runs (xcd_run > 0:
// the order of the writeout tiles).  Kind 0 reads (one word of the reads'
// xor is stored only if it matches a constant), 1 writes, 2 copies, 3 reads
// two vectors per vector written (the 2:1 byte mix of a 50 % compaction).
static __device__ __forceinline__ uint64_t probe_block(uint32_t xcd_run)
{
    const uint64_t b = blockIdx.x, n = gridDim.x;
    if (!xcd_run) return b;
    const uint64_t grp = kXcds * xcd_run, g = b / grp;
    if ((g + 1) * grp > n) return b;
    return g * grp + (b % kXcds) * xcd_run + (b / kXcds) % xcd_run;
}

template <int KIND, int U>
static __global__ __launch_bounds__(256) void k_stream_probe(const u32x4* __restrict__ src, u32x4* __restrict__ dst, uint64_t nv, uint32_t xcd_run)
{
    const uint64_t base = probe_block(xcd_run) * 256 * U + threadIdx.x;
    if constexpr (KIND == 0) {
        u32x4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < U; j++) {
            const uint64_t i = base + (uint64_t)j * 256;
            if (i < nv) acc ^= __builtin_nontemporal_load(src + i);
        }
        if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x9E3779B9u) dst[0] = acc;
    }
    else if constexpr (KIND == 1) {
#pragma unroll
        for (int j = 0; j < U; j++) {
            const uint64_t i = base + (uint64_t)j * 256;
            if (i < nv) __builtin_nontemporal_store(u32x4{(uint32_t)i, 1u, 2u, 3u}, dst + i);
        }
    }
    else if constexpr (KIND == 2) {
        u32x4 v[U];
#pragma unroll
        for (int j = 0; j < U; j++) {
            const uint64_t i = base + (uint64_t)j * 256;
            v[j] = i < nv ? __builtin_nontemporal_load(src + i) : u32x4{0, 0, 0, 0};
        }
#pragma unroll
        for (int j = 0; j < U; j++) {
            const uint64_t i = base + (uint64_t)j * 256;
            if (i < nv) __builtin_nontemporal_store(v[j], dst + i);
        }
    }
    else {
        // output vector o <- input vectors 2 o and 2 o + 1 (nv counts outputs)
        u32x4 a[U], c[U];
#pragma unroll
        for (int j = 0; j < U; j++) {
            const uint64_t o = base + (uint64_t)j * 256;
            a[j] = o < nv ? __builtin_nontemporal_load(src + 2 * o) : u32x4{0, 0, 0, 0};
            c[j] = o < nv ? __builtin_nontemporal_load(src + 2 * o + 1) : u32x4{0, 0, 0, 0};
        }
#pragma unroll
        for (int j = 0; j < U; j++) {
            const uint64_t o = base + (uint64_t)j * 256;
            if (o < nv) __builtin_nontemporal_store(a[j] ^ c[j], dst + o);
        }
    }
}

template <int KIND> static void launch_probe_kind(const u32x4* src, u32x4* dst, uint64_t nv, uint32_t vpt, uint32_t xcd_run, hipStream_t s)
{
    const uint64_t per = 256ull * vpt;
    const dim3 grid((unsigned)((nv + per - 1) / per));
    if (vpt == 4)
        launch_k(k_stream_probe<KIND, 4>, grid, dim3(256), s, src, dst, nv, xcd_run);
    else
        launch_k(k_stream_probe<KIND, 1>, grid, dim3(256), s, src, dst, nv, xcd_run);
}

hipError_t launch_stream_probe(int kind, const void* src, void* dst, uint64_t bytes, uint32_t vpt, uint32_t xcd_run, hipStream_t s)
{
    auto in = static_cast<const u32x4*>(src);
    auto out = static_cast<u32x4*>(dst);
    const uint64_t nv = bytes / 16;
    switch (kind) {
        case 0: launch_probe_kind<0>(in, out, nv, vpt, xcd_run, s); break;
        case 1: launch_probe_kind<1>(in, out, nv, vpt, xcd_run, s); break;
        case 2: launch_probe_kind<2>(in, out, nv, vpt, xcd_run, s); break;
        default: launch_probe_kind<3>(in, out, nv / 2, vpt, xcd_run, s); break;
    }
    return hipGetLastError();
}

} // namespace csx

namespace csx {

// base[i] = base[i-1] + total[i-1] (base[0] = 0): the carry of the multi-stream
// split, one thread, after the previous slice's scan and carry.  The
// reference's carry (src/kernels/kernel_streaming_add.cuh:8-11) races with a
// read-modify-write on the same word from another stream; here every word has
// one writer and the ordering comes from stream events.
static __global__ void k_carry(uint64_t* dst, const uint64_t* base, const uint64_t* total, uint32_t i)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = i == 0 ? 0ull : base[i - 1] + total[i - 1];
}

hipError_t launch_carry(uint64_t* dst, const uint64_t* base, const uint64_t* total, uint32_t i, hipStream_t s)
{
    k_carry<<<1, 64, 0, s>>>(dst, base, total, i);
    return hipGetLastError();
}

} // namespace csx
