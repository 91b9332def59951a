// csx_kernels.hpp -- gfx950 kernels of the three-phase bitmask compaction
// pipeline and of the fused single-pass form.
//
// Reference path (RememberOfLife/gpu_compressstore2):
//   phase 1 popcount   src/kernels/kernel_3pass.cuh:12-57
//   phase 2 scan       src/kernels/kernel_3pass.cuh:59-214, CUB ExclusiveSum
//                      (src/benchmarks.cuh:308,374,432)
//   phase 3 writeout   src/kernels/kernel_3pass.cuh:216-693
// The reference works in 32-lane warps over 1024-element strides; these
// kernels are laid out for 64-lane waves, 16-byte loads and LDS staging.
#pragma once

#include "csx_device.hpp"

namespace csx {

// ======================================================================
// phase 1: per-chunk popcount
// ======================================================================
//
// Each thread owns one 16-byte mask vector (128 elements).  For a chunk of
// C = 2^log2c elements:
//   C = 32, 64      one vector holds 4 / 2 whole chunks
//   128 <= C <= 8192  a chunk spans C/128 lanes of one wave: xor-reduce
//   C > 8192        one wave per chunk, looping over its vectors
// Bits at element index >= n do not count, whatever the buffer holds.
constexpr int kPopcBlock = 256;

// When `hdr` is not null the kernel also zeroes the ticket, the error word
// and `init_tiles` status words for a look-back scan launched after it.
template <int ORDER, bool ALIGNED>
__global__ __launch_bounds__(kPopcBlock) void k_popc(const uint8_t* __restrict__ mask, uint32_t* __restrict__ counts, uint64_t n, uint32_t log2c,
                                                     WsHeader* hdr, uint64_t* status, uint64_t init_tiles)
{
    if (hdr) {
        const uint64_t tid = blockIdx.x * (uint64_t)kPopcBlock + threadIdx.x;
        if (tid == 0) {
            hdr->ticket = 0;
            hdr->error = 0;
        }
        for (uint64_t i = tid; i < init_tiles; i += (uint64_t)gridDim.x * kPopcBlock) status[i] = 0;
    }
    const uint64_t nbytes = (n + 7) / 8;
    const uint64_t nvec = (nbytes + 15) / 16;
    const uint64_t chunks = (n + (1ull << log2c) - 1) >> log2c;
    const uint32_t lane = lane_id();
    const uint64_t wave_global = (blockIdx.x * (uint64_t)kPopcBlock + threadIdx.x) / kWave;
    const uint64_t waves_total = (uint64_t)gridDim.x * (kPopcBlock / kWave);

    auto vec_words = [&](uint64_t v, uint32_t w[4]) {
        u32x4 x = load_mask16<ALIGNED>(mask, v * 16, nbytes);
        w[0] = to_lsb_word<ORDER>(x.x);
        w[1] = to_lsb_word<ORDER>(x.y);
        w[2] = to_lsb_word<ORDER>(x.z);
        w[3] = to_lsb_word<ORDER>(x.w);
        const uint64_t e0 = v * 128;
        if (e0 + 128 > n) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint64_t b = e0 + 32 * k;
                w[k] = b >= n ? 0u : low_bits(w[k], (uint32_t)(n - b));
            }
        }
    };

    if (log2c <= 6) {
        // C = 32 or 64: whole chunks inside one vector
        for (uint64_t v = blockIdx.x * (uint64_t)kPopcBlock + threadIdx.x; v < nvec; v += (uint64_t)gridDim.x * kPopcBlock) {
            uint32_t w[4];
            vec_words(v, w);
            if (log2c == 5) {
                const uint64_t c0 = v * 4;
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (c0 + k < chunks) counts[c0 + k] = popc32(w[k]);
            }
            else {
                const uint64_t c0 = v * 2;
                if (c0 < chunks) counts[c0] = popc32(w[0]) + popc32(w[1]);
                if (c0 + 1 < chunks) counts[c0 + 1] = popc32(w[2]) + popc32(w[3]);
            }
        }
    }
    else if (log2c <= 13) {
        // C/128 lanes per chunk (1..64); loop bounds are uniform per wave
        const uint32_t lanes_per_chunk = 1u << (log2c - 7);
        for (uint64_t vb = wave_global * kWave; vb < nvec; vb += waves_total * kWave) {
            const uint64_t v = vb + lane;
            uint32_t cnt = 0;
            if (v < nvec) {
                uint32_t w[4];
                vec_words(v, w);
                cnt = popc32(w[0]) + popc32(w[1]) + popc32(w[2]) + popc32(w[3]);
            }
            for (uint32_t d = lanes_per_chunk >> 1; d >= 1; d >>= 1) cnt += __shfl_xor(cnt, d, kWave);
            if ((lane & (lanes_per_chunk - 1)) == 0 && v < nvec) {
                const uint64_t c = v >> (log2c - 7);
                if (c < chunks) counts[c] = cnt;
            }
        }
    }
    else {
        // one wave per chunk
        const uint64_t vecs_per_chunk = 1ull << (log2c - 7);
        for (uint64_t c = wave_global; c < chunks; c += waves_total) {
            uint32_t cnt = 0;
            const uint64_t v0 = c * vecs_per_chunk;
            for (uint64_t v = v0 + lane; v < v0 + vecs_per_chunk && v < nvec; v += kWave) {
                uint32_t w[4];
                vec_words(v, w);
                cnt += popc32(w[0]) + popc32(w[1]) + popc32(w[2]) + popc32(w[3]);
            }
            cnt = wave_sum(cnt);
            if (lane == 0) counts[c] = cnt;
        }
    }
}

// ======================================================================
// phase 2: exclusive scan of chunk counts (u32 -> u64), decoupled look-back
// ======================================================================
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

// OutT = uint64_t, or uint32_t for the reference's 32-bit offsets.  counts
// and offsets may alias (in-place scan): every element is read by the
// thread that later writes it, before the look-back.
template <typename OutT>
__global__ __launch_bounds__(kScanBlock) void k_scan(const uint32_t* counts, OutT* offsets, uint64_t m, uint64_t* __restrict__ total, WsHeader* hdr,
                                                     uint64_t* status, const uint64_t* m_dev = nullptr)
{
    __shared__ uint64_t s_scr[kScanBlock / kWave];
    __shared__ uint64_t s_tile;
    __shared__ uint64_t s_prefix;
    if (threadIdx.x == 0) s_tile = atomicAdd(&hdr->ticket, 1u);
    __syncthreads();
    const uint64_t tile = s_tile;
    // m_dev: the length lives in device memory; the grid covers an upper bound
    // and the surplus workgroups leave here (uniformly, before any barrier)
    if (m_dev) m = *m_dev;
    const uint64_t ntiles = (m + kScanTile - 1) / kScanTile;
    if (tile >= ntiles) return;
    const uint64_t base = tile * kScanTile + threadIdx.x * (uint64_t)kScanItems;

    uint32_t v[kScanItems];
    if (base + kScanItems <= m && (((uintptr_t)(counts + base)) & 15) == 0) {
#pragma unroll
        for (int q = 0; q < kScanItems / 4; q++) {
            u32x4 x = *reinterpret_cast<const u32x4*>(counts + base + 4 * q);
            v[4 * q + 0] = x.x;
            v[4 * q + 1] = x.y;
            v[4 * q + 2] = x.z;
            v[4 * q + 3] = x.w;
        }
    }
    else {
#pragma unroll
        for (int k = 0; k < kScanItems; k++) v[k] = base + k < m ? counts[base + k] : 0u;
    }
    uint64_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) sum += v[k];
    uint64_t tile_total;
    const uint64_t excl = block_exclusive_sum<kScanBlock>(sum, s_scr, &tile_total);

    if (threadIdx.x < kWave) {
        uint64_t pre = 0;
        if (tile == 0) {
            if (threadIdx.x == 0) publish_status(status, kFlagPrefix | tile_total);
        }
        else {
            if (threadIdx.x == 0) publish_status(status + tile, kFlagAggregate | tile_total);
            pre = lookback_exclusive(status, tile, &hdr->error);
            if (threadIdx.x == 0) publish_status(status + tile, kFlagPrefix | (pre + tile_total));
        }
        if (threadIdx.x == 0) s_prefix = pre;
    }
    __syncthreads();
    uint64_t run = s_prefix + excl;
    if (base + kScanItems <= m) {
#pragma unroll
        for (int k = 0; k < kScanItems; k++) {
            offsets[base + k] = (OutT)run;
            run += v[k];
        }
    }
    else {
        for (int k = 0; k < kScanItems; k++) {
            if (base + k < m) offsets[base + k] = (OutT)run;
            run += v[k];
        }
    }
    if (tile == ntiles - 1 && threadIdx.x == 0 && total) *total = s_prefix + tile_total;
}

// ======================================================================
// phases 1 + 2 of the compaction pipelines: block totals, unit offsets,
// block prefixes
// ======================================================================
//
// K1 (k_popc_blocks).  The column is cut into blocks of BE = 2^log2b
// elements (popb_log2), one workgroup of THREADS threads each, taken in
// rounds of THREADS * 4 mask vectors (16 bytes, 128 elements each).  In a
// round, wave w owns vectors [256 w, 256 w + 256): its load j reads vector
// 256 w + 64 j + lane, so every wave load reads 1 KiB contiguous, and the
// four loads of the next round are issued before the current round is
// processed.  A DPP scan of each load's popcounts, the four loads' totals
// added in order, and one exchange of the four wave totals through LDS give
// every unit of U = 2^log2u elements (32 <= U) its exclusive offset within
// the block: uoffs[unit] is written by the lane that holds the unit's first
// element.  The block's selected count goes to btotals[b].
//
// K2 (k_block_prefix).  One workgroup turns the block counts into exclusive
// block prefixes (u64) and writes the total selected count.  A column of at
// most kMidBlocks blocks skips it: its writeout (kModeFewBlocks) sums the
// block counts before each tile's block in every wave, and its last tile
// writes the total.
//
// No workgroup waits for another, so there is no look-back state to
// initialise; a writeout tile's output offset is bprefix[its block] +
// uoffs[its first unit], two words every wave requests before its mask
// (scalar loads with no branch around them, so no wait for them comes
// before the mask load, and no barrier shares them).  Bits at
// element index >= n do not count, whatever the buffer holds.
constexpr int kPopbThreads = 256;
constexpr int kPopbLoads = 4;  // vectors per thread and round
static_assert((1u << kPopbMinLog2) == kPopbThreads * kPopbLoads * 128, "a round is one 2^17-element block");

// `hdr` (nullable): its error word is cleared, so that the workspace error
// query answers for this call too (no kernel of these pipelines sets it).
// THREADS = 256, 512 for blocks of 2^18 elements, 1024 for 2^19 and more: a
// block of up to 2^19 elements is one round, with the same loads in flight
// as 256-thread workgroups over 2^17-element blocks (whose larger blocks
// would take their rounds one after another).
template <int ORDER, bool ALIGNED, int THREADS = kPopbThreads>
__global__ __launch_bounds__(THREADS) void k_popc_blocks(const uint8_t* __restrict__ mask, uint64_t n, uint32_t log2b, uint32_t log2u,
                                                         uint32_t* __restrict__ uoffs, uint32_t* __restrict__ btotals, WsHeader* hdr)
{
    constexpr int kPopbWaves = THREADS / kWave;
    constexpr uint32_t kRoundVecs = THREADS * kPopbLoads;  // 16-byte mask vectors per round
    __shared__ uint32_t s_wsum[2][kPopbWaves];
    if (hdr && blockIdx.x == 0 && threadIdx.x == 0) hdr->error = 0;
    const uint32_t t = threadIdx.x, lane = lane_id(), wid = t / kWave;
    const uint64_t nbytes = (n + 7) / 8;
    const uint64_t e_block = (uint64_t)blockIdx.x << log2b;
    const uint64_t e_end = min(n, e_block + (1ull << log2b));
    const uint64_t nv = (e_end - e_block + 127) / 128;  // vectors holding elements of this block
    // 256 and 512 threads are launched for blocks of exactly one round (2^17
    // and 2^18 elements): a constant round count drops the next round's
    // prefetch registers, so 8 workgroups of 256 fit per CU in both bit orders
    const uint32_t rounds = THREADS < 1024 ? 1u : (uint32_t)((nv + kRoundVecs - 1) / kRoundVecs);
    const uint8_t* bmask = mask + e_block / 8;
    const uint64_t bbytes = nbytes - e_block / 8;
    // vector of load j in round r (block-relative)
    auto vec_of = [&](uint32_t r, int j) { return (uint64_t)r * kRoundVecs + wid * 256 + j * kWave + lane; };
    u32x4 x[kPopbLoads];
#pragma unroll
    for (int j = 0; j < kPopbLoads; j++) x[j] = vec_of(0, j) < nv ? load_mask16<ALIGNED>(bmask, vec_of(0, j) * 16, bbytes) : u32x4{0, 0, 0, 0};
    uint32_t carry = 0;  // selected elements of the block in earlier rounds
    for (uint32_t r = 0; r < rounds; r++) {
        uint32_t w[kPopbLoads][4];
#pragma unroll
        for (int j = 0; j < kPopbLoads; j++) {
            w[j][0] = to_lsb_word<ORDER>(x[j].x);
            w[j][1] = to_lsb_word<ORDER>(x[j].y);
            w[j][2] = to_lsb_word<ORDER>(x[j].z);
            w[j][3] = to_lsb_word<ORDER>(x[j].w);
        }
        // the next round's loads, in flight while this round is scanned
        if (r + 1 < rounds) {
#pragma unroll
            for (int j = 0; j < kPopbLoads; j++)
                x[j] = vec_of(r + 1, j) < nv ? load_mask16<ALIGNED>(bmask, vec_of(r + 1, j) * 16, bbytes) : u32x4{0, 0, 0, 0};
        }
        uint32_t ex[kPopbLoads];
        uint32_t wave_total = 0;  // uniform
#pragma unroll
        for (int j = 0; j < kPopbLoads; j++) {
            const uint64_t e0 = e_block + vec_of(r, j) * 128;
            if (e0 + 128 > n) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint64_t b = e0 + 32 * k;
                    w[j][k] = b >= n ? 0u : low_bits(w[j][k], (uint32_t)(n - b));
                }
            }
            const uint32_t c = popc32(w[j][0]) + popc32(w[j][1]) + popc32(w[j][2]) + popc32(w[j][3]);
            const uint32_t inc = wave_inclusive_sum_u32(c);
            ex[j] = wave_total + inc - c;
            wave_total += (uint32_t)__builtin_amdgcn_readlane((int)inc, kWave - 1);
        }
        // two LDS buffers: a round's wave totals are overwritten two rounds
        // later, after every thread has passed the barrier between
        if (lane == 0) s_wsum[r & 1][wid] = wave_total;
        __syncthreads();
        uint32_t before = 0, all = 0;
#pragma unroll
        for (int k = 0; k < kPopbWaves; k++) {
            const uint32_t s = s_wsum[r & 1][k];
            before += (uint32_t)k < wid ? s : 0u;
            all += s;
        }
#pragma unroll
        for (int j = 0; j < kPopbLoads; j++) {
            const uint64_t e0 = e_block + vec_of(r, j) * 128;
            if (e0 >= n) continue;
            const uint32_t o = carry + before + ex[j];
            if (log2u >= 7) {
                if ((e0 & ((1ull << log2u) - 1)) == 0) uoffs[e0 >> log2u] = o;
            }
            else {
                const uint32_t p0 = popc32(w[j][0]), p1 = popc32(w[j][1]), p2 = popc32(w[j][2]);
                if (log2u == 6) {
                    const uint64_t u = e0 >> 6;
                    uoffs[u] = o;
                    if (e0 + 64 < n) uoffs[u + 1] = o + p0 + p1;
                }
                else {
                    const uint64_t u = e0 >> 5;
                    if (e0 + 128 <= n) {
                        *reinterpret_cast<u32x4*>(uoffs + u) = u32x4{o, o + p0, o + p0 + p1, o + p0 + p1 + p2};
                    }
                    else {
                        const uint32_t q[4] = {o, o + p0, o + p0 + p1, o + p0 + p1 + p2};
                        for (int k = 0; k < 4; k++)
                            if (e0 + 32 * k < n) uoffs[u + k] = q[k];
                    }
                }
            }
        }
        carry += all;
    }
    if (t == 0) btotals[blockIdx.x] = carry;
}

// ======================================================================
// phase 3 / fused: the compaction tile engine
// ======================================================================
//
// A tile is TE consecutive elements handled by one 256-thread workgroup:
//   * mask: TE/8 bytes, loaded as 16-byte vectors by the first TE/128
//     threads, converted to LSB-first words and scanned (word popcounts) in
//     LDS -> for every word the number of selected elements before it;
//   * data: each thread owns VPT 16-byte vectors, striped across the block
//     (vector j*256 + t) so every load instruction of a wave reads 1 KiB
//     contiguous; a vector whose mask bits are all zero is not loaded;
//   * the tile's output offset comes from the chunk offsets or the block
//     prefixes and unit offsets (three-phase writeout), the block counts
//     before its block and its unit offset (few blocks), a decoupled
//     look-back over tile counts (fused), the tile
//     and group counts (two-pass), the caller (given) or, with no mask in
//     memory, a periodic pattern whose mask words the tile generates;
//   * STAGED: selected elements are written to LDS at their rank, shifted so
//     LDS byte p maps to a 16-byte-aligned global address, then stored with
//     16-byte stores (element stores only in the two boundary vectors that
//     the neighbouring tiles share).  !STAGED: each selected element is
//     stored straight from registers to out[offset + rank].
template <int W, int VPT_ = kTileVpt<W>, int BLOCK_ = 256> struct TileCfg {
    static constexpr int BLOCK = BLOCK_;             // threads of the tile's workgroup (128, 256 or 512)
    static constexpr int VPT = VPT_;                 // 16-byte data vectors per thread
    static constexpr int EPV = 16 / W;               // elements per vector
    static constexpr int TE = BLOCK * VPT * EPV;     // elements per tile
    static constexpr int NWORDS = TE / 32;           // 32-bit mask words per tile
    static constexpr int MASK_THREADS = NWORDS / 4;  // threads loading 16 mask bytes
    static constexpr int STAGE_BYTES = TE * W + 16;
};

template <typename C> struct TileShared {
    alignas(16) uint8_t stage[C::STAGE_BYTES];
    uint32_t words[C::NWORDS];
    uint32_t wpre[C::NWORDS];
    uint64_t scr[C::BLOCK / kWave];
    uint32_t wtot[C::BLOCK / kWave];
    uint32_t count;
    uint64_t tile;
    uint64_t off;
};

// what an absent offset array reads (zeros, never written; a global-memory
// variable so that the loads stay global, not flat, loads)
static __device__ uint64_t kZeros[2] __attribute__((aligned(16))) = {0, 0};

template <int W> __device__ __forceinline__ uint64_t vec_elem(const u32x4& x, int k)
{
    if constexpr (W == 8) {
        const uint32_t lo = k == 0 ? x.x : x.z;
        const uint32_t hi = k == 0 ? x.y : x.w;
        return ((uint64_t)hi << 32) | lo;
    }
    else {
        constexpr int per = 4 / W;
        const int comp = k / per;
        const uint32_t word = comp == 0 ? x.x : comp == 1 ? x.y : comp == 2 ? x.z : x.w;
        return (word >> (8 * W * (k % per))) & (W == 4 ? 0xFFFFFFFFu : ((1u << (8 * W)) - 1u));
    }
}

template <int W> __device__ __forceinline__ void store_elem(uint8_t* p, uint64_t v)
{
    using T = typename word_t<W>::type;
    *reinterpret_cast<T*>(p) = (T)v;
}

template <int W> __device__ __forceinline__ uint64_t load_elem(const uint8_t* p)
{
    using T = typename word_t<W>::type;
    return (uint64_t)*reinterpret_cast<const T*>(p);
}

// Selected elements before element e = tile * TE under a periodic pattern:
// whole periods times their count, plus the pattern's first (e mod L) bits.
// With tile = tq L + tr and TE = qT L + rT, e = (tile qT + tq rT) L + tr rT,
// so only 32-bit divisions by L are needed (tr rT < 1024; a tile index fits
// in 32 bits, the grid being one block per tile).
template <uint32_t TE> __device__ __forceinline__ uint64_t pattern_offset(const TileArgs& a, uint64_t tile)
{
    const uint32_t L = a.pat_len, t32 = (uint32_t)tile;
    const uint32_t qT = TE / L, rT = TE % L, tq = t32 / L, tr = t32 % L;
    const uint64_t q = (uint64_t)t32 * qT + (uint64_t)tq * rT + (tr * rT) / L;
    const uint32_t r = (tr * rT) % L;
    return q * a.pat_pop + (uint64_t)__builtin_popcountll(a.pat_bits & ((1ull << r) - 1ull));
}

// Selected elements before element tbase (a multiple of 128, below n): the
// popcount of mask bytes [0, tbase / 8), read as 16-byte vectors (a byte's
// popcount does not depend on its bit order).  prefix_start issues the first
// kPrefixLoads vectors per thread (64 KiB per tile) -- at the start of the
// tile, so they are in flight with its own mask and data; prefix_finish
// counts them and the rest, kPrefixLoads per thread at a time, and returns
// the block's sum to every thread.  Every thread of the block calls both;
// prefix_finish contains __syncthreads.
constexpr int kPrefixLoads = 16;
struct PrefixBatch {
    u32x4 x[kPrefixLoads];
};
template <typename C, bool ALIGNED> __device__ __forceinline__ void prefix_start(const TileArgs& a, uint64_t tbase, PrefixBatch& pb)
{
    const uint64_t nv = tbase / 128;
    const uint64_t nbytes = (a.n + 7) / 8;
#pragma unroll
    for (int j = 0; j < kPrefixLoads; j++) {
        const uint64_t v = (uint64_t)j * C::BLOCK + threadIdx.x;
        pb.x[j] = v < nv ? load_mask16<ALIGNED>(a.mask, v * 16, nbytes) : u32x4{0, 0, 0, 0};
    }
}
template <typename C, bool ALIGNED>
__device__ __forceinline__ uint64_t prefix_finish(const TileArgs& a, uint64_t tbase, const PrefixBatch& pb, TileShared<C>& sh)
{
    const uint32_t t = threadIdx.x;
    const uint64_t nv = tbase / 128;
    const uint64_t nbytes = (a.n + 7) / 8;
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < kPrefixLoads; j++) c += popc32(pb.x[j].x) + popc32(pb.x[j].y) + popc32(pb.x[j].z) + popc32(pb.x[j].w);
    for (uint64_t v0 = (uint64_t)kPrefixLoads * C::BLOCK; v0 < nv; v0 += (uint64_t)kPrefixLoads * C::BLOCK) {
        u32x4 x[kPrefixLoads];
#pragma unroll
        for (int j = 0; j < kPrefixLoads; j++) {
            const uint64_t v = v0 + (uint64_t)j * C::BLOCK + t;
            x[j] = v < nv ? load_mask16<ALIGNED>(a.mask, v * 16, nbytes) : u32x4{0, 0, 0, 0};
        }
#pragma unroll
        for (int j = 0; j < kPrefixLoads; j++) c += popc32(x[j].x) + popc32(x[j].y) + popc32(x[j].z) + popc32(x[j].w);
    }
    const uint32_t s = wave_sum_u32(c);
    if (lane_id() == 0) sh.wtot[t / kWave] = s;
    __syncthreads();
    uint64_t all = 0;
#pragma unroll
    for (int k = 0; k < C::BLOCK / kWave; k++) all += sh.wtot[k];
    return all;
}

// One tile: every thread of the block calls this with the same `tile`.
// Contains __syncthreads; LDS in `sh` may be reused once all threads return
// and pass a barrier.
template <int W, int ORDER, bool ALIGNED, int MODE, bool STAGED, int VPT, int BLOCK = 256>
__device__ __forceinline__ void tile_body(const TileArgs& a, TileShared<TileCfg<W, VPT, BLOCK>>& sh, uint64_t tile)
{
    using C = TileCfg<W, VPT, BLOCK>;
    const uint32_t t = threadIdx.x;
    const uint64_t n = a.n;
    const uint64_t nbytes = (n + 7) / 8;
    const uint64_t ntiles = (n + C::TE - 1) / C::TE;
    const uint64_t tbase = tile * (uint64_t)C::TE;
    const bool partial = tbase + C::TE > n;

    // Zero-chunk skipping needs no chunk counts here: data vectors are loaded
    // only where the mask selects something, and a tile whose mask words are
    // all zero returns right after its mask scan (no data read, no store).

    // two-pass: this thread's share of the selected count before the tile;
    // independent loads, issued before the mask load
    uint64_t psum = 0;
    if constexpr (MODE == kModeTwoPass) {
        const uint64_t g = tile >> a.log2g;
        const uint64_t t0 = g << a.log2g;
        uint64_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (uint64_t j0 = t; j0 < g; j0 += 8 * C::BLOCK) {
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const uint64_t j = j0 + (uint64_t)k * C::BLOCK;
                if (j < g) acc[k] += a.gsums[j];
            }
        }
        for (uint64_t j0 = t0 + t; j0 < tile; j0 += 8 * C::BLOCK) {
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const uint64_t j = j0 + (uint64_t)k * C::BLOCK;
                if (j < tile) acc[k] += a.tcounts[j];
            }
        }
#pragma unroll
        for (int k = 0; k < 8; k++) psum += acc[k];
    }

    // single launch: the first part of the mask before the tile, in flight
    // with the tile's own work
    PrefixBatch pfx;
    if constexpr (MODE == kModeSingle) prefix_start<C, ALIGNED>(a, tbase, pfx);

    // three-phase writeout with two-level offsets: the tile's block prefix
    // and its unit's offset, requested ahead of the mask.  The addresses are
    // chosen without a branch and loaded by every wave (uniform: scalar
    // loads), so that no wait for them lands before the mask load and no
    // barrier is needed to share them; an absent array reads kZeros.
    uint64_t bpre = 0;
    uint32_t uoff = 0;
    if constexpr (MODE == kModeWriteout) {
        bpre = *(a.bprefix ? a.bprefix + (tbase >> a.log2b) : kZeros);
        uoff = *(a.uoffs ? a.uoffs + (tbase >> a.log2c) : reinterpret_cast<const uint32_t*>(kZeros));
    }
    // Without block prefixes (kModeFewBlocks, a column of at most kMidBlocks
    // blocks): the totals of the blocks before this tile's, two per lane,
    // requested ahead of the mask so that its wait covers them; each wave
    // sums them where the offset is needed.
    uint32_t btot0 = 0, btot1 = 0;  // added only at the offset: an add here would wait for them
    if constexpr (MODE == kModeFewBlocks) {
        uoff = *(a.uoffs + (tbase >> a.log2c));
        const uint32_t blk = (uint32_t)(tbase >> a.log2b), l = lane_id();
        if (l < blk) btot0 = a.btotals[l];
        if (l + kWave < blk) btot1 = a.btotals[l + kWave];
    }
    // this wave's sum of the totals (all 64 lanes active)
    auto blocks_before = [&] { return (uint64_t)wave_sum_u32(btot0 + btot1); };

    // ---- mask words, word prefixes, tile count ----
    uint32_t w[4] = {0, 0, 0, 0};
    uint32_t wsum = 0;
    if (t < (uint32_t)C::MASK_THREADS) {
        if constexpr (MODE == kModePattern) {
            // 32 elements from phase (e mod L) of the period: bits [phase, phase + 32)
            // of the unrolled pattern (phase < L <= 32)
            // (tile * TE + 128 t) mod L in 32-bit arithmetic: a tile index fits in
            // 32 bits (the grid is one block per tile)
            const uint32_t L = a.pat_len;
            uint32_t phase = (((uint32_t)tile % L) * ((uint32_t)C::TE % L) + (128u * t) % L) % L;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                w[k] = (uint32_t)(a.pat_bits >> phase);
                phase += 32 % a.pat_len;
                if (phase >= a.pat_len) phase -= a.pat_len;
            }
        }
        else {
            const uint64_t off = tbase / 8 + 16ull * t;
            u32x4 x = load_mask16<ALIGNED>(a.mask, off, nbytes);
            w[0] = to_lsb_word<ORDER>(x.x);
            w[1] = to_lsb_word<ORDER>(x.y);
            w[2] = to_lsb_word<ORDER>(x.z);
            w[3] = to_lsb_word<ORDER>(x.w);
        }
        if (partial) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint64_t b = tbase + 128ull * t + 32 * k;
                w[k] = b >= n ? 0u : low_bits(w[k], (uint32_t)(n - b));
            }
        }
        wsum = popc32(w[0]) + popc32(w[1]) + popc32(w[2]) + popc32(w[3]);
    }
    // Scan of the mask threads' popcounts: each mask word's prefix in the
    // tile, and the tile count.  Two-pass folds its offset partial sums into
    // a 64-bit block scan of (psum << 16 | wsum) (the low field, at most 32768,
    // never carries into the high one).  The other modes scan 32 bits with DPP
    // row operations: in one wave when the tile's mask fits in 64 threads
    // (then one barrier publishes words, prefixes and count together), else in
    // every wave with the wave totals through LDS.
    uint32_t wex = 0, tile_count = 0;
    uint64_t prefix_total = 0;
    bool published = false;  // words, prefixes and count already visible to the block
    if constexpr (MODE == kModeTwoPass) {
        uint64_t packed_total;
        const uint64_t pex = block_exclusive_sum<C::BLOCK>((psum << 16) | wsum, sh.scr, &packed_total);
        wex = (uint32_t)(pex & 0xFFFF);
        tile_count = (uint32_t)(packed_total & 0xFFFF);
        prefix_total = packed_total >> 16;
    }
    else if constexpr (C::MASK_THREADS <= kWave) {
        if (t < (uint32_t)kWave) {
            const uint32_t inc = wave_inclusive_sum_u32(wsum);
            wex = inc - wsum;
            if (t == kWave - 1) sh.count = inc;
        }
    }
    else {
        const uint32_t inc = wave_inclusive_sum_u32(wsum);
        if (lane_id() == kWave - 1) sh.wtot[t / kWave] = inc;
        __syncthreads();
        uint32_t before = 0, all = 0;
#pragma unroll
        for (int k = 0; k < C::MASK_THREADS / kWave; k++) {
            const uint32_t s = sh.wtot[k];
            before += (uint32_t)k < t / kWave ? s : 0u;
            all += s;
        }
        wex = before + inc - wsum;
        tile_count = all;
    }
    if (t < (uint32_t)C::MASK_THREADS) {
        uint32_t run = wex;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            sh.words[4 * t + k] = w[k];
            sh.wpre[4 * t + k] = run;
            run += popc32(w[k]);
        }
    }
    // This thread's mask words and their prefixes, read together (one LDS
    // wait) once they are visible: the selected bits of each of its vectors
    // and the tile rank of each vector's first element.  Where one barrier
    // publishes words and count, they are read with the count, before the
    // empty-tile test, so that the data loads below issue back to back.
    // (The single launch, whose prefix loads hold 64 registers, reads its
    // words after the empty-tile test and the prefixes at the ranks, as
    // registers held across the test would cost it a wave per SIMD.)
    constexpr bool kEarlyRank = MODE != kModeSingle;
    bool words_read = false;
    uint32_t bits[C::VPT], rank0[C::VPT];
    auto read_words = [&] {
#pragma unroll
        for (int j = 0; j < C::VPT; j++) {
            const uint32_t e0 = (j * C::BLOCK + t) * C::EPV;
            const uint32_t word = sh.words[e0 >> 5];
            bits[j] = (word >> (e0 & 31)) & (uint32_t)((1ull << C::EPV) - 1);
            rank0[j] = kEarlyRank ? sh.wpre[e0 >> 5] + popc32(low_bits(word, e0 & 31)) : 0u;
        }
    };
    if constexpr (MODE != kModeTwoPass && C::MASK_THREADS <= kWave) {
        __syncthreads();
        tile_count = sh.count;
        published = true;
        if constexpr (kEarlyRank) {
            read_words();
            words_read = true;
            // an empty asm statement taking the values: the compiler would
            // otherwise sink the reads below the empty-tile branch, behind a
            // wait for the count alone
#pragma unroll
            for (int j = 0; j < C::VPT; j++) asm volatile("" ::"v"(bits[j]), "v"(rank0[j]));
        }
    }
    if constexpr (MODE != kModeFused) {
        // an empty tile has nothing to read or write (the fused form must
        // still publish its look-back status)
        if (tile_count == 0) {
            if (MODE == kModeTwoPass && t == 0 && tile == ntiles - 1 && a.count_out) *a.count_out = prefix_total;
            if (MODE == kModeGiven && t == 0 && tile == ntiles - 1 && a.count_out) *a.count_out = a.given_off - a.given_base;
            if (MODE == kModePattern && t == 0 && tile == ntiles - 1 && a.count_out) *a.count_out = pattern_offset<C::TE>(a, tile);
            if constexpr (MODE == kModeFewBlocks) {
                if (tile == ntiles - 1 && a.count_out) {
                    const uint64_t total = blocks_before() + uoff;
                    if (t == 0) *a.count_out = total;
                }
            }
            if constexpr (MODE == kModeSingle) {
                if (tile == ntiles - 1 && a.count_out) {
                    const uint64_t pre = prefix_finish<C, ALIGNED>(a, tbase, pfx, sh);
                    if (t == 0) *a.count_out = pre;
                }
            }
            return;
        }
    }
    if (!published) __syncthreads();
    if (!words_read) read_words();

    // ---- data loads (predicated on the vector's mask bits) ----
    u32x4 vec[C::VPT];
    const uint8_t* tin = a.in + tbase * W;
#pragma unroll
    for (int j = 0; j < C::VPT; j++) {
        const uint32_t v = j * C::BLOCK + t;
        const uint32_t e0 = v * C::EPV;
        vec[j] = u32x4{0, 0, 0, 0};
        if (bits[j]) {
            if (ALIGNED && !partial) {
                vec[j] = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(tin + (uint64_t)v * 16));
            }
            else {
                // element loads: unaligned base or the last (partial) tile;
                // unselected elements (including all at index >= n) are not read
                uint32_t parts[4] = {0, 0, 0, 0};
#pragma unroll
                for (int k = 0; k < C::EPV; k++) {
                    if ((bits[j] >> k) & 1) {
                        const uint64_t e = load_elem<W>(tin + ((uint64_t)e0 + k) * W);
                        if constexpr (W == 8) {
                            parts[2 * k] = (uint32_t)e;
                            parts[2 * k + 1] = (uint32_t)(e >> 32);
                        }
                        else {
                            constexpr int per = 4 / W;
                            parts[k / per] |= (uint32_t)e << (8 * W * (k % per));
                        }
                    }
                }
                vec[j] = u32x4{parts[0], parts[1], parts[2], parts[3]};
            }
        }
    }

    // ---- tile output offset ----
    // Two-pass (the block scan's total), given, pattern (arithmetic on the
    // tile index) and single launch (prefix_finish's result): every thread
    // holds the offset, so no barrier shares it.
    constexpr bool kOwnOffset = MODE == kModeTwoPass || MODE == kModeGiven || MODE == kModePattern || MODE == kModeSingle;
    uint64_t own_off = 0;
    if constexpr (MODE == kModeTwoPass) {
        own_off = prefix_total;
    }
    else if constexpr (MODE == kModeGiven) {
        own_off = a.given_off - a.given_base;
    }
    else if constexpr (MODE == kModePattern) {
        own_off = pattern_offset<C::TE>(a, tile);
    }
    else if constexpr (MODE == kModeSingle) {
        // finished once this tile's data loads are in flight
        own_off = prefix_finish<C, ALIGNED>(a, tbase, pfx, sh);
    }
    if constexpr (kOwnOffset) {
        if (t == 0 && tile == ntiles - 1 && a.count_out) *a.count_out = own_off + tile_count;
    }
    else if constexpr (MODE == kModeFused) {
        if (t < kWave) {
            uint64_t pre = 0;
            if (tile == 0) {
                if (t == 0) publish_status(a.status, kFlagPrefix | tile_count);
            }
            else {
                if (t == 0) publish_status(a.status + tile, kFlagAggregate | tile_count);
                pre = lookback_exclusive(a.status, tile, &a.hdr->error);
                if (t == 0) publish_status(a.status + tile, kFlagPrefix | (pre + tile_count));
            }
            if (t == 0) {
                sh.off = pre;
                if (tile == ntiles - 1 && a.count_out) *a.count_out = pre + tile_count;
            }
        }
        __syncthreads();
    }
    else if constexpr (MODE == kModeFewBlocks) {
        // every wave sums the totals before its block: no barrier either
        bpre = blocks_before();
        if (t == 0 && tile == ntiles - 1 && a.count_out) *a.count_out = bpre + uoff + tile_count;
    }
    else if (a.bprefix) {
        // every thread holds bpre and uoff: no barrier here
    }
    else {
        const uint64_t chunk = tbase >> a.log2c;
        const uint64_t cstart = chunk << a.log2c;
        uint64_t extra = 0;
        if (cstart < tbase) {
            // chunk longer than a tile: count the chunk's selected elements before this tile
            // (whole 16-byte vectors: tbase - cstart is a multiple of TE)
            const uint64_t v0 = cstart / 128, v1 = tbase / 128;
            uint32_t c = 0;
            for (uint64_t v = v0 + t; v < v1; v += C::BLOCK) {
                u32x4 x = load_mask16<ALIGNED>(a.mask, v * 16, nbytes);
                c += popc32(x.x) + popc32(x.y) + popc32(x.z) + popc32(x.w);
            }
            uint64_t tot;
            (void)block_exclusive_sum<C::BLOCK>((uint64_t)c, sh.scr, &tot);
            extra = tot;
        }
        if (t == 0) sh.off = (a.offsets ? a.offsets[chunk] : (uint64_t)a.offsets32[chunk]) + extra;
        __syncthreads();
    }
    const uint64_t tile_off = kOwnOffset ? own_off : MODE == kModeFewBlocks || (MODE == kModeWriteout && a.bprefix) ? bpre + uoff : sh.off;
    const uint64_t toff = tile_off + (a.out_base ? *a.out_base : 0ull);

    // ---- ranks, then staging (or direct stores) ----
    const uintptr_t gstart = (uintptr_t)(a.out + toff * W);
    const uint32_t shift = STAGED ? (uint32_t)(gstart & 15) : 0u;
#pragma unroll
    for (int j = 0; j < C::VPT; j++) {
        if (!bits[j]) continue;
        uint32_t r = rank0[j];
        if constexpr (!kEarlyRank) {
            const uint32_t e0 = (j * C::BLOCK + t) * C::EPV;
            r = sh.wpre[e0 >> 5] + popc32(low_bits(sh.words[e0 >> 5], e0 & 31));
        }
#pragma unroll
        for (int k = 0; k < C::EPV; k++) {
            if ((bits[j] >> k) & 1) {
                const uint64_t e = vec_elem<W>(vec[j], k);
                if constexpr (STAGED) {
                    store_elem<W>(sh.stage + shift + (uint64_t)r * W, e);
                }
                else {
                    store_elem<W>(a.out + (toff + r) * W, e);
                }
                r++;
            }
        }
    }
    if constexpr (STAGED) {
        __syncthreads();
        const uint32_t bytes = shift + (uint32_t)tile_count * W;
        const uint32_t q_end = (bytes + 15) / 16;
        uint8_t* g0 = reinterpret_cast<uint8_t*>(gstart - shift);
        for (uint32_t q = t; q < q_end; q += C::BLOCK) {
            const uint32_t lo = q * 16, hi = lo + 16;
            if (lo >= shift && hi <= bytes) {
                const u32x4 x = *reinterpret_cast<const u32x4*>(sh.stage + lo);
                __builtin_nontemporal_store(x, reinterpret_cast<u32x4*>(g0 + lo));
            }
            else {
                // boundary vector shared with a neighbouring tile: own elements only
                const uint32_t p0 = lo > shift ? lo : shift;
                const uint32_t p1 = hi < bytes ? hi : bytes;
                for (uint32_t p = p0; p < p1; p += W) store_elem<W>(g0 + p, load_elem<W>(sh.stage + p));
            }
        }
    }
}

// One tile per workgroup: tile = blockIdx.x (writeout) or a ticket (fused).
template <int W, int ORDER, bool ALIGNED, int MODE, bool STAGED, int VPT = kTileVpt<W>, int BLOCK = 256>
__global__ __launch_bounds__(BLOCK) void k_tile(TileArgs a)
{
    using C = TileCfg<W, VPT, BLOCK>;
    __shared__ TileShared<C> sh;
    uint64_t tile;
    if constexpr (MODE == kModeFused) {
        if (threadIdx.x == 0) sh.tile = atomicAdd(&a.hdr->ticket, 1u);
        __syncthreads();
        tile = sh.tile;
    }
    else if constexpr (MODE == kModeWriteout || MODE == kModeFewBlocks || MODE == kModePattern || MODE == kModeSingle) {
        // the grid is exactly one block per tile
        tile = xcd_runs(blockIdx.x, gridDim.x);
        // the single launch clears the workspace error word (nothing of it sets
        // the word, and the workspace error query answers for this call)
        if (MODE == kModeSingle && a.hdr && blockIdx.x == 0 && threadIdx.x == 0) a.hdr->error = 0;
    }
    else {
        tile = blockIdx.x;
    }
    tile_body<W, ORDER, ALIGNED, MODE, STAGED, VPT, BLOCK>(a, sh, tile);
}

// ======================================================================
// two-pass pipeline, pass 1: selected count of every tile and of every
// group of 2^log2g consecutive tiles (one workgroup per group)
// ======================================================================
constexpr uint32_t kMaxGroupTiles = 2048;

// Also clears the workspace header (ticket and error word) when `hdr` is set.
template <int W, int ORDER, bool ALIGNED>
__global__ __launch_bounds__(256) void k_tile_counts(const uint8_t* __restrict__ mask, uint64_t n, uint32_t* __restrict__ tcounts,
                                                     uint64_t* __restrict__ gsums, uint32_t log2g, WsHeader* hdr)
{
    constexpr uint64_t TE = kTileElems<W>;
    constexpr uint32_t VEC_PER_TILE = (uint32_t)(TE / 128);        // 16-byte mask vectors per tile
    constexpr uint32_t LANES_PER_TILE = VEC_PER_TILE < 64 ? VEC_PER_TILE : 64;
    constexpr uint32_t ROUND_TILES = 16;                            // tiles per unrolled round
    constexpr uint32_t VPL = ROUND_TILES * VEC_PER_TILE / 256;      // vectors per thread and round
    if (hdr && blockIdx.x == 0 && threadIdx.x == 0) {
        hdr->ticket = 0;
        hdr->error = 0;
    }
    __shared__ uint32_t tc[kMaxGroupTiles];
    __shared__ uint64_t scr[4];
    const uint32_t t = threadIdx.x, lane = lane_id();
    const uint64_t nbytes = (n + 7) / 8;
    const uint64_t nvec = (nbytes + 15) / 16;
    const uint64_t ntiles = (n + TE - 1) / TE;
    const uint64_t G = 1ull << log2g;
    const uint64_t tile0 = (uint64_t)blockIdx.x << log2g;
    const uint64_t tile1 = min(ntiles, tile0 + G);
    for (uint64_t j = t; j < G; j += 256) tc[j] = 0;
    __syncthreads();
    const uint64_t v_end = min(nvec, tile1 * VEC_PER_TILE);
    for (uint64_t r0 = tile0; r0 < tile1; r0 += ROUND_TILES) {
        const uint64_t vr = r0 * VEC_PER_TILE;  // first vector of the round
        uint32_t c[VPL];
        // issue every load of the round first
        u32x4 x[VPL];
#pragma unroll
        for (uint32_t j = 0; j < VPL; j++) {
            const uint64_t v = vr + j * 256 + t;
            x[j] = v < v_end ? load_mask16<ALIGNED>(mask, v * 16, nbytes) : u32x4{0, 0, 0, 0};
        }
#pragma unroll
        for (uint32_t j = 0; j < VPL; j++) {
            const uint64_t v = vr + j * 256 + t;
            uint32_t w[4] = {to_lsb_word<ORDER>(x[j].x), to_lsb_word<ORDER>(x[j].y), to_lsb_word<ORDER>(x[j].z), to_lsb_word<ORDER>(x[j].w)};
            const uint64_t e0 = v * 128;
            if (e0 + 128 > n) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint64_t b = e0 + 32 * k;
                    w[k] = b >= n ? 0u : low_bits(w[k], (uint32_t)(n - b));
                }
            }
            c[j] = popc32(w[0]) + popc32(w[1]) + popc32(w[2]) + popc32(w[3]);
        }
#pragma unroll
        for (uint32_t j = 0; j < VPL; j++) {
            uint32_t cj = c[j];
#pragma unroll
            for (uint32_t d = LANES_PER_TILE / 2; d >= 1; d >>= 1) cj += __shfl_xor(cj, d, kWave);
            const uint64_t v = vr + j * 256 + t;
            if ((lane & (LANES_PER_TILE - 1)) == 0 && v < v_end && cj) atomicAdd(&tc[(v - tile0 * VEC_PER_TILE) / VEC_PER_TILE], cj);
        }
    }
    __syncthreads();
    uint64_t s = 0;
    for (uint64_t j = t; j < tile1 - tile0; j += 256) {
        tcounts[tile0 + j] = tc[j];
        s += tc[j];
    }
    uint64_t total;
    (void)block_exclusive_sum<256>(s, scr, &total);
    if (t == 0) gsums[blockIdx.x] = total;
}

// ======================================================================
// phase 3, thread-per-chunk variant (src/kernels/kernel_3pass.cuh:578-639)
// ======================================================================
template <int W, int ORDER>
__global__ __launch_bounds__(256) void k_thread_per_chunk(TileArgs a)
{
    const uint8_t* __restrict__ in = a.in;
    const uint8_t* __restrict__ mask = a.mask;
    const uint32_t* __restrict__ counts = a.counts;
    uint8_t* __restrict__ out = a.out;
    const uint64_t n = a.n;
    const uint32_t log2c = a.log2c;
    const uint64_t base = a.out_base ? *a.out_base : 0ull;
    const uint64_t chunks = (n + (1ull << log2c) - 1) >> log2c;
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < chunks; c += (uint64_t)gridDim.x * blockDim.x) {
        if (counts && counts[c] == 0) continue;
        uint64_t o = base + (a.offsets ? a.offsets[c] : (uint64_t)a.offsets32[c]);
        const uint64_t e_begin = c << log2c;
        const uint64_t e_end = min(n, e_begin + (1ull << log2c));
        for (uint64_t i = e_begin; i < e_end; i += 8) {
            const uint32_t byte = mask[i / 8];
            const uint32_t lim = (uint32_t)min<uint64_t>(8, e_end - i);
            for (uint32_t b = 0; b < lim; b++) {
                const uint32_t bit = ORDER == kBitMsb ? (byte >> (7 - b)) & 1u : (byte >> b) & 1u;
                if (bit) {
                    store_elem<W>(out + o * W, load_elem<W>(in + (i + b) * W));
                    o++;
                }
            }
        }
    }
}

// ======================================================================
// phase 3, sparse variant: no LDS, for low selectivity
// ======================================================================
//
// A wave covers a step of 64 x 128 elements: lane l owns the 16-byte mask
// vector of elements [128 l, 128 l + 128) of the step.  A wave scan of the
// lanes' popcounts gives each lane its first output rank; the lane then
// gathers its selected elements (up to four loads in flight) and stores them
// one by one.  Without the staging buffer a workgroup holds no LDS, so many
// more steps are in flight than tiles of the staged writeout, whose LDS is
// sized for a fully selected tile: below a few percent selectivity, where the
// staged tile waits on load latency, this form is faster; above, the
// per-element gathers lose to whole-vector loads.
constexpr uint64_t kSparseStep = kSparseStepElems;

// FEW: a column of at most kMidBlocks blocks without block prefixes (as the
// tile engine's kModeFewBlocks): each step's wave sums the block counts
// before its block, and the wave of the last step writes the total count.
template <int W, int ORDER, bool ALIGNED, bool FEW = false>
__global__ __launch_bounds__(256) void k_sparse(TileArgs a)
{
    const uint8_t* __restrict__ in = a.in;
    uint8_t* __restrict__ out = a.out;
    const uint64_t n = a.n;
    const uint64_t nbytes = (n + 7) / 8;
    const uint32_t lane = lane_id();
    const uint64_t steps = (n + kSparseStep - 1) / kSparseStep;
    const uint64_t base_out = a.out_base ? *a.out_base : 0ull;
    const uint64_t waves = (uint64_t)gridDim.x * (256 / kWave);
    for (uint64_t step = (blockIdx.x * 256ull + threadIdx.x) / kWave; step < steps; step += waves) {
        const uint64_t e0 = step * kSparseStep;
        const uint64_t el = e0 + 128ull * lane;
        // two-level offsets: the step's block prefix and unit offset,
        // requested ahead of the mask with no branch around them (an absent
        // array reads kZeros), so that no wait for them precedes the mask load
        const uint64_t bpre = FEW ? 0ull : *(a.bprefix ? a.bprefix + (e0 >> a.log2b) : kZeros);
        const uint32_t uoff = *(a.uoffs ? a.uoffs + (e0 >> a.log2c) : reinterpret_cast<const uint32_t*>(kZeros));
        // FEW: the counts of the blocks before the step's, two per lane, also
        // ahead of the mask (added only where the offset is needed)
        uint32_t btot0 = 0, btot1 = 0;
        if constexpr (FEW) {
            const uint32_t blk = (uint32_t)(e0 >> a.log2b);
            if (lane < blk) btot0 = a.btotals[lane];
            if (lane + kWave < blk) btot1 = a.btotals[lane + kWave];
        }
        uint32_t w[4] = {0, 0, 0, 0};
        if (el < n) {
            const u32x4 x = load_mask16<ALIGNED>(a.mask, el / 8, nbytes);
            w[0] = to_lsb_word<ORDER>(x.x);
            w[1] = to_lsb_word<ORDER>(x.y);
            w[2] = to_lsb_word<ORDER>(x.z);
            w[3] = to_lsb_word<ORDER>(x.w);
            if (el + 128 > n) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint64_t b = el + 32 * k;
                    w[k] = b >= n ? 0u : low_bits(w[k], (uint32_t)(n - b));
                }
            }
        }
        // output offset of the step: with two-level offsets the totals of
        // the blocks before it plus its unit's offset (the unit divides the
        // step); else its chunk's offset, plus the selected elements of the
        // chunk before the step when the chunk is longer
        const uint64_t chunk = e0 >> a.log2c;
        const uint64_t cstart = chunk << a.log2c;
        uint64_t before = 0, step_off = 0;
        if constexpr (FEW) {
            step_off = (uint64_t)wave_sum_u32(btot0 + btot1) + uoff;
        }
        else {
            if (a.bprefix) {
                step_off = bpre + uoff;
            }
            else if (cstart < e0) {
                // whole 16-byte vectors below e0 < n: no tail bits among them
                uint32_t c = 0;
                for (uint64_t v = cstart / 128 + lane; v < e0 / 128; v += kWave) {
                    const u32x4 x = load_mask16<ALIGNED>(a.mask, v * 16, nbytes);
                    c += popc32(x.x) + popc32(x.y) + popc32(x.z) + popc32(x.w);
                }
                before = wave_sum((uint64_t)c);
            }
            if (!a.bprefix) step_off = (a.offsets ? a.offsets[chunk] : (uint64_t)a.offsets32[chunk]) + before;
        }
        const uint32_t cnt = popc32(w[0]) + popc32(w[1]) + popc32(w[2]) + popc32(w[3]);
        const uint32_t inc = wave_inclusive_sum_u32(cnt);
        if constexpr (FEW) {
            // the last step's lane 63 holds the step's count: the column's total follows
            if (step == steps - 1 && lane == kWave - 1 && a.count_out) *a.count_out = step_off + inc;
        }
        uint64_t r = base_out + step_off + inc - cnt;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            uint32_t bits = w[k];
            const uint8_t* src = in + (el + 32 * k) * W;
            while (bits) {
                // up to four selected elements per round, their loads issued together
                const uint32_t q0 = __builtin_ctz(bits);
                bits &= bits - 1;
                const bool h1 = bits != 0;
                const uint32_t q1 = h1 ? __builtin_ctz(bits) : 0u;
                bits &= h1 ? bits - 1 : bits;
                const bool h2 = bits != 0;
                const uint32_t q2 = h2 ? __builtin_ctz(bits) : 0u;
                bits &= h2 ? bits - 1 : bits;
                const bool h3 = bits != 0;
                const uint32_t q3 = h3 ? __builtin_ctz(bits) : 0u;
                bits &= h3 ? bits - 1 : bits;
                const uint64_t v0 = load_elem<W>(src + q0 * W);
                const uint64_t v1 = h1 ? load_elem<W>(src + q1 * W) : 0ull;
                const uint64_t v2 = h2 ? load_elem<W>(src + q2 * W) : 0ull;
                const uint64_t v3 = h3 ? load_elem<W>(src + q3 * W) : 0ull;
                store_elem<W>(out + r * W, v0);
                if (h1) store_elem<W>(out + (r + 1) * W, v1);
                if (h2) store_elem<W>(out + (r + 2) * W, v2);
                if (h3) store_elem<W>(out + (r + 3) * W, v3);
                r += 1u + h1 + h2 + h3;
            }
        }
    }
}

// ======================================================================
// bytewise comparison (src/kernels/data_generator.cuh:122-161)
// ======================================================================
template <int W>
__global__ __launch_bounds__(256) void k_validate(const uint8_t* __restrict__ a, const uint8_t* __restrict__ b, uint64_t len, uint64_t* failures)
{
    uint64_t bad = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < len; i += (uint64_t)gridDim.x * blockDim.x) {
        bad += load_elem<W>(a + i * W) != load_elem<W>(b + i * W);
    }
    bad = wave_sum(bad);
    if (lane_id() == 0 && bad) atomicAdd((unsigned long long*)failures, (unsigned long long)bad);
}

} // namespace csx
