// csx_batch.hip -- batched compaction of many independent columns.
//
// The columns are a device-resident table of segments {in, mask, out, n}.
// Every launch reads its sizes from device memory, so the sequence needs no
// host round trip and can be captured into a graph once and replayed with a
// new table (the host only bounds the total for grid and workspace sizes):
//   K0 plan    one workgroup: tile prefix over the segments, the total tile
//              count, zeroed per-segment counts and look-back state
//   K1 counts  persistent grid walking the (segment, tile) queue, one wave
//              load per tile (per group of tiles when a tile has fewer than
//              64 mask vectors): selected count and segment of every tile
//   K2 scan    the look-back scan over all tile counts (length from device)
//   K3 write   one workgroup per (segment, tile): the tile engine with the
//              tile's offset made local to its segment; the last tile of a
//              segment writes the segment's selected count; empty tiles and
//              tiles of at most 1/64 density leave at once
//   K4 sparse  persistent grid over the tiles of at most 1/64 density: one
//              wave per tile, selected elements gathered without LDS
// The reference's closest pattern is its two-stream split with a carry
// (src/benchmarks.cuh:195-293); here the split is by column, needing no carry.
#include "csx_internal.hpp"
#include "csx_kernels.hpp"

namespace csx {

constexpr uint32_t kPlanItems = 16;

// tile_prefix[s] = tiles of segments [0, s); tile_prefix[nseg] = total
__global__ __launch_bounds__(256) void k_batch_plan(const Segment* segs, uint32_t nseg, uint64_t te, uint64_t max_tiles, uint64_t* tile_prefix,
                                                    uint64_t* counts, WsHeader* hdr, uint64_t* status, uint64_t status_words)
{
    __shared__ uint64_t scr[4];
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        hdr->ticket = 0;
        hdr->error = 0;
    }
    for (uint64_t i = t; i < status_words; i += 256) status[i] = 0;
    uint64_t run = 0;
    for (uint64_t b = 0; b < nseg; b += 256 * kPlanItems) {
        uint64_t v[kPlanItems];
        uint64_t sum = 0;
#pragma unroll
        for (uint32_t k = 0; k < kPlanItems; k++) {
            const uint64_t s = b + (uint64_t)t * kPlanItems + k;
            v[k] = s < nseg ? (segs[s].n + te - 1) / te : 0;
            if (s < nseg) counts[s] = 0;
            sum += v[k];
        }
        uint64_t total;
        const uint64_t ex = block_exclusive_sum<256>(sum, scr, &total);
        uint64_t p = run + ex;
#pragma unroll
        for (uint32_t k = 0; k < kPlanItems; k++) {
            const uint64_t s = b + (uint64_t)t * kPlanItems + k;
            if (s < nseg) tile_prefix[s] = p;
            p += v[k];
        }
        run += total;
        __syncthreads();
    }
    if (t == 0) {
        if (run > max_tiles) {
            hdr->error = 2;  // more tiles than the host bound: the launches below cover max_tiles only
            run = max_tiles;
        }
        tile_prefix[nseg] = run;
    }
}

// Segment s with tile_prefix[s] <= g < tile_prefix[s + 1] (the last such s:
// empty segments repeat a prefix).  Called by all 64 lanes of a wave with the
// same g: each round the lanes probe 64 evenly spaced points of the interval
// and keep the stretch after the last probe that is <= g, so 4096 segments
// take two dependent loads instead of twelve.  Lane 0 probes lo, whose prefix
// is <= g throughout, so the ballot is never empty.
__device__ __forceinline__ uint32_t find_segment(const uint64_t* tile_prefix, uint32_t nseg, uint64_t g)
{
    const uint32_t lane = lane_id();
    uint32_t lo = 0, hi = nseg;  // answer in [lo, hi)
    while (hi - lo > 1) {
        const uint32_t step = (hi - lo + kWave - 1) / kWave;
        const uint32_t p = lo + lane * step;
        const uint64_t le = __ballot(p < hi && tile_prefix[p] <= g);
        const uint32_t last = 63u - (uint32_t)__builtin_clzll(le);
        lo = lo + last * step;
        hi = hi < lo + step ? hi : lo + step;
    }
    return lo;
}

// A wave's place in the segment table, in scalar registers: segment s holds
// tiles [first, end), and in, mask, n are its descriptor.  The table and the
// tile prefix are not written by the kernels that walk them, so every load
// here is a scalar load (counted apart from the mask loads in flight: moving
// to the next segment never waits for them).
struct SegCursor {
    uint32_t s;
    uint64_t first, end;
    const void* in;
    const uint8_t* mask;
    uint64_t n;
};

__device__ __forceinline__ void cursor_fill(SegCursor& c, const Segment* segs)
{
    c.in = load_invariant(&segs[c.s].in);
    c.mask = load_invariant(&segs[c.s].mask);
    c.n = load_invariant(&segs[c.s].n);
}

// the cursor at the segment holding tile g (g below the total)
__device__ __forceinline__ SegCursor cursor_at(const Segment* segs, const uint64_t* tile_prefix, uint32_t nseg, uint64_t g)
{
    SegCursor c;
    c.s = (uint32_t)__builtin_amdgcn_readfirstlane((int)find_segment(tile_prefix, nseg, g));
    c.first = load_invariant(tile_prefix + c.s);
    c.end = load_invariant(tile_prefix + c.s + 1);
    cursor_fill(c, segs);
    return c;
}

// the next non-empty segment (the caller knows one holds a tile below the total)
__device__ __forceinline__ void cursor_next(SegCursor& c, const Segment* segs, const uint64_t* tile_prefix)
{
    do {
        c.s++;
        c.first = c.end;
        c.end = load_invariant(tile_prefix + c.s + 1);
    } while (c.end == c.first);
    cursor_fill(c, segs);
}

// A tile's record, word `word` (0..3) of it: lanes 0..3 of a wave together
// make one 32-byte store.
__device__ __forceinline__ void store_rec(TileRec* rec, uint64_t g, uint32_t word, const void* in, const uint8_t* mask, uint64_t n, uint32_t seg,
                                          uint64_t first)
{
    const uint64_t v = word == 0 ? (uint64_t)(uintptr_t)in : word == 1 ? (uint64_t)(uintptr_t)mask : word == 2 ? n : (first << 32) | seg;
    reinterpret_cast<uint64_t*>(rec + g)[word] = v;
}

// Each wave counts a contiguous run of tiles: one segment search for its first
// tile, then the cursor advances as the run crosses segment ends (empty
// segments included), and kCountUnroll tiles' mask loads are in flight together.
constexpr int kCountUnroll = 4;

// the wave's run of tiles [g0, g1) of the total: `per` tiles each
__device__ __forceinline__ void wave_run(uint64_t total, uint64_t& g0, uint64_t& g1)
{
    // made uniform for the compiler, so that the cursor stays in scalar registers
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)((blockIdx.x * 256u + threadIdx.x) / kWave));
    const uint64_t waves = gridDim.x * (256ull / kWave);
    const uint64_t per = (total + waves - 1) / waves;
    g0 = wave * per;
    g1 = g0 + per < total ? g0 + per : total;
}

// Tiles of fewer than 64 mask vectors (VEC = TE / 128): one wave load covers
// TPW = 64 / VEC consecutive tiles, lane L on vector L % VEC of the load's
// tile L / VEC, and each group of VEC lanes reduces its own tile's count.  The
// segments that meet a load are visited in order by the wave's cursor, each
// handing its descriptor to the lanes whose tiles it holds (a column shorter
// than a load's tiles is one of several).
template <int W, int ORDER, uint32_t VEC, uint64_t TE>
__device__ __forceinline__ void batch_counts_packed(const Segment* segs, uint32_t nseg, const uint64_t* tile_prefix, uint32_t* tcounts,
                                                    TileRec* rec)
{
    constexpr uint32_t TPW = kWave / VEC;
    static_assert(VEC * TPW == kWave && TE == 128ull * VEC && VEC >= 4, "tiles pack a wave load exactly; a record takes 4 lanes");
    const uint32_t lane = lane_id();
    const uint32_t j = lane / VEC, v = lane % VEC;
    const uint64_t total = load_invariant(tile_prefix + nseg);
    uint64_t g0, g1;
    wave_run(total, g0, g1);
    if (g0 >= g1) return;  // wave-uniform
    SegCursor c = cursor_at(segs, tile_prefix, nseg, g0);
    for (uint64_t g = g0; g < g1; g += (uint64_t)kCountUnroll * TPW) {
        u32x4 x[kCountUnroll];
        uint64_t n_of[kCountUnroll], e_of[kCountUnroll];
        bool live[kCountUnroll];
#pragma unroll
        for (int u = 0; u < kCountUnroll; u++) {
            const uint64_t gb = g + (uint64_t)u * TPW;  // the load's first tile (wave-uniform)
            const uint64_t gu = gb + j;                  // this lane's tile
            const uint64_t glast = gb + TPW < g1 ? gb + TPW : g1;
            x[u] = u32x4{0, 0, 0, 0};
            n_of[u] = 0;
            e_of[u] = 0;
            live[u] = gu < g1;
            if (gb >= g1) continue;
            while (gb >= c.end) cursor_next(c, segs, tile_prefix);
            // the segments holding tiles [gb, glast): each lane keeps its own
            const void* lin = nullptr;
            const uint8_t* lmask = nullptr;
            uint64_t ln = 0, lf = 0;
            uint32_t lseg = 0;
            while (true) {
                if (gu >= c.first && gu < c.end) {
                    lin = c.in;
                    lmask = c.mask;
                    ln = c.n;
                    lf = c.first;
                    lseg = c.s;
                }
                if (glast <= c.end) break;  // uniform
                cursor_next(c, segs, tile_prefix);
            }
            if (live[u]) {
                // the tile's record, for the write kernels (its count below)
                if (v < 4) store_rec(rec, gu, v, lin, lmask, ln, lseg, lf);
                const uint64_t lt = gu - lf;
                const uint64_t nbytes = (ln + 7) / 8;
                const uint64_t vec = lt * VEC + v;
                x[u] = ((uintptr_t)lmask & 15) == 0 ? load_mask16<true>(lmask, vec * 16, nbytes) : load_mask16<false>(lmask, vec * 16, nbytes);
                n_of[u] = ln;
                e_of[u] = lt * TE;
            }
        }
#pragma unroll
        for (int u = 0; u < kCountUnroll; u++) {
            const uint64_t gb = g + (uint64_t)u * TPW;
            if (gb >= g1) break;  // wave-uniform
            uint32_t w[4] = {to_lsb_word<ORDER>(x[u].x), to_lsb_word<ORDER>(x[u].y), to_lsb_word<ORDER>(x[u].z), to_lsb_word<ORDER>(x[u].w)};
            const uint64_t e0 = e_of[u] + 128ull * v;
            if (live[u] && e0 + 128 > n_of[u]) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint64_t b = e0 + 32 * k;
                    w[k] = b >= n_of[u] ? 0u : low_bits(w[k], (uint32_t)(n_of[u] - b));
                }
            }
            uint32_t c4 = live[u] ? popc32(w[0]) + popc32(w[1]) + popc32(w[2]) + popc32(w[3]) : 0u;
#pragma unroll
            for (uint32_t d = VEC / 2; d >= 1; d >>= 1) c4 += __shfl_xor(c4, d, kWave);
            if (live[u] && v == 0) tcounts[gb + j] = c4;
        }
    }
}

template <int W, int ORDER, int VPT, int BLOCK>
__global__ __launch_bounds__(256) void k_batch_counts(const Segment* segs, uint32_t nseg, const uint64_t* tile_prefix, uint32_t* tcounts,
                                                      TileRec* rec)
{
    constexpr uint64_t TE = TileCfg<W, VPT, BLOCK>::TE;
    constexpr uint32_t VEC_PER_TILE = (uint32_t)(TE / 128);
    if constexpr (VEC_PER_TILE < kWave) {
        batch_counts_packed<W, ORDER, VEC_PER_TILE, TE>(segs, nseg, tile_prefix, tcounts, rec);
        return;
    }
    constexpr uint32_t VPL = (VEC_PER_TILE + kWave - 1) / kWave;  // mask vectors per lane per tile
    const uint32_t lane = lane_id();
    const uint64_t total = load_invariant(tile_prefix + nseg);
    uint64_t g0, g1;
    wave_run(total, g0, g1);
    if (g0 >= g1) return;  // wave-uniform
    SegCursor c = cursor_at(segs, tile_prefix, nseg, g0);
    // the counts of tiles pbase + l, l < 64, held by lane l and stored 64 at
    // a time: a store per tile at the end of each round would make the next
    // round's first loads wait for it
    uint32_t pend = 0;
    uint64_t pbase = g0;
    for (uint64_t g = g0; g < g1; g += kCountUnroll) {
        u32x4 x[kCountUnroll][VPL];
        uint64_t n_of[kCountUnroll], e_of[kCountUnroll];
#pragma unroll
        for (int u = 0; u < kCountUnroll; u++) {
            const uint64_t gu = g + u;
            n_of[u] = 0;
            e_of[u] = 0;
#pragma unroll
            for (uint32_t q = 0; q < VPL; q++) x[u][q] = u32x4{0, 0, 0, 0};
            if (gu >= g1) continue;
            while (gu >= c.end) cursor_next(c, segs, tile_prefix);
            const uint64_t lt = gu - c.first;
            const uint64_t nbytes = (c.n + 7) / 8;
            const bool aligned = ((uintptr_t)c.mask & 15) == 0;
            n_of[u] = c.n;
            e_of[u] = lt * TE;
            if (lane < 4) store_rec(rec, gu, lane, c.in, c.mask, c.n, c.s, c.first);  // for the write kernels
#pragma unroll
            for (uint32_t q = 0; q < VPL; q++) {
                const uint32_t v = q * kWave + lane;
                if (v < VEC_PER_TILE) {
                    const uint64_t vec = lt * VEC_PER_TILE + v;
                    x[u][q] = aligned ? load_mask16<true>(c.mask, vec * 16, nbytes) : load_mask16<false>(c.mask, vec * 16, nbytes);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kCountUnroll; u++) {
            const uint64_t gu = g + u;
            if (gu >= g1) break;
            uint32_t cnt = 0;
#pragma unroll
            for (uint32_t q = 0; q < VPL; q++) {
                const uint32_t v = q * kWave + lane;
                uint32_t w[4] = {to_lsb_word<ORDER>(x[u][q].x), to_lsb_word<ORDER>(x[u][q].y), to_lsb_word<ORDER>(x[u][q].z),
                                 to_lsb_word<ORDER>(x[u][q].w)};
                const uint64_t e0 = e_of[u] + 128ull * v;
                if (v >= VEC_PER_TILE) continue;
                if (e0 + 128 > n_of[u]) {
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const uint64_t b = e0 + 32 * k;
                        w[k] = b >= n_of[u] ? 0u : low_bits(w[k], (uint32_t)(n_of[u] - b));
                    }
                }
                cnt += popc32(w[0]) + popc32(w[1]) + popc32(w[2]) + popc32(w[3]);
            }
            cnt = wave_sum(cnt);
            const uint32_t slot = (uint32_t)(gu - pbase);  // uniform, below 64
            if (lane == slot) pend = cnt;
            if (slot == kWave - 1 || gu + 1 == g1) {
                if (lane <= slot) tcounts[pbase + lane] = pend;
                pbase += kWave;
            }
        }
    }
}

// A tile with at most TE / kSparseDensity selected elements (about 1.6 %)
// is written by the sparse kernel (k_batch_sparse) instead of the tile engine.
constexpr uint32_t kSparseDensity = 64;
// grid bound of the sparse kernel
constexpr uint64_t kSparseBlocks = 2048;

template <int W, int ORDER, int VPT, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_batch_write(const Segment* segs, uint32_t nseg, const uint64_t* tile_prefix, const uint64_t* goff,
                                                       const uint32_t* tcounts, const TileRec* rec, uint64_t* counts)
{
    using C = TileCfg<W, VPT, BLOCK>;
    __shared__ TileShared<C> sh;
    // runs of consecutive tiles per XCD (see xcd_runs)
    const uint64_t g = xcd_runs(blockIdx.x, gridDim.x);
    // One round of wave-uniform scalar loads addresses the tile: the real tile
    // count, and the tile's record (written by the count kernel), global
    // offset and count (g is below the grid, the max_tiles entries of each).
    // The mask and data loads wait for this round only; the segment's output
    // and the offset of its first tile, needed at the stores, follow in their
    // shadow.  No search, no barrier.  The empty asm statement holds the
    // round's values, so the compiler issues its loads together instead of
    // one after another; it also takes the kernel arguments used later, which
    // would otherwise be fetched after the second round and wait for it.  A
    // surplus tile of the host's bound has an unwritten record and leaves once
    // the total is in, before using it.
    const uint64_t total = load_invariant(tile_prefix + nseg);
    const TileRec* r = rec + g;
    const uint8_t* in = static_cast<const uint8_t*>(load_invariant(&r->in));
    const uint8_t* mask = load_invariant(&r->mask);
    const uint64_t n = load_invariant(&r->n);
    const uint64_t sf = load_invariant(reinterpret_cast<const uint64_t*>(r) + 3);  // seg | first << 32
    const uint64_t gofs = load_invariant(goff + g);
    const uint32_t tc = load_invariant(tcounts + g);
    asm volatile("" ::"s"(total), "s"(in), "s"(mask), "s"(n), "s"(sf), "s"(gofs), "s"(tc), "s"(segs), "s"(counts));
    // both below 2^32 (the grid is one launch; the total is clamped to it):
    // a 32-bit test stays scalar, where a 64-bit one would move the total to
    // a vector register and wait for it ahead of the other loads
    if ((uint32_t)g >= (uint32_t)total) return;  // uniform
    const uint32_t s = (uint32_t)sf;
    const uint64_t first = sf >> 32;
    const uint64_t tile = g - first;
    const uint64_t gfirst = load_invariant(goff + first);
    if (tc <= C::TE / kSparseDensity) {
        // empty, or sparse and written by k_batch_sparse: only the segment's
        // last tile has something to do here, the segment's count
        if (threadIdx.x == 0 && tile == (n + C::TE - 1) / C::TE - 1) counts[s] = gofs - gfirst + tc;
        return;
    }
    TileArgs a{};
    a.in = in;
    a.mask = mask;
    a.out = static_cast<uint8_t*>(load_invariant(&segs[s].out));
    a.n = n;
    a.given_off = gofs;
    a.given_base = gfirst;
    a.count_out = counts + s;
    if ((((uintptr_t)in) | ((uintptr_t)mask)) & 15)
        tile_body<W, ORDER, false, kModeGiven, true, VPT, BLOCK>(a, sh, tile);
    else
        tile_body<W, ORDER, true, kModeGiven, true, VPT, BLOCK>(a, sh, tile);
}

// The sparse tiles (1 <= count <= TE / kSparseDensity): a persistent grid of
// W waves, tiles interleaved across them -- lane l of wave w reads the count
// of tile base + l W, base = w, w + 64 W, ... -- so that the consecutive
// sparse tiles of one column spread over many waves instead of queueing in
// one.  Each sparse tile in the wave's ballot is written in turn: its segment
// from the lane that read its count, then a lane per 16-byte mask vector (128
// elements) of the tile, ranks from a wave scan, the selected elements
// gathered (up to four loads in flight) and stored one by one -- no LDS, so
// the tile costs a few wave instructions instead of a workgroup of the tile
// engine waiting on its loads.
template <int W, int ORDER, int VPT, int BLOCK>
__global__ __launch_bounds__(256) void k_batch_sparse(const Segment* segs, uint32_t nseg, const uint64_t* tile_prefix, const uint32_t* tcounts,
                                                      const TileRec* rec, const uint64_t* goff)
{
    constexpr uint64_t TE = TileCfg<W, VPT, BLOCK>::TE;
    constexpr uint32_t VEC = (uint32_t)(TE / 128);
    constexpr uint32_t THR = (uint32_t)(TE / kSparseDensity);
    const uint32_t lane = lane_id();
    const uint64_t total = tile_prefix[nseg];
    const uint64_t wave = (blockIdx.x * 256ull + threadIdx.x) / kWave;
    const uint64_t waves = gridDim.x * (256ull / kWave);
    for (uint64_t base = wave; base < total; base += kWave * waves) {
        const uint64_t gl = base + lane * waves;
        const uint32_t tc = gl < total ? tcounts[gl] : 0u;
        const bool sparse = tc >= 1 && tc <= THR;
        // the lane's tile's segment, read only for a sparse tile: a batch
        // with none reads the counts alone
        const uint32_t sl = sparse ? rec[gl].seg : 0u;
        uint64_t todo = __ballot(sparse);
        while (todo) {
            const uint32_t idx = (uint32_t)__builtin_ctzll(todo);
            const uint64_t g = base + (uint64_t)idx * waves;
            todo &= todo - 1;
            const uint32_t s = (uint32_t)__builtin_amdgcn_readlane((int)sl, (int)idx);
            const uint64_t s_first = tile_prefix[s];
            const Segment seg = segs[s];
            const uint64_t n = seg.n, nbytes = (n + 7) / 8;
            const uint64_t e_t = (g - s_first) * TE;
            const bool aligned = ((uintptr_t)seg.mask & 15) == 0;
            const uint8_t* __restrict__ in = static_cast<const uint8_t*>(seg.in);
            uint8_t* __restrict__ out = static_cast<uint8_t*>(seg.out);
            uint64_t r0 = goff[g] - goff[s_first];  // the tile's offset in its segment
            for (uint32_t v0 = 0; v0 < VEC; v0 += kWave) {
                const uint32_t v = v0 + lane;
                const uint64_t el = e_t + 128ull * v;
                uint32_t w[4] = {0, 0, 0, 0};
                if (v < VEC && el < n) {
                    const u32x4 x = aligned ? load_mask16<true>(seg.mask, el / 8, nbytes) : load_mask16<false>(seg.mask, el / 8, nbytes);
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
                const uint32_t cnt = popc32(w[0]) + popc32(w[1]) + popc32(w[2]) + popc32(w[3]);
                const uint32_t inc = wave_inclusive_sum_u32(cnt);
                uint64_t r = r0 + inc - cnt;
                r0 += (uint32_t)__builtin_amdgcn_readlane((int)inc, kWave - 1);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    uint32_t bits = w[k];
                    const uint8_t* src = in + (el + 32 * k) * W;
                    while (bits) {
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
                        const uint64_t x0 = load_elem<W>(src + q0 * W);
                        const uint64_t x1 = h1 ? load_elem<W>(src + q1 * W) : 0ull;
                        const uint64_t x2 = h2 ? load_elem<W>(src + q2 * W) : 0ull;
                        const uint64_t x3 = h3 ? load_elem<W>(src + q3 * W) : 0ull;
                        store_elem<W>(out + r * W, x0);
                        if (h1) store_elem<W>(out + (r + 1) * W, x1);
                        if (h2) store_elem<W>(out + (r + 2) * W, x2);
                        if (h3) store_elem<W>(out + (r + 3) * W, x3);
                        r += 1u + h1 + h2 + h3;
                    }
                }
            }
        }
    }
}

uint64_t batch_tile_elems(uint32_t w, uint32_t vpt, uint32_t block) { return (uint64_t)block * vpt * (16ull / w); }

uint64_t batch_max_tiles(uint32_t w, uint32_t nseg, uint64_t max_total, uint32_t vpt, uint32_t block)
{
    return max_total / batch_tile_elems(w, vpt, block) + nseg;
}

template <int W, int ORDER, int VPT, int BLOCK>
static hipError_t batch_impl(const Segment* segs, uint32_t nseg, uint64_t max_tiles, uint64_t* counts, uint64_t* tile_prefix, uint32_t* tcounts,
                             TileRec* rec, uint64_t* goff, WsHeader* hdr, uint64_t* status, uint32_t count_blocks, hipStream_t s)
{
    const uint64_t status_words = (max_tiles + kScanTile - 1) / kScanTile;
    k_batch_plan<<<1, 256, 0, s>>>(segs, nseg, TileCfg<W, VPT, BLOCK>::TE, max_tiles, tile_prefix, counts, hdr, status, status_words);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || max_tiles == 0) return e;
    uint64_t g1 = (max_tiles + 3) / 4;
    if (g1 > count_blocks) g1 = count_blocks;
    k_batch_counts<W, ORDER, VPT, BLOCK><<<(unsigned)g1, 256, 0, s>>>(segs, nseg, tile_prefix, tcounts, rec);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_scan<uint64_t><<<(unsigned)status_words, kScanBlock, 0, s>>>(tcounts, goff, max_tiles, nullptr, hdr, status, tile_prefix + nseg);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_batch_write<W, ORDER, VPT, BLOCK><<<(unsigned)max_tiles, BLOCK, 0, s>>>(segs, nseg, tile_prefix, goff, tcounts, rec, counts);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // the sparse tiles: a persistent grid resident at once (8 workgroups per
    // CU; with the count kernel's larger grid the waves queue in rounds and a
    // batch without sparse tiles took 1 % longer)
    const uint64_t g4 = g1 < kSparseBlocks ? g1 : kSparseBlocks;
    k_batch_sparse<W, ORDER, VPT, BLOCK><<<(unsigned)g4, 256, 0, s>>>(segs, nseg, tile_prefix, tcounts, rec, goff);
    return hipGetLastError();
}

template <int W, int VPT, int BLOCK>
static hipError_t batch_order(int order, const Segment* segs, uint32_t nseg, uint64_t max_tiles, uint64_t* counts, uint64_t* tile_prefix,
                              uint32_t* tcounts, TileRec* rec, uint64_t* goff, WsHeader* hdr, uint64_t* status, uint32_t count_blocks, hipStream_t s)
{
    return order == kBitMsb
               ? batch_impl<W, kBitMsb, VPT, BLOCK>(segs, nseg, max_tiles, counts, tile_prefix, tcounts, rec, goff, hdr, status, count_blocks, s)
               : batch_impl<W, kBitLsb, VPT, BLOCK>(segs, nseg, max_tiles, counts, tile_prefix, tcounts, rec, goff, hdr, status, count_blocks, s);
}

template <int W, int VPT>
static hipError_t batch_block(uint32_t block, int order, const Segment* segs, uint32_t nseg, uint64_t max_tiles, uint64_t* counts,
                              uint64_t* tile_prefix, uint32_t* tcounts, TileRec* rec, uint64_t* goff, WsHeader* hdr, uint64_t* status, uint32_t count_blocks,
                              hipStream_t s)
{
    switch (block) {
        case 128: return batch_order<W, VPT, 128>(order, segs, nseg, max_tiles, counts, tile_prefix, tcounts, rec, goff, hdr, status, count_blocks, s);
        case 256: return batch_order<W, VPT, 256>(order, segs, nseg, max_tiles, counts, tile_prefix, tcounts, rec, goff, hdr, status, count_blocks, s);
        case 512: return batch_order<W, VPT, 512>(order, segs, nseg, max_tiles, counts, tile_prefix, tcounts, rec, goff, hdr, status, count_blocks, s);
        default: return hipErrorInvalidValue;
    }
}

template <int W>
static hipError_t batch_vpt(uint32_t vpt, uint32_t block, int order, const Segment* segs, uint32_t nseg, uint64_t max_tiles, uint64_t* counts,
                            uint64_t* tile_prefix, uint32_t* tcounts, TileRec* rec, uint64_t* goff, WsHeader* hdr, uint64_t* status, uint32_t count_blocks,
                            hipStream_t s)
{
    switch (vpt) {
        case 2: return batch_block<W, 2>(block, order, segs, nseg, max_tiles, counts, tile_prefix, tcounts, rec, goff, hdr, status, count_blocks, s);
        case 4: return batch_block<W, 4>(block, order, segs, nseg, max_tiles, counts, tile_prefix, tcounts, rec, goff, hdr, status, count_blocks, s);
        case 8: return batch_block<W, 8>(block, order, segs, nseg, max_tiles, counts, tile_prefix, tcounts, rec, goff, hdr, status, count_blocks, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_batch(uint32_t w, int order, const Segment* segs, uint32_t nseg, uint64_t max_tiles, uint64_t* counts, uint64_t* tile_prefix,
                        uint32_t* tcounts, TileRec* rec, uint64_t* goff, WsHeader* hdr, uint64_t* status, hipStream_t s, uint32_t vpt, uint32_t count_blocks,
                        uint32_t block)
{
    if (count_blocks < 1) count_blocks = 1;
    switch (w) {
        case 1: return batch_vpt<1>(vpt, block, order, segs, nseg, max_tiles, counts, tile_prefix, tcounts, rec, goff, hdr, status, count_blocks, s);
        case 2: return batch_vpt<2>(vpt, block, order, segs, nseg, max_tiles, counts, tile_prefix, tcounts, rec, goff, hdr, status, count_blocks, s);
        case 4: return batch_vpt<4>(vpt, block, order, segs, nseg, max_tiles, counts, tile_prefix, tcounts, rec, goff, hdr, status, count_blocks, s);
        default: return batch_vpt<8>(vpt, block, order, segs, nseg, max_tiles, counts, tile_prefix, tcounts, rec, goff, hdr, status, count_blocks, s);
    }
}

} // namespace csx
