// csx_types.hpp -- types shared by the host dispatch code and the kernels.
#pragma once

#include <stdint.h>

namespace csx {

// Workspace header of every single-pass (look-back) kernel, 256 bytes.
struct WsHeader {
    uint32_t ticket;  // dynamic tile counter
    uint32_t error;   // non-zero: a look-back gave up
    uint64_t count;   // selected count when the caller passes no d_count
    uint32_t pad[60];
};
static_assert(sizeof(WsHeader) == 256, "header is one 256-byte block");

// 16-byte input vectors per thread of the compaction tile engine, per width.
template <int W> constexpr int kTileVpt = 8;
// Elements per tile: 256 threads x kTileVpt vectors of 16 bytes.
template <int W> constexpr uint64_t kTileElems = 256ull * kTileVpt<W> * (16ull / W);

// Elements per tile of the chunk-count scan.
constexpr uint64_t kScanTileElems = 256ull * 16ull;

// Elements per wave step of the sparse writeout (64 lanes x 128 elements).
constexpr uint64_t kSparseStepElems = 64ull * 128ull;

// Blocks of the pipelines' popcount (k_popc_blocks) for n elements: log2 of
// the block length, 2^17 elements while that makes at most kPopbMaxBlocks
// blocks (n <= 2^32), longer beyond (at most 2^25 elements for n <= 2^40).
constexpr uint32_t kPopbMinLog2 = 17;
constexpr uint32_t kPopbMaxBlocks = 32768;
inline uint32_t popb_log2(uint64_t n)
{
    uint32_t l = kPopbMinLog2;
    while (((n + (1ull << l) - 1) >> l) > kPopbMaxBlocks) l++;
    return l;
}
inline uint64_t popb_blocks(uint64_t n) { return (n + (1ull << popb_log2(n)) - 1) >> popb_log2(n); }
// columns of up to this many blocks skip the block-prefix kernel: each
// writeout wave sums the totals before its block, two per lane
constexpr uint32_t kMidBlocks = 128;

// kModeGiven: the tile's output offset is supplied in TileArgs::given_off
// (less TileArgs::given_base)
// kModePattern: no mask in memory; mask words and the tile's offset follow
// from a periodic pattern (TileArgs::pat_*)
// kModeSingle: the tile's offset is the popcount of the whole mask before
// it, counted by the tile itself (one launch, no other kernel's results)
// kModeFewBlocks: the three-phase writeout of a column of at most kMidBlocks
// popcount blocks, with no block prefixes (TileArgs::btotals)
enum TileMode { kModeWriteout = 0, kModeFused = 1, kModeTwoPass = 2, kModeGiven = 3, kModePattern = 4, kModeSingle = 5, kModeFewBlocks = 6 };

// Arguments of one tile-engine launch.
struct TileArgs {
    const uint8_t* in;
    const uint8_t* mask;
    uint8_t* out;
    uint64_t n;
    // three-phase writeout
    const uint64_t* offsets;    // exclusive per-chunk offsets (or:)
    const uint32_t* offsets32;  // the same as 32-bit values, used when offsets is null
    const uint32_t* counts;     // nullable: skip tiles whose chunks are empty
    uint32_t log2c;
    const uint64_t* out_base;   // nullable: device value added to every output offset
    // two-level offsets of the pipelines (k_popc_blocks + k_block_prefix),
    // used instead of offsets/offsets32 when bprefix is set: a tile's (or
    // sparse step's) output offset is bprefix[first element >> log2b] (the
    // selected count of the blocks before its own) plus
    // uoffs[first element >> log2c] (its unit's offset within the block;
    // 2^log2c divides the tile)
    const uint64_t* bprefix;
    const uint32_t* uoffs;
    uint32_t log2b;
    // kModeFewBlocks, instead of bprefix, for columns of at most kMidBlocks
    // blocks (and a total below 2^32): the blocks' selected counts, summed by
    // every wave of a tile for the blocks before its own, and the total count
    // written by the last tile to count_out (no block-prefix kernel)
    const uint32_t* btotals;
    // fused
    WsHeader* hdr;
    uint64_t* status;
    uint64_t* count_out;  // nullable: total selected count (device)
    // two-pass: per-tile counts and per-group sums of 2^log2g tiles
    const uint32_t* tcounts;
    const uint64_t* gsums;
    uint32_t log2g;
    // given offset (kModeGiven): given_off - given_base, subtracted where the
    // offset is used (at the stores), so that a caller's load of either value
    // is waited for there rather than ahead of the tile's mask load
    uint64_t given_off;
    uint64_t given_base;
    // periodic pattern (kModePattern): bit k of pat_bits (LSB first) is the
    // selection of element k for k < 64, a period of pat_len elements with
    // pat_pop selected
    uint64_t pat_bits;
    uint32_t pat_len;
    uint32_t pat_pop;
};

// One column of a batched compaction (device-resident table).
struct Segment {
    const void* in;
    const uint8_t* mask;
    void* out;
    uint64_t n;
};

// One tile of a batched compaction, written by the count kernel: what the
// write kernels need to address the tile's mask and data, so that those loads
// wait for one record instead of the tile's segment index and then its
// descriptor.  Tile indices are below 2^24 (one launch), so `first` fits.
struct TileRec {
    const void* in;       // the segment's input
    const uint8_t* mask;  // the segment's mask
    uint64_t n;           // the segment's length
    uint32_t seg;         // the segment's index
    uint32_t first;       // the segment's first tile (index over all tiles)
};
static_assert(sizeof(TileRec) == 32, "a record is one 32-byte store");

// A predicate as one range test over the element type (csx_predicate.hip):
// selected = (!(flags & kPredLo) || v >(=) lo) && (!(flags & kPredHi) || v <(=) hi),
// xor (flags & kPredNegate); lo/hi hold the bounds' bytes (little-endian low bytes).
constexpr uint32_t kPredLo = 1, kPredLoIncl = 2, kPredHi = 4, kPredHiIncl = 8, kPredNegate = 16;
struct PredSpec {
    int type;  // csx_type
    uint64_t lo, hi;
    uint32_t flags;
};

// Tiles per group of the two-pass pipeline: a power of two >= 16 keeping
// the group count at most 2048 (the writeout sums at most 8 group sums per
// thread), capped at 2048 tiles per group (the pass-1 LDS table).
inline uint32_t two_pass_log2g(uint64_t ntiles)
{
    uint32_t l = 4;
    while (l < 11 && ((ntiles + (1ull << l) - 1) >> l) > 2048) l++;
    return l;
}

} // namespace csx
