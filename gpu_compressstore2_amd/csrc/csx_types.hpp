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

// Elements per tile of the compaction tile engine: 256 threads x 8 vectors
// of 16 bytes, i.e. 32 KiB of input per tile whatever the width.
template <int W> constexpr uint64_t kTileElems = 256ull * 8ull * (16ull / W);

// Elements per tile of the chunk-count scan.
constexpr uint64_t kScanTileElems = 256ull * 16ull;

enum TileMode { kModeWriteout = 0, kModeFused = 1 };

// Arguments of one tile-engine launch.
struct TileArgs {
    const uint8_t* in;
    const uint8_t* mask;
    uint8_t* out;
    uint64_t n;
    // three-phase writeout
    const uint64_t* offsets;  // exclusive per-chunk offsets
    const uint32_t* counts;   // nullable: skip tiles whose chunks are empty
    uint32_t log2c;
    // fused
    WsHeader* hdr;
    uint64_t* status;
    uint64_t* count_out;  // nullable: total selected count (device)
};

} // namespace csx
