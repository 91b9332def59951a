// csx_tile.hip -- tile-engine launchers for one element width.
// Built once per width with -DCSX_TILE_W=1|2|4|8.
#include "csx_internal.hpp"
#include "csx_kernels.hpp"

#ifndef CSX_TILE_W
#error "define CSX_TILE_W (element width in bytes)"
#endif

namespace csx {

template <int W, int ORDER, bool ALIGNED, int MODE, int VPT, int BLOCK> static void launch_writeout_shape(const TileArgs& a, hipStream_t s)
{
    using C = TileCfg<W, VPT, BLOCK>;
    const uint64_t tiles = (a.n + C::TE - 1) / C::TE;
    launch_k(k_tile<W, ORDER, ALIGNED, MODE, true, VPT, BLOCK>, dim3((unsigned)tiles), dim3(BLOCK), s, a);
}

// the staged writeout's tile (kModeWriteout or kModeFewBlocks): `block`
// threads (128, 256 or 512) x VPT vectors
template <int W, int ORDER, bool ALIGNED, int MODE, int VPT> static void launch_writeout_vpt(const TileArgs& a, int block, hipStream_t s)
{
    switch (block) {
        case 128: launch_writeout_shape<W, ORDER, ALIGNED, MODE, VPT, 128>(a, s); break;
        case 512: launch_writeout_shape<W, ORDER, ALIGNED, MODE, VPT, 512>(a, s); break;
        default: launch_writeout_shape<W, ORDER, ALIGNED, MODE, VPT, 256>(a, s); break;
    }
}

template <int W, int ORDER, bool ALIGNED, int MODE> static void launch_writeout_mode(const TileArgs& a, int vpt, int block, hipStream_t s)
{
    // three-phase writeout: a tile's offset comes from the chunk offsets (or
    // the block totals), so any tile size works with the same workspace
    switch (vpt) {
        case 2: launch_writeout_vpt<W, ORDER, ALIGNED, MODE, 2>(a, block, s); break;
        case 4: launch_writeout_vpt<W, ORDER, ALIGNED, MODE, 4>(a, block, s); break;
        case 16:
            if constexpr (W >= 2) launch_writeout_vpt<W, ORDER, ALIGNED, MODE, 16>(a, block, s);
            break;
        default: launch_writeout_vpt<W, ORDER, ALIGNED, MODE, 8>(a, block, s); break;
    }
}

// the single-launch pipeline's tile: 256 threads x VPT vectors
template <int W, int ORDER, bool ALIGNED, int VPT> static void launch_single_vpt(const TileArgs& a, hipStream_t s)
{
    const uint64_t tiles = (a.n + TileCfg<W, VPT>::TE - 1) / TileCfg<W, VPT>::TE;
    launch_k(k_tile<W, ORDER, ALIGNED, kModeSingle, true, VPT>, dim3((unsigned)tiles), dim3(256), s, a);
}

template <int W, bool ALIGNED, int VPT> static void launch_pattern_vpt(const TileArgs& a, hipStream_t s)
{
    const uint64_t tiles = (a.n + TileCfg<W, VPT>::TE - 1) / TileCfg<W, VPT>::TE;
    launch_k(k_tile<W, kBitMsb, ALIGNED, kModePattern, true, VPT>, dim3((unsigned)tiles), dim3(256), s, a);
}

template <int W, int ORDER, bool ALIGNED>
static void launch_tile_impl(int mode, bool staged, int vpt, int block, const TileArgs& a, unsigned grid, hipStream_t s)
{
    if (mode == kModeTwoPass) {
        launch_k(k_tile<W, ORDER, ALIGNED, kModeTwoPass, true>, dim3(grid), dim3(256), s, a);
    }
    else if (mode == kModeFused) {
        launch_k(k_tile<W, ORDER, ALIGNED, kModeFused, true>, dim3(grid), dim3(256), s, a);
    }
    else if (mode == kModeSingle) {
        switch (vpt) {
            case 2: launch_single_vpt<W, ORDER, ALIGNED, 2>(a, s); break;
            case 4: launch_single_vpt<W, ORDER, ALIGNED, 4>(a, s); break;
            default: launch_single_vpt<W, ORDER, ALIGNED, 8>(a, s); break;
        }
    }
    else if (mode == kModeGiven) {
        // one tile of 256 threads x vpt vectors at a.given_off (inputs of at
        // most one such tile: the whole call)
        switch (vpt) {
            case 2: launch_k(k_tile<W, ORDER, ALIGNED, kModeGiven, true, 2>, dim3(1), dim3(256), s, a); break;
            case 4: launch_k(k_tile<W, ORDER, ALIGNED, kModeGiven, true, 4>, dim3(1), dim3(256), s, a); break;
            case 16:
                if constexpr (W >= 2) launch_k(k_tile<W, ORDER, ALIGNED, kModeGiven, true, 16>, dim3(1), dim3(256), s, a);
                break;
            default: launch_k(k_tile<W, ORDER, ALIGNED, kModeGiven, true>, dim3(grid), dim3(256), s, a); break;
        }
    }
    else if (mode == kModePattern) {
        // no mask to order: one instantiation per tile size
        if constexpr (ORDER == kBitMsb) {
            switch (vpt) {
                case 2: launch_pattern_vpt<W, ALIGNED, 2>(a, s); break;
                case 4: launch_pattern_vpt<W, ALIGNED, 4>(a, s); break;
                case 16:
                    if constexpr (W >= 2) launch_pattern_vpt<W, ALIGNED, 16>(a, s);
                    break;
                default: launch_pattern_vpt<W, ALIGNED, 8>(a, s); break;
            }
        }
    }
    else if (mode == kModeFewBlocks) {
        launch_writeout_mode<W, ORDER, ALIGNED, kModeFewBlocks>(a, vpt, block, s);
    }
    else if (staged) {
        launch_writeout_mode<W, ORDER, ALIGNED, kModeWriteout>(a, vpt, block, s);
    }
    else {
        launch_k(k_tile<W, ORDER, ALIGNED, kModeWriteout, false>, dim3(grid), dim3(256), s, a);
    }
}

template <int W> hipError_t launch_tile(int mode, bool staged, int order, bool aligned, const TileArgs& a, hipStream_t s, int vpt, int block)
{
    if (vpt != 2 && vpt != 4 && vpt != 8 && !(vpt == 16 && W >= 2)) return hipErrorInvalidValue;
    if (mode == kModeSingle && vpt == 16) return hipErrorInvalidValue;
    // only the staged writeout's tile takes another workgroup size than 256
    if (block != 128 && block != 256 && block != 512) return hipErrorInvalidValue;
    if (block != 256 && !((mode == kModeWriteout || mode == kModeFewBlocks) && staged)) return hipErrorInvalidValue;
    if (mode == kModeFewBlocks && (!staged || !a.btotals || !a.uoffs)) return hipErrorInvalidValue;
    const uint64_t tiles = (a.n + kTileElems<W> - 1) / kTileElems<W>;
    if (tiles == 0) return hipSuccess;
    const unsigned grid = (unsigned)tiles;
    if (order == kBitMsb) {
        if (aligned)
            launch_tile_impl<W, kBitMsb, true>(mode, staged, vpt, block, a, grid, s);
        else
            launch_tile_impl<W, kBitMsb, false>(mode, staged, vpt, block, a, grid, s);
    }
    else {
        if (aligned)
            launch_tile_impl<W, kBitLsb, true>(mode, staged, vpt, block, a, grid, s);
        else
            launch_tile_impl<W, kBitLsb, false>(mode, staged, vpt, block, a, grid, s);
    }
    return hipGetLastError();
}

template <int W>
hipError_t launch_tile_counts(int order, bool aligned, const uint8_t* mask, uint64_t n, uint32_t* tcounts, uint64_t* gsums, uint32_t log2g,
                              WsHeader* hdr, hipStream_t s)
{
    const uint64_t ntiles = (n + kTileElems<W> - 1) / kTileElems<W>;
    const uint64_t groups = (ntiles + (1ull << log2g) - 1) >> log2g;
    if (groups == 0) return hipSuccess;
    if (order == kBitMsb) {
        if (aligned)
            k_tile_counts<W, kBitMsb, true><<<(unsigned)groups, 256, 0, s>>>(mask, n, tcounts, gsums, log2g, hdr);
        else
            k_tile_counts<W, kBitMsb, false><<<(unsigned)groups, 256, 0, s>>>(mask, n, tcounts, gsums, log2g, hdr);
    }
    else {
        if (aligned)
            k_tile_counts<W, kBitLsb, true><<<(unsigned)groups, 256, 0, s>>>(mask, n, tcounts, gsums, log2g, hdr);
        else
            k_tile_counts<W, kBitLsb, false><<<(unsigned)groups, 256, 0, s>>>(mask, n, tcounts, gsums, log2g, hdr);
    }
    return hipGetLastError();
}

template <int W> hipError_t launch_thread_per_chunk(int order, const TileArgs& a, hipStream_t s)
{
    const uint64_t chunks = (a.n + (1ull << a.log2c) - 1) >> a.log2c;
    uint64_t g = (chunks + 255) / 256;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    if (order == kBitMsb)
        launch_k(k_thread_per_chunk<W, kBitMsb>, dim3((unsigned)g), dim3(256), s, a);
    else
        launch_k(k_thread_per_chunk<W, kBitLsb>, dim3((unsigned)g), dim3(256), s, a);
    return hipGetLastError();
}

template <int W, bool FEW> static void launch_sparse_few(int order, bool mask_aligned, const TileArgs& a, unsigned g, hipStream_t s)
{
    if (order == kBitMsb) {
        if (mask_aligned)
            launch_k(k_sparse<W, kBitMsb, true, FEW>, dim3(g), dim3(256), s, a);
        else
            launch_k(k_sparse<W, kBitMsb, false, FEW>, dim3(g), dim3(256), s, a);
    }
    else {
        if (mask_aligned)
            launch_k(k_sparse<W, kBitLsb, true, FEW>, dim3(g), dim3(256), s, a);
        else
            launch_k(k_sparse<W, kBitLsb, false, FEW>, dim3(g), dim3(256), s, a);
    }
}

// with a.btotals (a column of at most kMidBlocks blocks, no block prefixes)
// the waves sum the block counts themselves and the last step writes the count
template <int W> hipError_t launch_sparse(int order, bool mask_aligned, const TileArgs& a, hipStream_t s)
{
    const uint64_t steps = (a.n + kSparseStep - 1) / kSparseStep;
    uint64_t g = (steps + 3) / 4;  // one step per wave
    if (g < 1) g = 1;
    if (g > (1u << 22)) g = 1u << 22;  // beyond 2^37 elements the waves loop
    if (a.btotals)
        launch_sparse_few<W, true>(order, mask_aligned, a, (unsigned)g, s);
    else
        launch_sparse_few<W, false>(order, mask_aligned, a, (unsigned)g, s);
    return hipGetLastError();
}

template hipError_t launch_tile<CSX_TILE_W>(int, bool, int, bool, const TileArgs&, hipStream_t, int, int);
template hipError_t launch_sparse<CSX_TILE_W>(int, bool, const TileArgs&, hipStream_t);
template hipError_t launch_tile_counts<CSX_TILE_W>(int, bool, const uint8_t*, uint64_t, uint32_t*, uint64_t*, uint32_t, WsHeader*, hipStream_t);
template hipError_t launch_thread_per_chunk<CSX_TILE_W>(int, const TileArgs&, hipStream_t);

} // namespace csx
