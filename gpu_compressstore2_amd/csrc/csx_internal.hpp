// csx_internal.hpp -- host-side launchers behind the C ABI (not exported).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "csx.h"

#include "csx_types.hpp"

namespace csx {

// Kernel timing for csx_profile_begin: an event pair armed by the caller
// (csx_api.cpp) is consumed by the next launch_k on this host thread, which
// launches through hipExtLaunchKernelGGL so that the events carry the start
// and end of that kernel itself, not of the stream around it.
struct LaunchEvents {
    hipEvent_t start = nullptr, stop = nullptr;
    bool used = false;
};
extern thread_local LaunchEvents g_launch_events;
// the launch shape of the next kernel launched through launch_k while armed
// (csx_last_launch reports it)
struct LaunchShape {
    bool armed = false;
    uint64_t grid = 0;
    uint32_t block = 0;
};
extern thread_local LaunchShape g_launch_shape;
template <typename K, typename... A> inline void launch_k(K kernel, dim3 grid, dim3 block, hipStream_t s, A... args)
{
    LaunchShape& sh = g_launch_shape;
    if (sh.armed) {
        sh.armed = false;
        sh.grid = (uint64_t)grid.x * grid.y * grid.z;
        sh.block = block.x * block.y * block.z;
    }
    LaunchEvents& ev = g_launch_events;
    if (ev.start && !ev.used) {
        ev.used = true;
        hipExtLaunchKernelGGL(kernel, grid, block, 0, s, ev.start, ev.stop, 0, args...);
    }
    else {
        kernel<<<grid, block, 0, s>>>(args...);
    }
}

// the calling thread's csx_last_error() message (csx_api.cpp)
csx_status set_error(csx_status s, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
csx_status clear_error();
// csx_compact_config's body with no profiling event pairs (csx_api.cpp); the
// no-allocation entry point runs it in its static scratch
csx_status compact_in_workspace(const void* d_in, const uint8_t* d_mask, void* d_out, uint64_t n, uint32_t elem_size, int bit_order,
                                const csx_config* cfg, uint64_t* d_count, uint64_t* h_count, void* d_ws, size_t ws_bytes, hipStream_t stream);

// With `hdr` set, the popcount kernel also initialises the look-back state of
// `init_tiles` scan tiles (saves the init launch before launch_scan).
hipError_t launch_popc(const uint8_t* mask, uint32_t* counts, uint64_t n, uint32_t log2c, int order, hipStream_t s, WsHeader* hdr = nullptr,
                       uint64_t* status = nullptr, uint64_t init_tiles = 0);
inline uint64_t scan_tiles(uint64_t m) { return (m + kScanTileElems - 1) / kScanTileElems; }
// The pipelines' popcount: one workgroup per block of 2^log2b elements writes
// the block's selected count to btotals and every 2^log2u-element unit's
// exclusive offset within its block to uoffs (csx_kernels.hpp k_popc_blocks);
// clears hdr->error when hdr is set.  Then one workgroup turns the block
// counts into exclusive block prefixes and writes the total to *count
// (nullable); blocks <= kPopbMaxBlocks; `wide`: the total may reach 2^32.
hipError_t launch_popc_blocks(const uint8_t* mask, uint64_t n, uint32_t log2b, uint32_t log2u, uint32_t* uoffs, uint32_t* btotals, int order,
                              hipStream_t s, WsHeader* hdr = nullptr);
hipError_t launch_block_prefix(const uint32_t* btotals, uint64_t blocks, uint64_t* bprefix, uint64_t* count, bool wide, hipStream_t s);
hipError_t launch_lookback_init(WsHeader* hdr, uint64_t* status, uint64_t tiles, hipStream_t s);
// initialized: the look-back state was already zeroed (by launch_popc)
hipError_t launch_scan(const uint32_t* counts, uint64_t* offsets, uint64_t m, uint64_t* total, WsHeader* hdr, uint64_t* status, hipStream_t s,
                       bool initialized = false);
// 32-bit offsets (counts and offsets may alias)
hipError_t launch_scan32(const uint32_t* counts, uint32_t* offsets, uint64_t m, uint64_t* total, WsHeader* hdr, uint64_t* status, hipStream_t s);
hipError_t launch_tree_from_offsets(uint32_t* pss, const uint64_t* ex, uint64_t m, hipStream_t s);
hipError_t launch_tree_offsets(const uint32_t* tree, uint32_t* out, uint64_t m, hipStream_t s);
hipError_t launch_store_u64(uint64_t* dst, uint64_t v, hipStream_t s);
// both pointers 16-byte aligned
hipError_t launch_copy(void* dst, const void* src, uint64_t bytes, hipStream_t s);
// csx_stream_probe's kernels (kind 0 read, 1 write, 2 copy, 3 2:1 mix)
hipError_t launch_stream_probe(int kind, const void* src, void* dst, uint64_t bytes, uint32_t vpt, uint32_t xcd_run, hipStream_t s);
// *dst = base[i-1] + total[i-1] (0 for i = 0): the multi-stream carry
hipError_t launch_carry(uint64_t* dst, const uint64_t* base, const uint64_t* total, uint32_t i, hipStream_t s);
// batched compaction (csx_batch.hip)
// vpt in {2, 4, 8}, block in {128, 256, 512}: tile of block * vpt * 16/w
// elements written by a workgroup of `block` threads; count_blocks bounds the
// persistent grid of the per-tile count kernel
uint64_t batch_tile_elems(uint32_t w, uint32_t vpt, uint32_t block);
uint64_t batch_max_tiles(uint32_t w, uint32_t nseg, uint64_t max_total, uint32_t vpt, uint32_t block);
hipError_t launch_batch(uint32_t w, int order, const Segment* segs, uint32_t nseg, uint64_t max_tiles, uint64_t* counts, uint64_t* tile_prefix,
                        uint32_t* tcounts, TileRec* rec, uint64_t* goff, WsHeader* hdr, uint64_t* status, hipStream_t s, uint32_t vpt = 8,
                        uint32_t count_blocks = 4096, uint32_t block = 256);
// device generators (csx_gen.hip)
hipError_t launch_gen_bernoulli(uint8_t* mask, uint64_t n, uint64_t thr, uint64_t seed, int order, hipStream_t s);
hipError_t launch_gen_pattern(uint8_t* mask, uint64_t n, uint32_t pattern, uint32_t L, int order, hipStream_t s);
hipError_t launch_gen_zipf(uint8_t* mask, uint64_t n, uint32_t block_len, uint64_t seed, const uint32_t* cdf, const uint64_t* thr, uint32_t levels,
                           int order, hipStream_t s);
hipError_t launch_gen_runs(uint8_t* mask, uint64_t n, uint32_t thr, uint64_t seed, int order, uint32_t* counts, uint64_t* before, WsHeader* hdr,
                           uint64_t* status, hipStream_t s);
hipError_t launch_gen_values(void* out, uint64_t n, uint32_t w, int kind, uint64_t seed, hipStream_t s);
// column + comparison -> mask (csx_predicate.hip); counts per 1024-element
// chunk and the running total are optional, hdr/status as in launch_popc
hipError_t launch_predicate(const void* in, uint64_t n, const PredSpec& p, uint8_t* mask, int order, uint32_t* counts, uint64_t* total, WsHeader* hdr,
                            uint64_t* status, uint64_t init_tiles, hipStream_t s);
// periodic-pattern compaction (csx_pattern.hip); k = selected count
hipError_t launch_validate(const void* a, const void* b, uint64_t len, uint32_t w, uint64_t* failures, hipStream_t s);

// tile engine, instantiated once per element width (csx_tile.hip)
// vpt: 16-byte vectors per thread of the three-phase writeout tile (2, 4, 8, or 16
// for W >= 2), block: its workgroup size (128, 256, 512); the other modes use
// kTileVpt and 256
template <int W>
hipError_t launch_tile(int mode, bool staged, int order, bool aligned, const TileArgs& a, hipStream_t s, int vpt = kTileVpt<W>, int block = 256);
template <int W>
hipError_t launch_tile_counts(int order, bool aligned, const uint8_t* mask, uint64_t n, uint32_t* tcounts, uint64_t* gsums, uint32_t log2g,
                              WsHeader* hdr, hipStream_t s);
template <int W> hipError_t launch_thread_per_chunk(int order, const TileArgs& a, hipStream_t s);
// sparse writeout (three-phase offsets, no LDS); mask_aligned: 16-byte mask pointer
template <int W> hipError_t launch_sparse(int order, bool mask_aligned, const TileArgs& a, hipStream_t s);

extern template hipError_t launch_tile<1>(int, bool, int, bool, const TileArgs&, hipStream_t, int, int);
extern template hipError_t launch_tile<2>(int, bool, int, bool, const TileArgs&, hipStream_t, int, int);
extern template hipError_t launch_tile<4>(int, bool, int, bool, const TileArgs&, hipStream_t, int, int);
extern template hipError_t launch_tile<8>(int, bool, int, bool, const TileArgs&, hipStream_t, int, int);
extern template hipError_t launch_tile_counts<1>(int, bool, const uint8_t*, uint64_t, uint32_t*, uint64_t*, uint32_t, WsHeader*, hipStream_t);
extern template hipError_t launch_thread_per_chunk<1>(int, const TileArgs&, hipStream_t);
extern template hipError_t launch_sparse<1>(int, bool, const TileArgs&, hipStream_t);
extern template hipError_t launch_tile_counts<2>(int, bool, const uint8_t*, uint64_t, uint32_t*, uint64_t*, uint32_t, WsHeader*, hipStream_t);
extern template hipError_t launch_thread_per_chunk<2>(int, const TileArgs&, hipStream_t);
extern template hipError_t launch_sparse<2>(int, bool, const TileArgs&, hipStream_t);
extern template hipError_t launch_tile_counts<4>(int, bool, const uint8_t*, uint64_t, uint32_t*, uint64_t*, uint32_t, WsHeader*, hipStream_t);
extern template hipError_t launch_thread_per_chunk<4>(int, const TileArgs&, hipStream_t);
extern template hipError_t launch_sparse<4>(int, bool, const TileArgs&, hipStream_t);
extern template hipError_t launch_tile_counts<8>(int, bool, const uint8_t*, uint64_t, uint32_t*, uint64_t*, uint32_t, WsHeader*, hipStream_t);
extern template hipError_t launch_thread_per_chunk<8>(int, const TileArgs&, hipStream_t);
extern template hipError_t launch_sparse<8>(int, bool, const TileArgs&, hipStream_t);

} // namespace csx
