// csx_api.cpp -- the C ABI (include/csx.h): argument checks, workspace
// planning and dispatch to the gfx950 kernels.  No call here allocates
// device memory; scratch lives in the caller's workspace.
#include "csx.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "csx_internal.hpp"

namespace {

thread_local std::string g_last_error;

csx_status vfail(csx_status s, const char* fmt, va_list ap)
{
    char buf[512];
    vsnprintf(buf, sizeof(buf), fmt, ap);
    g_last_error = buf;
    return s;
}

csx_status fail(csx_status s, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vfail(s, fmt, ap);
    va_end(ap);
    return s;
}

csx_status ok()
{
    g_last_error.clear();
    return CSX_OK;
}

csx_status hip_fail(hipError_t e, const char* where)
{
    return fail(CSX_ERR_HIP, "%s: %s", where, hipGetErrorString(e));
}

#define CSX_HIP(call, where)                                 \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return hip_fail(e_, (where));  \
    } while (0)

constexpr size_t kAlign = 256;

size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

bool pow2(uint64_t v) { return v && !(v & (v - 1)); }

uint32_t log2u(uint64_t v)
{
    uint32_t r = 0;
    while ((1ull << r) < v) r++;
    return r;
}

csx_status check_chunk(uint32_t chunk_len)
{
    if (!pow2(chunk_len) || chunk_len < 32)
        return fail(CSX_ERR_INVALID, "chunk length %u must be a power of two >= 32", chunk_len);
    return CSX_OK;
}

csx_status check_width(uint32_t w)
{
    if (w != 1 && w != 2 && w != 4 && w != 8) return fail(CSX_ERR_INVALID, "element size %u must be 1, 2, 4 or 8", w);
    return CSX_OK;
}

csx_status check_order(int order)
{
    if (order != CSX_MSB_FIRST && order != CSX_LSB_FIRST) return fail(CSX_ERR_INVALID, "bit order %d is neither MSB- nor LSB-first", order);
    return CSX_OK;
}

// a configuration's workgroup size: 128, 256 or 512, or 0 for 256
csx_status check_block(uint32_t block)
{
    if (block != 0 && block != 128 && block != 256 && block != 512)
        return fail(CSX_ERR_INVALID, "workgroup size %u (128, 256 or 512; 0 for 256)", block);
    return CSX_OK;
}

csx_status check_n(uint64_t n)
{
    if (n > CSX_MAX_ELEMENTS) return fail(CSX_ERR_INVALID, "element count %llu exceeds the maximum %llu", (unsigned long long)n,
                                          (unsigned long long)CSX_MAX_ELEMENTS);
    return CSX_OK;
}

uint64_t tile_elems(uint32_t w)
{
    switch (w) {
        case 1: return csx::kTileElems<1>;
        case 2: return csx::kTileElems<2>;
        case 4: return csx::kTileElems<4>;
        default: return csx::kTileElems<8>;
    }
}

uint64_t chunk_count(uint64_t n, uint32_t chunk_len) { return (n + chunk_len - 1) / chunk_len; }

// Scan workspace: header | status[scan tiles] | u64[chunks] (tree scratch)
struct ScanPlan {
    uint64_t tiles;
    size_t status_off, tmp_off, total;
};

ScanPlan plan_scan(uint64_t chunks)
{
    ScanPlan p;
    p.tiles = (chunks + csx::kScanTileElems - 1) / csx::kScanTileElems;
    p.status_off = sizeof(csx::WsHeader);
    p.tmp_off = align_up(p.status_off + p.tiles * 8);
    p.total = align_up(p.tmp_off + chunks * 8);
    return p;
}

// Compaction workspace:
// header | status[max(fused tiles, scan tiles)] | u32 counts[chunks] |
// u64 offsets[chunks] | u32 tile counts[tiles] | u64 group sums[groups] |
// u32 unit offsets[units] | u32 block totals[blocks] | u64 block prefixes[blocks].
// The pipelines share the header and the status words.  Units (of the
// three-phase and sparse pipelines' popcount) are min(chunk, writeout tile)
// long; the smallest tile (128 threads x 2 vectors) is 4096 / w elements, so
// there are at most n / min(chunk, 4096 / w) of them.
struct CompactPlan {
    uint64_t fused_tiles, scan_tiles, chunks, groups, units;
    uint32_t log2g;
    size_t status_off, counts_off, offsets_off, tcounts_off, gsums_off, uoffs_off, btot_off, bpre_off, total;
};

CompactPlan plan_compact(uint64_t n, uint32_t w, uint32_t chunk_len)
{
    CompactPlan p;
    const uint64_t te = tile_elems(w);
    p.fused_tiles = (n + te - 1) / te;
    p.chunks = chunk_count(n, chunk_len);
    p.scan_tiles = (p.chunks + csx::kScanTileElems - 1) / csx::kScanTileElems;
    p.log2g = csx::two_pass_log2g(p.fused_tiles);
    p.groups = (p.fused_tiles + (1ull << p.log2g) - 1) >> p.log2g;
    const uint64_t st = p.fused_tiles > p.scan_tiles ? p.fused_tiles : p.scan_tiles;
    p.status_off = sizeof(csx::WsHeader);
    p.counts_off = align_up(p.status_off + st * 8);
    p.offsets_off = align_up(p.counts_off + p.chunks * 4);
    p.tcounts_off = align_up(p.offsets_off + p.chunks * 8);
    p.gsums_off = align_up(p.tcounts_off + p.fused_tiles * 4);
    const uint64_t umin = std::min<uint64_t>(chunk_len, 4096 / w);
    p.units = (n + umin - 1) / umin;
    p.uoffs_off = align_up(p.gsums_off + p.groups * 8);
    p.btot_off = align_up(p.uoffs_off + p.units * 4);
    p.bpre_off = align_up(p.btot_off + 4 * csx::popb_blocks(n));
    p.total = align_up(p.bpre_off + 8 * csx::popb_blocks(n));
    return p;
}

csx_status check_ws(const void* ws, size_t have, size_t need)
{
    if (!ws) return fail(CSX_ERR_WORKSPACE, "workspace pointer is null (%zu bytes needed)", need);
    if (((uintptr_t)ws & 15) != 0) return fail(CSX_ERR_WORKSPACE, "workspace must be 16-byte aligned");
    if (have < need) return fail(CSX_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu bytes needed", have, need);
    return CSX_OK;
}

// Whether the byte ranges [a, a + na) and [b, b + nb) share a byte.  The
// kernels never compact in place: a tile's output can land on input or mask
// bytes that another tile, running at the same time, has not read yet.
bool overlaps(const void* a, uint64_t na, const void* b, uint64_t nb)
{
    if (!a || !b || na == 0 || nb == 0) return false;
    const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
    return x >= y ? x - y < nb : y - x < na;
}

hipError_t dispatch_tile(uint32_t w, int mode, bool staged, int order, bool aligned, const csx::TileArgs& a, hipStream_t s, int vpt = 8,
                         int block = 256)
{
    switch (w) {
        case 1: return csx::launch_tile<1>(mode, staged, order, aligned, a, s, vpt, block);
        case 2: return csx::launch_tile<2>(mode, staged, order, aligned, a, s, vpt, block);
        case 4: return csx::launch_tile<4>(mode, staged, order, aligned, a, s, vpt, block);
        default: return csx::launch_tile<8>(mode, staged, order, aligned, a, s, vpt, block);
    }
}

// a configuration's workgroup size: 0 is the default, 256
uint32_t block_of(uint32_t b) { return b ? b : 256; }

hipError_t dispatch_tile_counts(uint32_t w, int order, bool aligned, const uint8_t* mask, uint64_t n, uint32_t* tcounts, uint64_t* gsums,
                                uint32_t log2g, csx::WsHeader* hdr, hipStream_t s)
{
    switch (w) {
        case 1: return csx::launch_tile_counts<1>(order, aligned, mask, n, tcounts, gsums, log2g, hdr, s);
        case 2: return csx::launch_tile_counts<2>(order, aligned, mask, n, tcounts, gsums, log2g, hdr, s);
        case 4: return csx::launch_tile_counts<4>(order, aligned, mask, n, tcounts, gsums, log2g, hdr, s);
        default: return csx::launch_tile_counts<8>(order, aligned, mask, n, tcounts, gsums, log2g, hdr, s);
    }
}

hipError_t dispatch_tpc(uint32_t w, int order, const csx::TileArgs& a, hipStream_t s)
{
    switch (w) {
        case 1: return csx::launch_thread_per_chunk<1>(order, a, s);
        case 2: return csx::launch_thread_per_chunk<2>(order, a, s);
        case 4: return csx::launch_thread_per_chunk<4>(order, a, s);
        default: return csx::launch_thread_per_chunk<8>(order, a, s);
    }
}

hipError_t dispatch_sparse(uint32_t w, int order, bool mask_aligned, const csx::TileArgs& a, hipStream_t s)
{
    switch (w) {
        case 1: return csx::launch_sparse<1>(order, mask_aligned, a, s);
        case 2: return csx::launch_sparse<2>(order, mask_aligned, a, s);
        case 4: return csx::launch_sparse<4>(order, mask_aligned, a, s);
        default: return csx::launch_sparse<8>(order, mask_aligned, a, s);
    }
}

// a grid of `tiles` workgroups of `threads` holds fewer than 2^32 work-items
// (the dispatch's 32-bit grid size)
bool launch_fits(uint64_t tiles, uint32_t threads = 256) { return tiles < (1ull << 32) && tiles * threads < (1ull << 32); }

bool known_pipeline(int p)
{
    return p == CSX_PIPELINE_FUSED || p == CSX_PIPELINE_THREE_PHASE || p == CSX_PIPELINE_TWO_PASS || p == CSX_PIPELINE_SPARSE || p == CSX_PIPELINE_SINGLE;
}

// ---- dominant-kernel timing (csx_profile_begin / csx_profile_end) ----
struct Profiler {
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
};
thread_local Profiler g_prof;
constexpr size_t kMaxProfiled = 1u << 16;

} // namespace
namespace csx {
thread_local LaunchEvents g_launch_events;
thread_local LaunchShape g_launch_shape;
}
namespace {

// `fn` launches the call's dominant kernel through launch_k, which records
// its grid and workgroup size for csx_last_launch (0 and 0 when it launches
// nothing).
template <typename F> hipError_t captured(F&& fn)
{
    csx::g_launch_shape = csx::LaunchShape{true, 0, 0};
    const hipError_t e = fn();
    csx::g_launch_shape.armed = false;
    return e;
}

// As captured(); when profiling, the launch's start and end are also
// recorded by the armed event pair.  A call that launches nothing (an empty
// input) leaves no pair.
template <typename F> hipError_t timed(hipStream_t s, F&& fn)
{
    (void)s;
    if (!g_prof.on || g_prof.events.size() >= kMaxProfiled) return captured(fn);
    hipEvent_t a, b;
    hipError_t e = hipEventCreate(&a);
    if (e != hipSuccess) return e;
    e = hipEventCreate(&b);
    if (e != hipSuccess) {
        (void)hipEventDestroy(a);
        return e;
    }
    csx::g_launch_events = csx::LaunchEvents{a, b, false};
    e = captured(fn);
    const bool used = csx::g_launch_events.used;
    csx::g_launch_events = csx::LaunchEvents{};
    if (used) {
        g_prof.events.emplace_back(a, b);
    }
    else {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    return e;
}

// When profiling: `fn` (several kernels) between an event pair recorded on
// the stream before and after it.
template <typename F> hipError_t timed_span(hipStream_t s, F&& fn)
{
    if (!g_prof.on || g_prof.events.size() >= kMaxProfiled) return fn();
    hipEvent_t a, b;
    hipError_t e = hipEventCreate(&a);
    if (e != hipSuccess) return e;
    e = hipEventCreate(&b);
    if (e != hipSuccess) {
        (void)hipEventDestroy(a);
        return e;
    }
    g_prof.events.emplace_back(a, b);
    if ((e = hipEventRecord(a, s)) != hipSuccess) return e;
    if ((e = fn()) != hipSuccess) return e;
    return hipEventRecord(b, s);
}

// ---- tuning table (csx_autotune, csx_tune_*) ----
struct TuneCell {
    uint32_t w;
    uint32_t log2n;
    double sel;
    csx_config cfg;
    double ms;
};

struct TuneTable {
    std::mutex mu;
    std::vector<TuneCell> cells;
    csx_config batch[9];  // by element width; vpt 0 = not tuned
};

TuneTable& tune_table()
{
    static TuneTable t;
    return t;
}

constexpr csx_config kDefaultConfig = {CSX_PIPELINE_THREE_PHASE, 1024, 8, 0, 256};

uint32_t log2_floor(uint64_t v)
{
    uint32_t r = 0;
    while (v >>= 1) r++;
    return r;
}

csx_config lookup_config(uint32_t w, uint64_t n, double sel)
{
    if (sel < 0) sel = 0.5;
    TuneTable& t = tune_table();
    std::lock_guard<std::mutex> g(t.mu);
    const uint32_t l = log2_floor(n ? n : 1);
    const TuneCell* best = nullptr;
    for (const TuneCell& c : t.cells) {
        if (c.w != w) continue;
        // the nearest cell may be smaller: never past the single launch's cap
        if (c.cfg.pipeline == CSX_PIPELINE_SINGLE && n > CSX_SINGLE_MAX_ELEMENTS) continue;
        if (!best) {
            best = &c;
            continue;
        }
        const uint32_t dl = c.log2n > l ? c.log2n - l : l - c.log2n, bl = best->log2n > l ? best->log2n - l : l - best->log2n;
        if (dl < bl || (dl == bl && std::fabs(c.sel - sel) < std::fabs(best->sel - sel))) best = &c;
    }
    return best ? best->cfg : kDefaultConfig;
}

void record_cell(uint32_t w, uint64_t n, double sel, const csx_config& cfg, double ms)
{
    TuneTable& t = tune_table();
    std::lock_guard<std::mutex> g(t.mu);
    const uint32_t l = log2_floor(n ? n : 1);
    for (TuneCell& c : t.cells)
        if (c.w == w && c.log2n == l && c.sel == sel) {
            c.cfg = cfg;
            c.ms = ms;
            return;
        }
    t.cells.push_back(TuneCell{w, l, sel, cfg, ms});
}

csx_config batch_config(uint32_t w)
{
    TuneTable& t = tune_table();
    std::lock_guard<std::mutex> g(t.mu);
    if (w < 9 && t.batch[w].vpt) return t.batch[w];
    return csx_config{CSX_PIPELINE_THREE_PHASE, 0, 8, 4096, 256};
}

// csx_predicate -> one range test; returns the element size (0 if invalid)
uint32_t pred_spec(const csx_predicate* q, csx::PredSpec* p)
{
    static const uint32_t sizes[] = {1, 2, 4, 8, 1, 2, 4, 8, 4, 8};
    if (q->type < CSX_TYPE_U8 || q->type > CSX_TYPE_F64) return 0;
    p->type = q->type;
    p->lo = p->hi = q->a.u64;
    using namespace csx;
    switch (q->op) {
        case CSX_PRED_GT: p->flags = kPredLo; break;
        case CSX_PRED_GE: p->flags = kPredLo | kPredLoIncl; break;
        case CSX_PRED_LT: p->flags = kPredHi; break;
        case CSX_PRED_LE: p->flags = kPredHi | kPredHiIncl; break;
        case CSX_PRED_EQ: p->flags = kPredLo | kPredLoIncl | kPredHi | kPredHiIncl; break;
        case CSX_PRED_NE: p->flags = kPredLo | kPredLoIncl | kPredHi | kPredHiIncl | kPredNegate; break;
        case CSX_PRED_BETWEEN:
            p->flags = kPredLo | kPredLoIncl | kPredHi | kPredHiIncl;
            p->hi = q->b.u64;
            break;
        default: return 0;
    }
    return sizes[q->type];
}

csx_status check_pred(const csx_predicate* q, csx::PredSpec* p, uint32_t* w)
{
    if (!q) return fail(CSX_ERR_INVALID, "null predicate");
    if (q->type < CSX_TYPE_U8 || q->type > CSX_TYPE_F64) return fail(CSX_ERR_INVALID, "unknown element type %d", q->type);
    if (!(*w = pred_spec(q, p))) return fail(CSX_ERR_INVALID, "unknown comparison %d", q->op);
    return CSX_OK;
}

// predicate compaction workspace: the csx_compact plan at chunk 1024 | mask
constexpr uint32_t kPredChunk = 1024;
size_t pred_mask_off(uint64_t n, uint32_t w) { return align_up(plan_compact(n, w, kPredChunk).total); }
size_t pred_ws_total(uint64_t n, uint32_t w) { return pred_mask_off(n, w) + align_up((n + 7) / 8); }

} // namespace

namespace csx {
csx_status set_error(csx_status s, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vfail(s, fmt, ap);
    va_end(ap);
    return s;
}
csx_status clear_error()
{
    g_last_error.clear();
    return CSX_OK;
}
} // namespace csx

extern "C" {

const char* csx_status_string(csx_status s)
{
    switch (s) {
        case CSX_OK: return "ok";
        case CSX_ERR_INVALID: return "invalid argument";
        case CSX_ERR_WORKSPACE: return "workspace missing or too small";
        case CSX_ERR_HIP: return "HIP runtime error";
        case CSX_ERR_UNSUPPORTED: return "unsupported request";
        case CSX_ERR_DEVICE: return "device-side failure";
    }
    return "unknown status";
}

const char* csx_last_error(void) { return g_last_error.c_str(); }

int csx_version(void) { return 100; }

uint64_t csx_chunk_count(uint64_t n, uint32_t chunk_len) { return chunk_len ? chunk_count(n, chunk_len) : 0; }

csx_status csx_popc(const uint8_t* d_mask, uint32_t* d_counts, uint64_t n, uint32_t chunk_len, int bit_order, hipStream_t stream)
{
    csx_status s;
    if ((s = check_chunk(chunk_len)) || (s = check_order(bit_order)) || (s = check_n(n))) return s;
    if (n == 0) return ok();
    if (!d_mask || !d_counts) return fail(CSX_ERR_INVALID, "null mask or counts pointer");
    CSX_HIP(csx::launch_popc(d_mask, d_counts, n, log2u(chunk_len), bit_order, stream), "csx_popc launch");
    return ok();
}

csx_status csx_scan_workspace_size(uint64_t chunks, size_t* bytes)
{
    if (!bytes) return fail(CSX_ERR_INVALID, "null size pointer");
    *bytes = plan_scan(chunks).total;
    return ok();
}

csx_status csx_scan(const uint32_t* d_counts, uint64_t* d_offsets, uint64_t chunks, uint64_t* d_total, void* d_ws, size_t ws_bytes,
                    hipStream_t stream)
{
    if (chunks == 0) {
        if (d_total) CSX_HIP(csx::launch_store_u64(d_total, 0, stream), "csx_scan total");
        return ok();
    }
    if (!d_counts || !d_offsets) return fail(CSX_ERR_INVALID, "null counts or offsets pointer");
    const ScanPlan p = plan_scan(chunks);
    csx_status s;
    if ((s = check_ws(d_ws, ws_bytes, p.total))) return s;
    auto base = static_cast<uint8_t*>(d_ws);
    CSX_HIP(csx::launch_scan(d_counts, d_offsets, chunks, d_total, reinterpret_cast<csx::WsHeader*>(base),
                             reinterpret_cast<uint64_t*>(base + p.status_off), stream),
            "csx_scan launch");
    return ok();
}

csx_status csx_scan_u32(const uint32_t* d_counts, uint32_t* d_offsets, uint64_t chunks, uint64_t* d_total, void* d_ws, size_t ws_bytes,
                        hipStream_t stream)
{
    if (chunks == 0) {
        if (d_total) CSX_HIP(csx::launch_store_u64(d_total, 0, stream), "csx_scan_u32 total");
        return ok();
    }
    if (!d_counts || !d_offsets) return fail(CSX_ERR_INVALID, "null counts or offsets pointer");
    const ScanPlan p = plan_scan(chunks);
    csx_status s;
    if ((s = check_ws(d_ws, ws_bytes, p.total))) return s;
    auto base = static_cast<uint8_t*>(d_ws);
    CSX_HIP(csx::launch_scan32(d_counts, d_offsets, chunks, d_total, reinterpret_cast<csx::WsHeader*>(base),
                               reinterpret_cast<uint64_t*>(base + p.status_off), stream),
            "csx_scan_u32 launch");
    return ok();
}

csx_status csx_scan_tree(uint32_t* d_pss,uint64_t chunks, uint64_t* d_total, void* d_ws, size_t ws_bytes, hipStream_t stream)
{
    if (chunks == 0) {
        if (d_total) CSX_HIP(csx::launch_store_u64(d_total, 0, stream), "csx_scan_tree total");
        return ok();
    }
    if (!d_pss) return fail(CSX_ERR_INVALID, "null pss pointer");
    const ScanPlan p = plan_scan(chunks);
    csx_status s;
    if ((s = check_ws(d_ws, ws_bytes, p.total))) return s;
    auto base = static_cast<uint8_t*>(d_ws);
    auto ex = reinterpret_cast<uint64_t*>(base + p.tmp_off);
    CSX_HIP(csx::launch_scan(d_pss, ex, chunks, d_total, reinterpret_cast<csx::WsHeader*>(base), reinterpret_cast<uint64_t*>(base + p.status_off),
                             stream),
            "csx_scan_tree scan");
    CSX_HIP(csx::launch_tree_from_offsets(d_pss, ex, chunks, stream), "csx_scan_tree tree");
    return ok();
}

csx_status csx_tree_offsets(const uint32_t* d_tree, uint32_t* d_offsets, uint64_t chunks, hipStream_t stream)
{
    if (chunks == 0) return ok();
    if (!d_tree || !d_offsets) return fail(CSX_ERR_INVALID, "null tree or offsets pointer");
    if (chunks > (1ull << 40)) return fail(CSX_ERR_INVALID, "too many chunks");
    CSX_HIP(csx::launch_tree_offsets(d_tree, d_offsets, chunks, stream), "csx_tree_offsets launch");
    return ok();
}

csx_status csx_writeout_ex(const csx_writeout_args* wa, hipStream_t stream)
{
    if (!wa) return fail(CSX_ERR_INVALID, "null argument block");
    const uint32_t elem_size = wa->elem_size, chunk_len = wa->chunk_len;
    const uint64_t n = wa->n;
    const int variant = wa->variant, bit_order = wa->bit_order;
    csx_status s;
    if ((s = check_width(elem_size)) || (s = check_chunk(chunk_len)) || (s = check_order(bit_order)) || (s = check_n(n))) return s;
    if (variant < CSX_WRITEOUT_STAGED || variant > CSX_WRITEOUT_SPARSE) return fail(CSX_ERR_INVALID, "unknown writeout variant %d", variant);
    if (n == 0) return ok();
    if (!wa->in || !wa->mask || !wa->offsets || !wa->out) return fail(CSX_ERR_INVALID, "null input, mask, offsets or output pointer");
    if (overlaps(wa->out, n * elem_size, wa->in, n * elem_size) || overlaps(wa->out, n * elem_size, wa->mask, (n + 7) / 8))
        return fail(CSX_ERR_INVALID, "the output overlaps the input or the mask (the writeout is not in place)");
    if (((uintptr_t)wa->in % elem_size) || ((uintptr_t)wa->out % elem_size))
        return fail(CSX_ERR_INVALID, "input and output must be aligned to the element size");
    csx::TileArgs a{};
    a.in = static_cast<const uint8_t*>(wa->in);
    a.mask = wa->mask;
    a.out = static_cast<uint8_t*>(wa->out);
    a.n = n;
    if (wa->offsets_u32)
        a.offsets32 = static_cast<const uint32_t*>(wa->offsets);
    else
        a.offsets = static_cast<const uint64_t*>(wa->offsets);
    a.counts = wa->counts;
    a.log2c = log2u(chunk_len);
    a.out_base = wa->out_base;
    if (variant == CSX_WRITEOUT_THREAD_PER_CHUNK) {
        CSX_HIP(timed(stream, [&] { return dispatch_tpc(elem_size, bit_order, a, stream); }), "csx_writeout launch");
        return ok();
    }
    if (variant == CSX_WRITEOUT_SPARSE) {
        const bool mask_aligned = ((uintptr_t)wa->mask & 15) == 0;
        CSX_HIP(timed(stream, [&] { return dispatch_sparse(elem_size, bit_order, mask_aligned, a, stream); }), "csx_writeout launch");
        return ok();
    }
    if ((s = check_block(wa->block_size))) return s;
    const uint32_t block = block_of(wa->block_size);
    if (block != 256 && variant != CSX_WRITEOUT_STAGED) return fail(CSX_ERR_INVALID, "only the staged writeout takes a workgroup size other than 256");
    // the staged tile: block threads x 8 vectors
    const uint64_t te = (uint64_t)block * 8 * (16 / elem_size);
    const uint64_t tiles = (n + te - 1) / te;
    if (!launch_fits(tiles, block)) return fail(CSX_ERR_UNSUPPORTED, "%llu tiles exceed one launch", (unsigned long long)tiles);
    const bool aligned = (((uintptr_t)wa->in | (uintptr_t)wa->mask) & 15) == 0;
    CSX_HIP(timed(stream,
                  [&] { return dispatch_tile(elem_size, csx::kModeWriteout, variant == CSX_WRITEOUT_STAGED, bit_order, aligned, a, stream, 8, (int)block); }),
            "csx_writeout launch");
    return ok();
}

csx_status csx_writeout(const void* d_in, const uint8_t* d_mask, const uint64_t* d_offsets, const uint32_t* d_counts, void* d_out, uint64_t n,
                        uint32_t elem_size, uint32_t chunk_len, int bit_order, int variant, hipStream_t stream)
{
    csx_writeout_args wa{};
    wa.in = d_in;
    wa.mask = d_mask;
    wa.out = d_out;
    wa.offsets = d_offsets;
    wa.offsets_u32 = 0;
    wa.counts = d_counts;
    wa.out_base = nullptr;
    wa.n = n;
    wa.elem_size = elem_size;
    wa.chunk_len = chunk_len;
    wa.bit_order = bit_order;
    wa.variant = variant;
    return csx_writeout_ex(&wa, stream);
}

#define CSX_WRITEOUT_W(W, T)                                                                                                                  \
    csx_status csx_writeout_w##W(const T* d_in, const uint8_t* d_mask, const uint64_t* d_offsets, const uint32_t* d_counts, T* d_out,       \
                                 uint64_t n, uint32_t chunk_len, int bit_order, int variant, hipStream_t stream)                              \
    {                                                                                                                                         \
        return csx_writeout(d_in, d_mask, d_offsets, d_counts, d_out, n, W, chunk_len, bit_order, variant, stream);                           \
    }
CSX_WRITEOUT_W(1, uint8_t)
CSX_WRITEOUT_W(2, uint16_t)
CSX_WRITEOUT_W(4, uint32_t)
CSX_WRITEOUT_W(8, uint64_t)

csx_status csx_workspace_size(uint64_t n, uint32_t elem_size, uint32_t chunk_len, size_t* bytes)
{
    csx_status s;
    if ((s = check_width(elem_size)) || (s = check_chunk(chunk_len)) || (s = check_n(n))) return s;
    if (!bytes) return fail(CSX_ERR_INVALID, "null size pointer");
    *bytes = plan_compact(n, elem_size, chunk_len).total;
    return ok();
}

static csx_status finish_host_count(uint64_t* h_count, const uint64_t* d_count_src, const csx::WsHeader* hdr, hipStream_t stream)
{
    uint32_t err = 0;
    CSX_HIP(hipMemcpyAsync(h_count, d_count_src, sizeof(uint64_t), hipMemcpyDeviceToHost, stream), "count copy");
    if (hdr) CSX_HIP(hipMemcpyAsync(&err, &hdr->error, sizeof(uint32_t), hipMemcpyDeviceToHost, stream), "error copy");
    CSX_HIP(hipStreamSynchronize(stream), "stream synchronize");
    if (err) return fail(CSX_ERR_DEVICE, "a look-back wait gave up (error word %u)", err);
    return ok();
}

static csx_status check_vpt(uint32_t vpt, uint32_t elem_size)
{
    if (vpt != 2 && vpt != 4 && vpt != 8 && !(vpt == 16 && elem_size >= 2))
        return fail(CSX_ERR_INVALID, "tile of %u vectors per thread is not built for %u-byte elements (2, 4, 8; 16 from 2 bytes)", vpt, elem_size);
    return CSX_OK;
}


// the tile shapes a pipeline is built for: the three-phase writeout takes
// every tile size and workgroup size, the single launch tiles of 2, 4 or 8
// vectors per 256-thread workgroup, the others the 8-vector, 256-thread tile;
// null when (vpt, block) is one of them
static const char* pipeline_shape_error(int pipeline, uint32_t vpt, uint32_t block)
{
    block = block_of(block);
    if (pipeline == CSX_PIPELINE_THREE_PHASE) return nullptr;
    if (block != 256) return "only the three-phase writeout takes a workgroup size other than 256";
    if (pipeline == CSX_PIPELINE_SINGLE) return vpt == 16 ? "the single-launch pipeline takes tiles of 2, 4 or 8 vectors per thread" : nullptr;
    return vpt != 8 ? "only the three-phase and single-launch pipelines take a tile size other than 8" : nullptr;
}

static csx_status compact_impl(const void* d_in, const uint8_t* d_mask, void* d_out, uint64_t n, uint32_t elem_size, uint32_t chunk_len,
                               int bit_order, int pipeline, uint32_t vpt, uint32_t block, uint64_t* d_count, uint64_t* h_count, void* d_ws,
                               size_t ws_bytes, hipStream_t stream, bool profile = true)
{
    // the dominant kernel between csx_profile_begin's event pairs (creating
    // the events is the only resource this path acquires; `profile` = false
    // keeps it out)
    auto timed = [&](hipStream_t st, auto&& fn) -> hipError_t { return profile ? ::timed(st, fn) : captured(fn); };
    csx_status s;
    if ((s = check_width(elem_size)) || (s = check_chunk(chunk_len)) || (s = check_order(bit_order)) || (s = check_n(n))) return s;
    if (!known_pipeline(pipeline)) return fail(CSX_ERR_INVALID, "unknown pipeline %d", pipeline);
    if (pipeline == CSX_PIPELINE_SINGLE && n > CSX_SINGLE_MAX_ELEMENTS)
        return fail(CSX_ERR_UNSUPPORTED, "the single-launch pipeline takes at most %llu elements", (unsigned long long)CSX_SINGLE_MAX_ELEMENTS);
    if ((s = check_vpt(vpt, elem_size)) || (s = check_block(block))) return s;
    if (const char* why = pipeline_shape_error(pipeline, vpt, block)) return fail(CSX_ERR_INVALID, "%s", why);
    block = block_of(block);
    const CompactPlan p = plan_compact(n, elem_size, chunk_len);
    if ((s = check_ws(d_ws, ws_bytes, p.total))) return s;
    if (n > 0 && (!d_in || !d_mask || !d_out)) return fail(CSX_ERR_INVALID, "null input, mask or output pointer");
    if (overlaps(d_out, n * elem_size, d_in, n * elem_size) || overlaps(d_out, n * elem_size, d_mask, (n + 7) / 8))
        return fail(CSX_ERR_INVALID, "the output overlaps the input or the mask (compaction is not in place)");
    if (((uintptr_t)d_in % elem_size) || ((uintptr_t)d_out % elem_size))
        return fail(CSX_ERR_INVALID, "input and output must be aligned to the element size");
    // the writeout's grid is one workgroup per tile of the chosen size
    const uint64_t wte = (uint64_t)block * vpt * (16 / elem_size);
    const uint64_t wtiles = (n + wte - 1) / wte;
    if (!launch_fits(p.fused_tiles) || !launch_fits(wtiles, block))
        return fail(CSX_ERR_UNSUPPORTED, "%llu tiles exceed one launch", (unsigned long long)std::max(p.fused_tiles, wtiles));

    auto base = static_cast<uint8_t*>(d_ws);
    auto hdr = reinterpret_cast<csx::WsHeader*>(base);
    auto status = reinterpret_cast<uint64_t*>(base + p.status_off);
    uint64_t* count_dst = d_count ? d_count : &hdr->count;
    auto in = static_cast<const uint8_t*>(d_in);
    auto out = static_cast<uint8_t*>(d_out);
    const bool aligned = (((uintptr_t)d_in | (uintptr_t)d_mask) & 15) == 0;

    if (n == 0) {
        CSX_HIP(csx::launch_lookback_init(hdr, status, 0, stream), "csx_compact init");
        CSX_HIP(csx::launch_store_u64(count_dst, 0, stream), "csx_compact count");
    }
    else if (n <= 256ull * vpt * (16 / elem_size)) {
        // the whole input is one tile of the configured size (256 threads x
        // vpt vectors): one workgroup, output offset 0, every pipeline alike
        csx::TileArgs a{};
        a.in = in;
        a.mask = d_mask;
        a.out = out;
        a.n = n;
        a.count_out = count_dst;
        a.given_off = 0;
        CSX_HIP(timed(stream, [&] { return dispatch_tile(elem_size, csx::kModeGiven, true, bit_order, aligned, a, stream, (int)vpt); }),
                "csx_compact single tile");
        if (h_count) return finish_host_count(h_count, count_dst, nullptr, stream);
        return ok();
    }
    else if (pipeline == CSX_PIPELINE_FUSED) {
        CSX_HIP(csx::launch_lookback_init(hdr, status, p.fused_tiles, stream), "csx_compact init");
        csx::TileArgs a{};
        a.in = in;
        a.mask = d_mask;
        a.out = out;
        a.n = n;
        a.hdr = hdr;
        a.status = status;
        a.count_out = count_dst;
        CSX_HIP(timed(stream, [&] { return dispatch_tile(elem_size, csx::kModeFused, true, bit_order, aligned, a, stream); }),
                "csx_compact fused");
    }
    else if (pipeline == CSX_PIPELINE_SINGLE) {
        csx::TileArgs a{};
        a.in = in;
        a.mask = d_mask;
        a.out = out;
        a.n = n;
        a.count_out = count_dst;
        a.hdr = hdr;
        CSX_HIP(timed(stream, [&] { return dispatch_tile(elem_size, csx::kModeSingle, true, bit_order, aligned, a, stream, (int)vpt); }),
                "csx_compact single launch");
    }
    else if (pipeline == CSX_PIPELINE_TWO_PASS) {
        auto tcounts = reinterpret_cast<uint32_t*>(base + p.tcounts_off);
        auto gsums = reinterpret_cast<uint64_t*>(base + p.gsums_off);
        CSX_HIP(dispatch_tile_counts(elem_size, bit_order, aligned, d_mask, n, tcounts, gsums, p.log2g, hdr, stream), "csx_compact tile counts");
        csx::TileArgs a{};
        a.in = in;
        a.mask = d_mask;
        a.out = out;
        a.n = n;
        a.count_out = count_dst;
        a.tcounts = tcounts;
        a.gsums = gsums;
        a.log2g = p.log2g;
        CSX_HIP(timed(stream, [&] { return dispatch_tile(elem_size, csx::kModeTwoPass, true, bit_order, aligned, a, stream); }),
                "csx_compact two-pass writeout");
    }
    else {
        // phases 1 + 2: per block of the column, its selected count and the
        // offsets of its units (min(chunk, writeout tile) elements) within
        // it; then the blocks' exclusive prefixes and the total count
        auto uoffs = reinterpret_cast<uint32_t*>(base + p.uoffs_off);
        auto btotals = reinterpret_cast<uint32_t*>(base + p.btot_off);
        auto bprefix = reinterpret_cast<uint64_t*>(base + p.bpre_off);
        const uint64_t te = pipeline == CSX_PIPELINE_SPARSE ? csx::kSparseStepElems : wte;
        const uint32_t log2unit = std::min(log2u(chunk_len), log2u(te));
        // A column of few blocks skips the block prefixes: each writeout
        // wave (tile or sparse step) sums the totals before its block, and
        // the last tile or step writes the count.  Blocks grow up to 2^19
        // elements (one round of a 1024-thread popcount workgroup) to stay
        // within kMidBlocks: up to 2^26 elements.
        uint32_t log2b = csx::popb_log2(n);
        bool few_blocks = false;
        if (n < (1ull << 32)) {
            while (log2b < csx::kPopbMinLog2 + 2 && ((n + (1ull << log2b) - 1) >> log2b) > csx::kMidBlocks) log2b++;
            few_blocks = ((n + (1ull << log2b) - 1) >> log2b) <= csx::kMidBlocks;
            if (!few_blocks) log2b = csx::popb_log2(n);
        }
        CSX_HIP(csx::launch_popc_blocks(d_mask, n, log2b, log2unit, uoffs, btotals, bit_order, stream, hdr), "csx_compact popc");
        if (!few_blocks)
            CSX_HIP(csx::launch_block_prefix(btotals, csx::popb_blocks(n), bprefix, count_dst, n >= (1ull << 32), stream), "csx_compact block prefixes");
        csx::TileArgs a{};
        a.in = in;
        a.mask = d_mask;
        a.out = out;
        a.n = n;
        a.bprefix = few_blocks ? nullptr : bprefix;
        a.btotals = few_blocks ? btotals : nullptr;
        a.count_out = few_blocks ? count_dst : nullptr;
        a.uoffs = uoffs;
        a.log2b = log2b;
        a.log2c = log2unit;
        if (pipeline == CSX_PIPELINE_SPARSE) {
            const bool mask_aligned = ((uintptr_t)d_mask & 15) == 0;
            CSX_HIP(timed(stream, [&] { return dispatch_sparse(elem_size, bit_order, mask_aligned, a, stream); }), "csx_compact sparse writeout");
        }
        else {
            const int mode = few_blocks ? csx::kModeFewBlocks : csx::kModeWriteout;
            CSX_HIP(timed(stream, [&] { return dispatch_tile(elem_size, mode, true, bit_order, aligned, a, stream, (int)vpt, (int)block); }),
                    "csx_compact writeout");
        }
    }
    if (h_count) return finish_host_count(h_count, count_dst, hdr, stream);
    return ok();
}

csx_status csx_compact(const void* d_in, const uint8_t* d_mask, void* d_out, uint64_t n, uint32_t elem_size, uint32_t chunk_len, int bit_order,
                       int pipeline, uint64_t* d_count, uint64_t* h_count, void* d_ws, size_t ws_bytes, hipStream_t stream)
{
    return compact_impl(d_in, d_mask, d_out, n, elem_size, chunk_len, bit_order, pipeline, 8, 256, d_count, h_count, d_ws, ws_bytes, stream);
}

csx_status csx_compact_config(const void* d_in, const uint8_t* d_mask, void* d_out, uint64_t n, uint32_t elem_size, int bit_order,
                              const csx_config* cfg, uint64_t* d_count, uint64_t* h_count, void* d_ws, size_t ws_bytes, hipStream_t stream)
{
    if (!cfg) return fail(CSX_ERR_INVALID, "null configuration");
    return compact_impl(d_in, d_mask, d_out, n, elem_size, cfg->chunk_len, bit_order, cfg->pipeline, cfg->vpt, cfg->block_size, d_count, h_count,
                        d_ws, ws_bytes, stream);
}

} // extern "C"

namespace csx {
csx_status compact_in_workspace(const void* d_in, const uint8_t* d_mask, void* d_out, uint64_t n, uint32_t elem_size, int bit_order,
                                const csx_config* cfg, uint64_t* d_count, uint64_t* h_count, void* d_ws, size_t ws_bytes, hipStream_t stream)
{
    if (!cfg) return fail(CSX_ERR_INVALID, "null configuration");
    return compact_impl(d_in, d_mask, d_out, n, elem_size, cfg->chunk_len, bit_order, cfg->pipeline, cfg->vpt, cfg->block_size, d_count, h_count,
                        d_ws, ws_bytes, stream, false);
}
} // namespace csx

extern "C" {

#define CSX_COMPACT_W(W, T)                                                                                                                   \
    csx_status csx_compact_w##W(const T* d_in, const uint8_t* d_mask, T* d_out, uint64_t n, uint32_t chunk_len, int bit_order, int pipeline, \
                                uint64_t* d_count, uint64_t* h_count, void* d_ws, size_t ws_bytes, hipStream_t stream)                       \
    {                                                                                                                                         \
        return csx_compact(d_in, d_mask, d_out, n, W, chunk_len, bit_order, pipeline, d_count, h_count, d_ws, ws_bytes, stream);              \
    }
CSX_COMPACT_W(1, uint8_t)
CSX_COMPACT_W(2, uint16_t)
CSX_COMPACT_W(4, uint32_t)
CSX_COMPACT_W(8, uint64_t)

csx_status csx_workspace_error(const void* d_ws, uint32_t* h_error, hipStream_t stream)
{
    if (!d_ws || !h_error) return fail(CSX_ERR_INVALID, "null workspace or result pointer");
    auto hdr = static_cast<const csx::WsHeader*>(d_ws);
    CSX_HIP(hipMemcpyAsync(h_error, &hdr->error, sizeof(uint32_t), hipMemcpyDeviceToHost, stream), "error copy");
    CSX_HIP(hipStreamSynchronize(stream), "stream synchronize");
    return ok();
}

csx_status csx_validate(const void* d_expected, const void* d_actual, uint64_t len, uint32_t elem_size, uint64_t* d_failures, hipStream_t stream)
{
    csx_status s;
    if ((s = check_width(elem_size))) return s;
    if (len == 0) return ok();
    if (!d_expected || !d_actual || !d_failures) return fail(CSX_ERR_INVALID, "null pointer");
    CSX_HIP(csx::launch_validate(d_expected, d_actual, len, elem_size, d_failures, stream), "csx_validate launch");
    return ok();
}

csx_status csx_pattern_count(uint64_t n, uint32_t pattern, uint32_t pattern_length, uint64_t* count)
{
    if (pattern_length < 1 || pattern_length > 32) return fail(CSX_ERR_INVALID, "pattern length %u must be in [1, 32]", pattern_length);
    if (!count) return fail(CSX_ERR_INVALID, "null count pointer");
    uint64_t per = 0, head = 0;
    const uint64_t rem = n % pattern_length;
    for (uint32_t p = 0; p < pattern_length; p++) {
        if ((pattern >> (31 - p)) & 1u) {
            per++;
            if (p < rem) head++;
        }
    }
    *count = (n / pattern_length) * per + head;
    return ok();
}

csx_status csx_pattern_compact(const void* d_in, void* d_out, uint64_t n, uint32_t elem_size, uint32_t pattern, uint32_t pattern_length,
                               uint64_t* d_count, uint64_t* h_count, hipStream_t stream)
{
    csx_status s;
    uint64_t k = 0;
    if ((s = check_width(elem_size)) || (s = check_n(n)) || (s = csx_pattern_count(n, pattern, pattern_length, &k))) return s;
    if (k > 0 && (!d_in || !d_out)) return fail(CSX_ERR_INVALID, "null input or output pointer");
    // the count is known here, and nothing at or past d_out[k] is written
    if (k > 0 && overlaps(d_out, k * elem_size, d_in, n * elem_size))
        return fail(CSX_ERR_INVALID, "the output overlaps the input (compaction is not in place)");
    if (((uintptr_t)d_in % elem_size) || ((uintptr_t)d_out % elem_size))
        return fail(CSX_ERR_INVALID, "input and output must be aligned to the element size");
    // tile size, measured at 2^29 on 32-bit patterns of 12.5 % and 62.5 %
    // density: 4 vectors per thread for 1- and 2-byte elements; for 4- and
    // 8-byte ones 16 from half density up (4 % faster there), else 2 (2-7 % faster)
    uint32_t pop = 0;
    for (uint32_t p = 0; p < pattern_length; p++) pop += (pattern >> (31 - p)) & 1u;
    const uint32_t vpt = elem_size <= 2 ? 4 : (2 * pop >= pattern_length ? 16 : 2);
    const uint64_t te = 256ull * vpt * (16 / elem_size);
    const uint64_t tiles = (n + te - 1) / te;
    if (!launch_fits(tiles)) return fail(CSX_ERR_UNSUPPORTED, "%llu tiles exceed one launch", (unsigned long long)tiles);
    if (k == 0) {
        if (d_count) CSX_HIP(csx::launch_store_u64(d_count, 0, stream), "csx_pattern_compact count");
    }
    else {
        // the tile engine with mask words generated from the pattern and each
        // tile's offset in closed form: one kernel, no mask, no scan
        csx::TileArgs a{};
        a.in = static_cast<const uint8_t*>(d_in);
        a.out = static_cast<uint8_t*>(d_out);
        a.n = n;
        a.count_out = d_count;
        a.pat_len = pattern_length;
        for (uint32_t b = 0; b < 64; b++)
            if ((pattern >> (31 - b % pattern_length)) & 1u) a.pat_bits |= 1ull << b;
        for (uint32_t p = 0; p < pattern_length; p++) a.pat_pop += (pattern >> (31 - p)) & 1u;
        const bool aligned = ((uintptr_t)d_in & 15) == 0;
        CSX_HIP(timed(stream, [&] { return dispatch_tile(elem_size, csx::kModePattern, true, CSX_MSB_FIRST, aligned, a, stream, (int)vpt); }),
                "csx_pattern_compact");
    }
    if (h_count) *h_count = k;
    return ok();
}

// ---------------------------------------------------------------- multi-stream split

// Slices of max(tile, chunk) elements; per slice: its own scan header and
// status words; shared: u64 base[S], u64 total[S], u32 counts and u64
// offsets over all chunks (each slice uses its own range).
struct StreamsPlan {
    uint32_t S;
    uint64_t unit, units, per_slice;
    size_t base_off, total_off, counts_off, offsets_off, scan_off, scan_stride, total;
    uint64_t begin(uint32_t i, uint64_t n) const
    {
        const uint64_t b = (uint64_t)i * per_slice * unit;
        return b < n ? b : n;
    }
    uint64_t end(uint32_t i, uint64_t n) const { return i + 1 == S ? n : begin(i + 1, n); }
};

static StreamsPlan plan_streams(uint64_t n, uint32_t w, uint32_t chunk_len, uint32_t S)
{
    StreamsPlan p;
    p.S = S;
    p.unit = tile_elems(w) > chunk_len ? tile_elems(w) : chunk_len;
    p.units = (n + p.unit - 1) / p.unit;
    p.per_slice = p.units / S;
    const uint64_t chunks = chunk_count(n, chunk_len);
    // largest slice: the last one (it takes the remainder)
    const uint64_t max_slice_chunks = (n - (uint64_t)(S - 1) * p.per_slice * p.unit + chunk_len - 1) / chunk_len;
    p.base_off = 0;
    p.total_off = align_up(8ull * S);
    p.counts_off = align_up(p.total_off + 8ull * S);
    p.offsets_off = align_up(p.counts_off + 4 * chunks);
    p.scan_off = align_up(p.offsets_off + 8 * chunks);
    p.scan_stride = align_up(sizeof(csx::WsHeader) + 8 * csx::scan_tiles(max_slice_chunks + 1));
    p.total = p.scan_off + (size_t)S * p.scan_stride;
    return p;
}

csx_status csx_streams_workspace_size(uint64_t n, uint32_t elem_size, uint32_t chunk_len, uint32_t stream_count, size_t* bytes)
{
    csx_status s;
    if ((s = check_width(elem_size)) || (s = check_chunk(chunk_len)) || (s = check_n(n))) return s;
    if (stream_count < 1 || stream_count > 64) return fail(CSX_ERR_INVALID, "stream count %u must be in [1, 64]", stream_count);
    if (!bytes) return fail(CSX_ERR_INVALID, "null size pointer");
    *bytes = plan_streams(n, elem_size, chunk_len, stream_count).total;
    return ok();
}

csx_status csx_compact_streams(const void* d_in, const uint8_t* d_mask, void* d_out, uint64_t n, uint32_t elem_size, uint32_t chunk_len,
                               int bit_order, const hipStream_t* streams, uint32_t stream_count, uint64_t* d_count, uint64_t* h_count, void* d_ws,
                               size_t ws_bytes)
{
    csx_status s;
    if ((s = check_width(elem_size)) || (s = check_chunk(chunk_len)) || (s = check_order(bit_order)) || (s = check_n(n))) return s;
    if (stream_count < 1 || stream_count > 64 || !streams) return fail(CSX_ERR_INVALID, "need 1 to 64 streams");
    const StreamsPlan p = plan_streams(n, elem_size, chunk_len, stream_count);
    if ((s = check_ws(d_ws, ws_bytes, p.total))) return s;
    if (n > 0 && (!d_in || !d_mask || !d_out)) return fail(CSX_ERR_INVALID, "null input, mask or output pointer");
    if (overlaps(d_out, n * elem_size, d_in, n * elem_size) || overlaps(d_out, n * elem_size, d_mask, (n + 7) / 8))
        return fail(CSX_ERR_INVALID, "the output overlaps the input or the mask (compaction is not in place)");
    if (((uintptr_t)d_in % elem_size) || ((uintptr_t)d_out % elem_size))
        return fail(CSX_ERR_INVALID, "input and output must be aligned to the element size");
    auto base = static_cast<uint8_t*>(d_ws);
    auto bases = reinterpret_cast<uint64_t*>(base + p.base_off);
    auto totals = reinterpret_cast<uint64_t*>(base + p.total_off);
    auto counts = reinterpret_cast<uint32_t*>(base + p.counts_off);
    auto offsets = reinterpret_cast<uint64_t*>(base + p.offsets_off);
    const uint32_t log2c = log2u(chunk_len);
    auto in = static_cast<const uint8_t*>(d_in);
    auto out = static_cast<uint8_t*>(d_out);

    std::vector<hipEvent_t> carried(stream_count, nullptr), done(stream_count, nullptr);
    auto cleanup = [&] {
        for (auto e : carried)
            if (e) (void)hipEventDestroy(e);
        for (auto e : done)
            if (e) (void)hipEventDestroy(e);
    };
    for (uint32_t i = 0; i < stream_count; i++) {
        if (hipEventCreateWithFlags(&carried[i], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&done[i], hipEventDisableTiming) != hipSuccess) {
            cleanup();
            return fail(CSX_ERR_HIP, "event creation failed");
        }
    }
    hipError_t e = hipSuccess;
    for (uint32_t i = 0; i < stream_count && e == hipSuccess; i++) {
        hipStream_t st = streams[i];
        const uint64_t e0 = p.begin(i, n), e1 = p.end(i, n), ni = e1 - e0;
        const uint64_t c0 = e0 >> log2c, ci = chunk_count(ni, chunk_len);
        auto hdr = reinterpret_cast<csx::WsHeader*>(base + p.scan_off + i * p.scan_stride);
        auto status = reinterpret_cast<uint64_t*>(base + p.scan_off + i * p.scan_stride + sizeof(csx::WsHeader));
        // phases 1-2 of the slice run concurrently with the other slices
        if (ni > 0) {
            e = csx::launch_popc(d_mask + e0 / 8, counts + c0, ni, log2c, bit_order, st, hdr, status, csx::scan_tiles(ci));
            if (e == hipSuccess) e = csx::launch_scan(counts + c0, offsets + c0, ci, totals + i, hdr, status, st, true);
        }
        else {
            e = csx::launch_store_u64(totals + i, 0, st);
        }
        // the carry waits for the previous slice's carry (which follows its scan)
        if (e == hipSuccess && i > 0) e = hipStreamWaitEvent(st, carried[i - 1], 0);
        if (e == hipSuccess) e = csx::launch_carry(bases + i, bases, totals, i, st);
        if (e == hipSuccess) e = hipEventRecord(carried[i], st);
        if (e == hipSuccess && ni > 0) {
            csx::TileArgs a{};
            a.in = in + e0 * elem_size;
            a.mask = d_mask + e0 / 8;
            a.out = out;
            a.n = ni;
            a.offsets = offsets + c0;
            a.log2c = log2c;
            a.out_base = bases + i;
            const bool aligned = (((uintptr_t)a.in | (uintptr_t)a.mask) & 15) == 0;
            e = timed(st, [&] { return dispatch_tile(elem_size, csx::kModeWriteout, true, bit_order, aligned, a, st); });
        }
        if (e == hipSuccess) e = hipEventRecord(done[i], st);
    }
    // join on streams[0], then the total
    for (uint32_t i = 1; i < stream_count && e == hipSuccess; i++) e = hipStreamWaitEvent(streams[0], done[i], 0);
    uint64_t* count_dst = d_count ? d_count : &reinterpret_cast<csx::WsHeader*>(base + p.scan_off)->count;
    if (e == hipSuccess) e = csx::launch_carry(count_dst, bases, totals, stream_count, streams[0]);
    cleanup();
    if (e != hipSuccess) return hip_fail(e, "csx_compact_streams");
    if (h_count) {
        CSX_HIP(hipMemcpyAsync(h_count, count_dst, sizeof(uint64_t), hipMemcpyDeviceToHost, streams[0]), "count copy");
        CSX_HIP(hipStreamSynchronize(streams[0]), "stream synchronize");
    }
    return ok();
}

// ---------------------------------------------------------------- batched columns

static_assert(sizeof(csx_segment) == sizeof(csx::Segment), "segment layouts agree");

struct BatchPlan {
    uint64_t max_tiles, status_words;
    size_t prefix_off, tcounts_off, rec_off, goff_off, status_off, total;
};

// the workspace is sized for the smallest tile (128 threads x 2 vectors), so
// every batched configuration fits in csx_batch_workspace_size's answer
constexpr uint32_t kBatchMinVpt = 2, kBatchMinBlock = 128;

static BatchPlan plan_batch(uint32_t nseg, uint64_t max_total, uint32_t w, uint32_t vpt, uint32_t block)
{
    BatchPlan p;
    p.max_tiles = csx::batch_max_tiles(w, nseg, max_total, vpt, block);
    p.status_words = csx::scan_tiles(p.max_tiles);
    p.prefix_off = sizeof(csx::WsHeader);
    p.tcounts_off = align_up(p.prefix_off + 8ull * (nseg + 1));
    p.rec_off = align_up(p.tcounts_off + 4 * p.max_tiles);  // each tile's record
    p.goff_off = align_up(p.rec_off + sizeof(csx::TileRec) * p.max_tiles);
    p.status_off = align_up(p.goff_off + 8 * p.max_tiles);
    p.total = align_up(p.status_off + 8 * (p.status_words + 1));
    return p;
}

csx_status csx_batch_workspace_size(uint32_t nseg, uint64_t max_total_elements, uint32_t elem_size, size_t* bytes)
{
    csx_status s;
    if ((s = check_width(elem_size)) || (s = check_n(max_total_elements))) return s;
    if (!bytes) return fail(CSX_ERR_INVALID, "null size pointer");
    *bytes = plan_batch(nseg, max_total_elements, elem_size, kBatchMinVpt, kBatchMinBlock).total;
    return ok();
}

static csx_status batch_impl(const csx_segment* d_segments, uint32_t nseg, uint64_t max_total_elements, uint32_t elem_size, int bit_order,
                             const csx_config& cfg, uint64_t* d_counts, void* d_ws, size_t ws_bytes, hipStream_t stream)
{
    csx_status s;
    if ((s = check_width(elem_size)) || (s = check_order(bit_order)) || (s = check_n(max_total_elements))) return s;
    if (cfg.vpt != 2 && cfg.vpt != 4 && cfg.vpt != 8) return fail(CSX_ERR_INVALID, "batched tile of %u vectors per thread (2, 4 or 8)", cfg.vpt);
    if ((s = check_block(cfg.block_size))) return s;
    if (nseg == 0) return ok();
    if (!d_segments || !d_counts) return fail(CSX_ERR_INVALID, "null segment table or counts pointer");
    const uint32_t block = block_of(cfg.block_size);
    const BatchPlan p = plan_batch(nseg, max_total_elements, elem_size, cfg.vpt, block);
    if ((s = check_ws(d_ws, ws_bytes, p.total))) return s;
    if (!launch_fits(p.max_tiles, block)) return fail(CSX_ERR_UNSUPPORTED, "%llu tiles exceed one launch", (unsigned long long)p.max_tiles);
    auto base = static_cast<uint8_t*>(d_ws);
    const uint32_t count_blocks = cfg.count_blocks ? cfg.count_blocks : 4096;
    CSX_HIP(timed_span(stream, [&] {
                return csx::launch_batch(elem_size, bit_order, reinterpret_cast<const csx::Segment*>(d_segments), nseg, p.max_tiles, d_counts,
                                         reinterpret_cast<uint64_t*>(base + p.prefix_off), reinterpret_cast<uint32_t*>(base + p.tcounts_off),
                                         reinterpret_cast<csx::TileRec*>(base + p.rec_off), reinterpret_cast<uint64_t*>(base + p.goff_off),
                                         reinterpret_cast<csx::WsHeader*>(base),
                                         reinterpret_cast<uint64_t*>(base + p.status_off), stream, cfg.vpt, count_blocks, block);
            }),
            "csx_compact_batch");
    return ok();
}

csx_status csx_compact_batch(const csx_segment* d_segments, uint32_t nseg, uint64_t max_total_elements, uint32_t elem_size, int bit_order,
                             uint64_t* d_counts, void* d_ws, size_t ws_bytes, hipStream_t stream)
{
    return batch_impl(d_segments, nseg, max_total_elements, elem_size, bit_order, batch_config(elem_size), d_counts, d_ws, ws_bytes, stream);
}

csx_status csx_compact_batch_config(const csx_segment* d_segments, uint32_t nseg, uint64_t max_total_elements, uint32_t elem_size, int bit_order,
                                    const csx_config* cfg, uint64_t* d_counts, void* d_ws, size_t ws_bytes, hipStream_t stream)
{
    if (!cfg) return fail(CSX_ERR_INVALID, "null configuration");
    return batch_impl(d_segments, nseg, max_total_elements, elem_size, bit_order, *cfg, d_counts, d_ws, ws_bytes, stream);
}

// ---------------------------------------------------------------- generators

static uint64_t bernoulli_threshold(double p) { return (uint64_t)(p * UINT32_MAX); }

csx_status csx_gen_mask_bernoulli(uint8_t* d_mask, uint64_t n, double p, uint64_t seed, int bit_order, hipStream_t stream)
{
    csx_status s;
    if ((s = check_order(bit_order)) || (s = check_n(n))) return s;
    if (!(p >= 0.0 && p <= 1.0)) return fail(CSX_ERR_INVALID, "selectivity must be in [0, 1]");
    if (n == 0) return ok();
    if (!d_mask) return fail(CSX_ERR_INVALID, "null mask pointer");
    CSX_HIP(csx::launch_gen_bernoulli(d_mask, n, bernoulli_threshold(p), seed, bit_order, stream), "csx_gen_mask_bernoulli");
    return ok();
}

csx_status csx_gen_mask_pattern(uint8_t* d_mask, uint64_t n, uint32_t pattern, uint32_t pattern_length, int bit_order, hipStream_t stream)
{
    csx_status s;
    if ((s = check_order(bit_order)) || (s = check_n(n))) return s;
    if (pattern_length < 1 || pattern_length > 32) return fail(CSX_ERR_INVALID, "pattern length %u must be in [1, 32]", pattern_length);
    if (n == 0) return ok();
    if (!d_mask) return fail(CSX_ERR_INVALID, "null mask pointer");
    CSX_HIP(csx::launch_gen_pattern(d_mask, n, pattern, pattern_length, bit_order, stream), "csx_gen_mask_pattern");
    return ok();
}

csx_status csx_zipf_table(double s_exp, uint32_t levels, uint32_t* cdf, uint64_t* thr)
{
    if (levels < 2 || levels > 128) return fail(CSX_ERR_INVALID, "zipf levels %u must be in [2, 128]", levels);
    if (!cdf || !thr) return fail(CSX_ERR_INVALID, "null table pointer");
    double norm = 0.0;
    for (uint32_t r = 1; r <= levels; r++) norm += std::pow((double)r, -s_exp);
    double acc = 0.0;
    for (uint32_t r = 1; r <= levels; r++) {
        acc += std::pow((double)r, -s_exp) / norm;
        const double t = acc * 4294967296.0;
        cdf[r - 1] = (r == levels || t >= 4294967295.0) ? UINT32_MAX : (uint32_t)t;
        thr[r - 1] = bernoulli_threshold((double)(r - 1) / (double)(levels - 1));
    }
    return ok();
}

csx_status csx_gen_mask_zipf_blocks(uint8_t* d_mask, uint64_t n, uint32_t block_len, double s_exp, uint32_t levels, uint64_t seed, int bit_order,
                                    hipStream_t stream)
{
    csx_status s;
    if ((s = check_order(bit_order)) || (s = check_n(n))) return s;
    if (block_len == 0 || block_len % 32) return fail(CSX_ERR_INVALID, "block length %u must be a positive multiple of 32", block_len);
    uint32_t cdf[128];
    uint64_t thr[128];
    if ((s = csx_zipf_table(s_exp, levels, cdf, thr))) return s;
    if (n == 0) return ok();
    if (!d_mask) return fail(CSX_ERR_INVALID, "null mask pointer");
    CSX_HIP(csx::launch_gen_zipf(d_mask, n, block_len, seed, cdf, thr, levels, bit_order, stream), "csx_gen_mask_zipf_blocks");
    return ok();
}

struct RunsPlan {
    uint64_t words;
    size_t counts_off, before_off, status_off, total;
};

static RunsPlan plan_runs(uint64_t n)
{
    RunsPlan p;
    p.words = (n + 31) / 32;
    p.counts_off = sizeof(csx::WsHeader);
    p.before_off = align_up(p.counts_off + p.words * 4);
    p.status_off = align_up(p.before_off + p.words * 8);
    p.total = align_up(p.status_off + csx::scan_tiles(p.words) * 8);
    return p;
}

csx_status csx_gen_runs_workspace_size(uint64_t n, size_t* bytes)
{
    if (!bytes) return fail(CSX_ERR_INVALID, "null size pointer");
    *bytes = plan_runs(n).total;
    return ok();
}

csx_status csx_gen_mask_runs(uint8_t* d_mask, uint64_t n, uint32_t mean_run, uint64_t seed, int bit_order, void* d_ws, size_t ws_bytes,
                             hipStream_t stream)
{
    csx_status s;
    if ((s = check_order(bit_order)) || (s = check_n(n))) return s;
    if (mean_run < 2) return fail(CSX_ERR_INVALID, "mean run length must be >= 2");
    if (n == 0) return ok();
    if (!d_mask) return fail(CSX_ERR_INVALID, "null mask pointer");
    const RunsPlan p = plan_runs(n);
    if ((s = check_ws(d_ws, ws_bytes, p.total))) return s;
    auto base = static_cast<uint8_t*>(d_ws);
    const uint32_t thr = (uint32_t)((1ull << 32) / mean_run);
    CSX_HIP(csx::launch_gen_runs(d_mask, n, thr, seed, bit_order, reinterpret_cast<uint32_t*>(base + p.counts_off),
                                 reinterpret_cast<uint64_t*>(base + p.before_off), reinterpret_cast<csx::WsHeader*>(base),
                                 reinterpret_cast<uint64_t*>(base + p.status_off), stream),
            "csx_gen_mask_runs");
    return ok();
}

csx_status csx_gen_values(void* d_out, uint64_t n, uint32_t elem_size, int kind, uint64_t seed, hipStream_t stream)
{
    csx_status s;
    if ((s = check_width(elem_size)) || (s = check_n(n))) return s;
    if (kind != CSX_VALUES_INTEGER && kind != CSX_VALUES_FLOAT) return fail(CSX_ERR_INVALID, "unknown value kind %d", kind);
    if (kind == CSX_VALUES_FLOAT && elem_size < 4) return fail(CSX_ERR_INVALID, "float values need 4- or 8-byte elements");
    if (n == 0) return ok();
    if (!d_out || ((uintptr_t)d_out % elem_size)) return fail(CSX_ERR_INVALID, "null or misaligned output pointer");
    CSX_HIP(csx::launch_gen_values(d_out, n, elem_size, kind, seed, stream), "csx_gen_values");
    return ok();
}

csx_status csx_copy(void* d_dst, const void* d_src, uint64_t bytes, hipStream_t stream)
{
    if (bytes == 0) return ok();
    if (!d_dst || !d_src) return fail(CSX_ERR_INVALID, "null pointer");
    if ((((uintptr_t)d_dst) | ((uintptr_t)d_src)) & 15) return fail(CSX_ERR_INVALID, "csx_copy needs 16-byte aligned pointers");
    CSX_HIP(timed(stream, [&] { return csx::launch_copy(d_dst, d_src, bytes, stream); }), "csx_copy launch");
    return ok();
}

csx_status csx_stream_probe(int kind, void* d_dst, const void* d_src, uint64_t bytes, uint32_t vectors_per_thread, uint32_t xcd_run,
                            hipStream_t stream)
{
    if (kind < CSX_PROBE_READ || kind > CSX_PROBE_MIX) return fail(CSX_ERR_INVALID, "unknown probe kind %d", kind);
    if (vectors_per_thread != 1 && vectors_per_thread != 4) return fail(CSX_ERR_INVALID, "vectors per thread must be 1 or 4");
    if (xcd_run > 64) return fail(CSX_ERR_INVALID, "XCD run of %u blocks (at most 64)", xcd_run);
    if (bytes % 32) return fail(CSX_ERR_INVALID, "probe size must be a multiple of 32 bytes");
    if (bytes / 16 / (256ull * vectors_per_thread) >= (1ull << 31)) return fail(CSX_ERR_UNSUPPORTED, "probe too large for one launch");
    if (bytes == 0) return ok();
    if (!d_dst || (kind != CSX_PROBE_WRITE && !d_src)) return fail(CSX_ERR_INVALID, "null pointer");
    if (((uintptr_t)d_dst | (uintptr_t)d_src) & 15) return fail(CSX_ERR_INVALID, "probe pointers must be 16-byte aligned");
    CSX_HIP(timed(stream, [&] { return csx::launch_stream_probe(kind, d_src, d_dst, bytes, vectors_per_thread, xcd_run, stream); }),
            "csx_stream_probe launch");
    return ok();
}

csx_status csx_profile_begin(void)
{
    for (auto& p : g_prof.events) {
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    g_prof.events.clear();
    g_prof.on = true;
    return ok();
}

csx_status csx_profile_end(double* total_ms, uint64_t* launches)
{
    g_prof.on = false;
    double sum = 0.0;
    hipError_t err = hipSuccess;
    for (auto& p : g_prof.events) {
        float ms = 0.f;
        if (err == hipSuccess) err = hipEventSynchronize(p.second);
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, p.first, p.second);
        sum += ms;
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    const uint64_t count = g_prof.events.size();
    g_prof.events.clear();
    if (total_ms) *total_ms = sum;
    if (launches) *launches = count;
    if (err != hipSuccess) return hip_fail(err, "csx_profile_end");
    return ok();
}

csx_status csx_last_launch(uint32_t* block_size, uint64_t* grid)
{
    if (!block_size || !grid) return fail(CSX_ERR_INVALID, "null argument");
    *block_size = csx::g_launch_shape.block;
    *grid = csx::g_launch_shape.grid;
    return ok();
}

csx_status csx_predicate_mask(const void* d_in, uint64_t n, const csx_predicate* pred, uint8_t* d_mask, int bit_order, uint64_t* d_count,
                              hipStream_t stream)
{
    csx_status s;
    csx::PredSpec p;
    uint32_t w = 0;
    if ((s = check_pred(pred, &p, &w)) || (s = check_order(bit_order)) || (s = check_n(n))) return s;
    if (n > 0 && (!d_in || !d_mask)) return fail(CSX_ERR_INVALID, "null input or mask pointer");
    if (n > 0 && overlaps(d_mask, (n + 7) / 8, d_in, n * w)) return fail(CSX_ERR_INVALID, "the mask overlaps the column");
    if ((uintptr_t)d_in % w) return fail(CSX_ERR_INVALID, "input must be aligned to the element size");
    if (d_count) CSX_HIP(csx::launch_store_u64(d_count, 0, stream), "csx_predicate_mask count");
    if (n == 0) return ok();
    CSX_HIP(csx::launch_predicate(d_in, n, p, d_mask, bit_order, nullptr, d_count, nullptr, nullptr, 0, stream), "csx_predicate_mask");
    return ok();
}

csx_status csx_predicate_workspace_size(uint64_t n, int type, size_t* bytes)
{
    csx_status s;
    if ((s = check_n(n))) return s;
    if (!bytes) return fail(CSX_ERR_INVALID, "null result pointer");
    csx_predicate q{};
    q.type = type;
    csx::PredSpec p;
    uint32_t w = 0;
    if ((s = check_pred(&q, &p, &w))) return s;
    *bytes = pred_ws_total(n, w);
    return ok();
}

csx_status csx_compact_predicate(const void* d_in, uint64_t n, const csx_predicate* pred, void* d_out, uint8_t* d_mask, int bit_order,
                                 uint64_t* d_count, uint64_t* h_count, void* d_ws, size_t ws_bytes, hipStream_t stream)
{
    csx_status s;
    csx::PredSpec p;
    uint32_t w = 0;
    if ((s = check_pred(pred, &p, &w)) || (s = check_order(bit_order)) || (s = check_n(n))) return s;
    if ((s = check_ws(d_ws, ws_bytes, pred_ws_total(n, w)))) return s;
    if (n > 0 && (!d_in || !d_out)) return fail(CSX_ERR_INVALID, "null input or output pointer");
    if (n > 0 && (overlaps(d_out, n * w, d_in, n * w) || overlaps(d_mask, (n + 7) / 8, d_in, n * w) || overlaps(d_mask, (n + 7) / 8, d_out, n * w)))
        return fail(CSX_ERR_INVALID, "the column, the output and the mask must not overlap (compaction is not in place)");
    if (((uintptr_t)d_in % w) || ((uintptr_t)d_out % w)) return fail(CSX_ERR_INVALID, "input and output must be aligned to the element size");
    const CompactPlan cp = plan_compact(n, w, kPredChunk);
    if (!launch_fits(cp.fused_tiles)) return fail(CSX_ERR_UNSUPPORTED, "%llu tiles exceed one launch", (unsigned long long)cp.fused_tiles);

    auto base = static_cast<uint8_t*>(d_ws);
    auto hdr = reinterpret_cast<csx::WsHeader*>(base);
    auto status = reinterpret_cast<uint64_t*>(base + cp.status_off);
    auto counts = reinterpret_cast<uint32_t*>(base + cp.counts_off);
    auto offsets = reinterpret_cast<uint64_t*>(base + cp.offsets_off);
    uint8_t* mask = d_mask ? d_mask : base + pred_mask_off(n, w);
    uint64_t* count_dst = d_count ? d_count : &hdr->count;
    if (n == 0) {
        CSX_HIP(csx::launch_lookback_init(hdr, status, 0, stream), "csx_compact_predicate init");
        CSX_HIP(csx::launch_store_u64(count_dst, 0, stream), "csx_compact_predicate count");
    }
    else {
        CSX_HIP(csx::launch_predicate(d_in, n, p, mask, bit_order, counts, nullptr, hdr, status, cp.scan_tiles, stream), "csx_compact_predicate mask");
        CSX_HIP(csx::launch_scan(counts, offsets, cp.chunks, count_dst, hdr, status, stream, true), "csx_compact_predicate scan");
        const bool aligned = (((uintptr_t)d_in | (uintptr_t)mask) & 15) == 0;
        // the writeout's tile size: the tuned three-phase choice for this width
        // and size at unknown selectivity (the chunk offsets serve any tile size)
        uint32_t vpt = 8, block = 256;
        const csx_config tc = lookup_config(w, n, -1.0);
        const uint64_t tte = (uint64_t)block_of(tc.block_size) * tc.vpt * (16 / w);
        if (tc.pipeline == CSX_PIPELINE_THREE_PHASE && launch_fits((n + tte - 1) / tte, block_of(tc.block_size))) {
            vpt = tc.vpt;
            block = block_of(tc.block_size);
        }
        csx::TileArgs a{};
        a.in = static_cast<const uint8_t*>(d_in);
        a.mask = mask;
        a.out = static_cast<uint8_t*>(d_out);
        a.n = n;
        if ((uint64_t)block * vpt * (16 / w) % kPredChunk == 0) {
            // every tile starts a chunk, so the chunk's exclusive offset is the
            // tile's: the writeout takes it as its block prefix (a scalar load
            // ahead of the mask, no barrier), with no unit offset
            a.bprefix = offsets;
            a.log2b = log2u(kPredChunk);
        }
        else {
            a.offsets = offsets;
            a.counts = counts;
            a.log2c = log2u(kPredChunk);
        }
        CSX_HIP(timed(stream, [&] { return dispatch_tile(w, csx::kModeWriteout, true, bit_order, aligned, a, stream, (int)vpt, (int)block); }),
                "csx_compact_predicate writeout");
    }
    if (h_count) return finish_host_count(h_count, count_dst, hdr, stream);
    return ok();
}

// ---------------------------------------------------------------- tuning

csx_status csx_auto_workspace_size(uint64_t n, uint32_t elem_size, size_t* bytes)
{
    return csx_workspace_size(n, elem_size, 256, bytes);  // the smallest chunk the tuner tries
}

csx_status csx_tune_lookup(uint32_t elem_size, uint64_t n, double selectivity, csx_config* cfg)
{
    csx_status s;
    if ((s = check_width(elem_size))) return s;
    if (!cfg) return fail(CSX_ERR_INVALID, "null configuration pointer");
    *cfg = lookup_config(elem_size, n, selectivity);
    return ok();
}

csx_status csx_compact_auto(const void* d_in, const uint8_t* d_mask, void* d_out, uint64_t n, uint32_t elem_size, int bit_order, double selectivity,
                            uint64_t* d_count, uint64_t* h_count, void* d_ws, size_t ws_bytes, hipStream_t stream)
{
    csx_status s;
    if ((s = check_width(elem_size))) return s;
    const csx_config c = lookup_config(elem_size, n, selectivity);
    return compact_impl(d_in, d_mask, d_out, n, elem_size, c.chunk_len, bit_order, c.pipeline, c.vpt, c.block_size, d_count, h_count, d_ws, ws_bytes,
                        stream);
}

extern "C++" {
namespace {

// Device buffers owned by one tuning call.
struct TuneBuffers {
    std::vector<void*> ptrs;
    ~TuneBuffers()
    {
        for (void* p : ptrs) (void)hipFree(p);
    }
    hipError_t alloc(void** p, size_t bytes)
    {
        hipError_t e = hipMalloc(p, bytes ? bytes : 16);
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// median over reps of the time of fn() on stream s, after one warm-up run
// Per-call time of `fn` as the device sees it in a stream of calls: `reps`
// calls enqueued back to back between two events, in three rounds after a
// warm-up; the median round's time divided by reps.  (One call bracketed by
// events alone would also time the host's launch path, which at mid sizes
// outweighs the kernels and would hide their differences.)
template <typename F> csx_status time_median(F&& fn, uint32_t reps, hipStream_t s, double* ms)
{
    EventPair ev;
    CSX_HIP(hipEventCreate(&ev.a), "tune event");
    CSX_HIP(hipEventCreate(&ev.b), "tune event");
    csx_status st;
    if ((st = fn())) return st;
    if (reps == 0) reps = 1;
    std::vector<double> t;
    for (int round = 0; round < 3; round++) {
        CSX_HIP(hipEventRecord(ev.a, s), "tune record");
        for (uint32_t r = 0; r < reps; r++)
            if ((st = fn())) return st;
        CSX_HIP(hipEventRecord(ev.b, s), "tune record");
        CSX_HIP(hipEventSynchronize(ev.b), "tune sync");
        float f = 0.f;
        CSX_HIP(hipEventElapsedTime(&f, ev.a, ev.b), "tune elapsed");
        t.push_back(f / reps);
    }
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    return CSX_OK;
}

// Every candidate of the autotuner timed on one column (median of reps after
// a warm-up, the whole call); all must give the same count, returned in *k.
// d_count null: the count goes to the workspace header.
csx_status tune_candidates(const void* in, const uint8_t* mask, void* out, uint64_t n, uint32_t elem_size, int bit_order, uint64_t* d_count, void* ws,
                           size_t ws_bytes, uint32_t reps, hipStream_t stream, csx_config* best, double* best_t, uint64_t* k)
{
    std::vector<csx_config> cand;
    for (uint32_t chunk : {256u, 1024u, 4096u})
        for (uint32_t vpt : {2u, 4u, 8u, 16u})
            if (vpt != 16 || elem_size >= 2) cand.push_back(csx_config{CSX_PIPELINE_THREE_PHASE, chunk, vpt, 0, 256});
    // the other workgroup sizes at the usual chunk length
    for (uint32_t block : {128u, 512u})
        for (uint32_t vpt : {2u, 4u, 8u, 16u})
            if (vpt != 16 || elem_size >= 2) cand.push_back(csx_config{CSX_PIPELINE_THREE_PHASE, 1024, vpt, 0, block});
    cand.push_back(csx_config{CSX_PIPELINE_TWO_PASS, 1024, 8, 0});
    cand.push_back(csx_config{CSX_PIPELINE_SPARSE, 1024, 8, 0});
    cand.push_back(csx_config{CSX_PIPELINE_SPARSE, 4096, 8, 0});
    cand.push_back(csx_config{CSX_PIPELINE_FUSED, 1024, 8, 0});
    if (n <= CSX_SINGLE_MAX_ELEMENTS)
        for (uint32_t vpt : {2u, 4u, 8u}) cand.push_back(csx_config{CSX_PIPELINE_SINGLE, 1024, vpt, 0});

    auto hdr = static_cast<csx::WsHeader*>(ws);
    const uint64_t* count_src = d_count ? d_count : &hdr->count;
    uint64_t first_count = 0;
    for (size_t i = 0; i < cand.size(); i++) {
        const csx_config c = cand[i];
        double t = 0.0;
        csx_status s = time_median(
            [&] {
                return compact_impl(in, mask, out, n, elem_size, c.chunk_len, bit_order, c.pipeline, c.vpt, c.block_size, d_count, nullptr, ws, ws_bytes,
                                    stream);
            },
            reps, stream, &t);
        if (s) return s;
        uint64_t kc = 0;
        if ((s = finish_host_count(&kc, count_src, hdr, stream))) return s;
        if (i == 0) first_count = kc;
        if (kc != first_count)
            return fail(CSX_ERR_DEVICE, "configurations disagree on the selected count (%llu, %llu)", (unsigned long long)first_count,
                        (unsigned long long)kc);
        if (i == 0 || t < *best_t) {
            *best_t = t;
            *best = c;
        }
    }
    *k = first_count;
    return CSX_OK;
}

} // namespace
} // extern "C++"

csx_status csx_autotune(uint32_t elem_size, uint64_t n, double selectivity, uint32_t reps, csx_config* best, double* best_ms, hipStream_t stream)
{
    csx_status s;
    if ((s = check_width(elem_size)) || (s = check_n(n))) return s;
    if (!(selectivity >= 0.0 && selectivity <= 1.0)) return fail(CSX_ERR_INVALID, "selectivity must be in [0, 1]");
    if (n == 0 || reps == 0) return fail(CSX_ERR_INVALID, "nothing to time (n = 0 or reps = 0)");
    if (!best) return fail(CSX_ERR_INVALID, "null result pointer");
    size_t ws_bytes = 0;
    if ((s = csx_auto_workspace_size(n, elem_size, &ws_bytes))) return s;
    TuneBuffers buf;
    void *in, *mask, *out, *ws, *cnt;
    CSX_HIP(buf.alloc(&in, n * elem_size), "tune alloc");
    CSX_HIP(buf.alloc(&mask, (n + 7) / 8), "tune alloc");
    CSX_HIP(buf.alloc(&out, n * elem_size), "tune alloc");
    CSX_HIP(buf.alloc(&ws, ws_bytes), "tune alloc");
    CSX_HIP(buf.alloc(&cnt, sizeof(uint64_t)), "tune alloc");
    if ((s = csx_gen_values(in, n, elem_size, CSX_VALUES_INTEGER, 0x74756e65ull, stream))) return s;
    if ((s = csx_gen_mask_bernoulli(static_cast<uint8_t*>(mask), n, selectivity, 0x6d61736bull, CSX_MSB_FIRST, stream))) return s;
    double best_t = 0.0;
    uint64_t k = 0;
    if ((s = tune_candidates(in, static_cast<const uint8_t*>(mask), out, n, elem_size, CSX_MSB_FIRST, static_cast<uint64_t*>(cnt), ws, ws_bytes, reps,
                             stream, best, &best_t, &k)))
        return s;
    record_cell(elem_size, n, selectivity, *best, best_t);
    if (best_ms) *best_ms = best_t;
    return ok();
}

csx_status csx_autotune_inputs(const void* d_in, const uint8_t* d_mask, void* d_out, uint64_t n, uint32_t elem_size, int bit_order, uint32_t reps,
                               void* d_ws, size_t ws_bytes, csx_config* best, double* best_ms, uint64_t* h_count, hipStream_t stream)
{
    csx_status s;
    if ((s = check_width(elem_size)) || (s = check_order(bit_order)) || (s = check_n(n))) return s;
    if (n == 0 || reps == 0) return fail(CSX_ERR_INVALID, "nothing to time (n = 0 or reps = 0)");
    if (!best) return fail(CSX_ERR_INVALID, "null result pointer");
    if (!d_in || !d_mask || !d_out) return fail(CSX_ERR_INVALID, "null input, mask or output pointer");
    size_t need = 0;
    if ((s = csx_auto_workspace_size(n, elem_size, &need)) || (s = check_ws(d_ws, ws_bytes, need))) return s;
    // the count goes to the workspace header (no d_count)
    double best_t = 0.0;
    uint64_t k = 0;
    if ((s = tune_candidates(d_in, d_mask, d_out, n, elem_size, bit_order, nullptr, d_ws, ws_bytes, reps, stream, best, &best_t, &k))) return s;
    record_cell(elem_size, n, (double)k / (double)n, *best, best_t);
    if (best_ms) *best_ms = best_t;
    if (h_count) *h_count = k;
    return ok();
}

csx_status csx_autotune_batch(const csx_segment* d_segments, uint32_t nseg, uint64_t max_total_elements, uint32_t elem_size, int bit_order,
                              uint64_t* d_counts, void* d_ws, size_t ws_bytes, uint32_t reps, csx_config* best, double* best_ms, hipStream_t stream)
{
    csx_status s;
    if ((s = check_width(elem_size))) return s;
    if (!best || reps == 0) return fail(CSX_ERR_INVALID, "null result pointer or reps = 0");
    double best_t = 0.0;
    bool have = false;
    std::vector<csx_config> cand;
    for (uint32_t vpt : {2u, 4u, 8u}) {
        for (uint32_t cb : {1024u, 2048u, 4096u, 8192u}) cand.push_back(csx_config{CSX_PIPELINE_THREE_PHASE, 0, vpt, cb, 256});
        for (uint32_t block : {128u, 512u})
            for (uint32_t cb : {4096u, 8192u}) cand.push_back(csx_config{CSX_PIPELINE_THREE_PHASE, 0, vpt, cb, block});
    }
    for (const csx_config& c : cand) {
        double t = 0.0;
        s = time_median([&] { return batch_impl(d_segments, nseg, max_total_elements, elem_size, bit_order, c, d_counts, d_ws, ws_bytes, stream); },
                        reps, stream, &t);
        if (s) return s;
        if (!have || t < best_t) {
            best_t = t;
            *best = c;
            have = true;
        }
    }
    {
        TuneTable& t = tune_table();
        std::lock_guard<std::mutex> g(t.mu);
        t.batch[elem_size] = *best;
    }
    if (best_ms) *best_ms = best_t;
    return ok();
}

csx_status csx_tune_save(const char* path)
{
    if (!path) return fail(CSX_ERR_INVALID, "null path");
    FILE* f = fopen(path, "w");
    if (!f) return fail(CSX_ERR_INVALID, "cannot open %s for writing", path);
    TuneTable& t = tune_table();
    std::lock_guard<std::mutex> g(t.mu);
    fprintf(f, "# csx tuning table 2\n# compact <elem_size> <log2 n> <selectivity> <pipeline> <chunk_len> <vpt> <block_size> <ms>\n"
               "# batch <elem_size> <vpt> <count_blocks> <block_size>\n");
    for (const TuneCell& c : t.cells)
        fprintf(f, "compact %u %u %.6f %d %u %u %u %.6f\n", c.w, c.log2n, c.sel, c.cfg.pipeline, c.cfg.chunk_len, c.cfg.vpt, block_of(c.cfg.block_size),
                c.ms);
    for (uint32_t w : {1u, 2u, 4u, 8u})
        if (t.batch[w].vpt) fprintf(f, "batch %u %u %u %u\n", w, t.batch[w].vpt, t.batch[w].count_blocks, block_of(t.batch[w].block_size));
    const bool bad = ferror(f) != 0;
    if (fclose(f) != 0 || bad) return fail(CSX_ERR_INVALID, "writing %s failed", path);
    return ok();
}

csx_status csx_tune_load(const char* path)
{
    if (!path) return fail(CSX_ERR_INVALID, "null path");
    FILE* f = fopen(path, "r");
    if (!f) return fail(CSX_ERR_INVALID, "cannot open %s", path);
    char line[512];
    std::vector<TuneCell> cells;
    std::vector<std::pair<uint32_t, csx_config>> batch;
    uint32_t lineno = 0;
    csx_status st = CSX_OK;
    while (fgets(line, sizeof line, f)) {
        lineno++;
        const size_t len = strlen(line);
        if (len == sizeof line - 1 && line[len - 1] != '\n' && !feof(f)) {
            // a row never comes near the buffer; a split line would parse as two
            st = fail(CSX_ERR_INVALID, "%s:%u: line longer than %zu bytes", path, lineno, sizeof line - 2);
            break;
        }
        const char* row = line + strspn(line, " \t\r\n");  // blank lines and comments may be indented
        if (*row == '#' || *row == '\0') continue;
        TuneCell c{};
        uint32_t w, vpt, cb, block = 256;
        int end = -1, bend = -1;  // offsets past the row's fields and trailing blanks: the row must end there
        // rows without a workgroup size (version 1 of the format) mean 256; the
        // shorter form is tried first, and it must end the row, so that a
        // version-2 row's integer block size is never read as the time
        const bool compact7 = sscanf(row, "compact %u %u %lf %d %u %u %lf %n", &c.w, &c.log2n, &c.sel, &c.cfg.pipeline, &c.cfg.chunk_len, &c.cfg.vpt,
                                     &c.ms, &end) == 7 &&
                              end >= 0 && row[end] == '\0';
        const bool compact8 = !compact7 && (end = -1, sscanf(row, "compact %u %u %lf %d %u %u %u %lf %n", &c.w, &c.log2n, &c.sel, &c.cfg.pipeline,
                                                             &c.cfg.chunk_len, &c.cfg.vpt, &block, &c.ms, &end) == 8) &&
                              end >= 0 && row[end] == '\0';
        if (compact7 || compact8) {
            c.cfg.block_size = block;
            if (check_width(c.w) || check_chunk(c.cfg.chunk_len) || check_vpt(c.cfg.vpt, c.w) || c.log2n > 40 || !(c.sel >= 0 && c.sel <= 1) ||
                !(c.ms >= 0 && std::isfinite(c.ms)) || (block != 128 && block != 256 && block != 512) ||
                !known_pipeline(c.cfg.pipeline) || (c.cfg.pipeline == CSX_PIPELINE_SINGLE && c.log2n > 24) ||
                pipeline_shape_error(c.cfg.pipeline, c.cfg.vpt, block)) {
                st = fail(CSX_ERR_INVALID, "%s:%u: configuration out of range", path, lineno);
                break;
            }
            cells.push_back(c);
        }
        else if ((sscanf(row, "batch %u %u %u %n", &w, &vpt, &cb, &bend) == 3 && bend >= 0 && row[bend] == '\0') ||
                 (bend = -1, sscanf(row, "batch %u %u %u %u %n", &w, &vpt, &cb, &block, &bend) == 4 && bend >= 0 && row[bend] == '\0')) {
            if (check_width(w) || (vpt != 2 && vpt != 4 && vpt != 8) || cb == 0 || (block != 128 && block != 256 && block != 512)) {
                st = fail(CSX_ERR_INVALID, "%s:%u: batch configuration out of range", path, lineno);
                break;
            }
            batch.push_back({w, csx_config{CSX_PIPELINE_THREE_PHASE, 0, vpt, cb, block}});
        }
        else {
            st = fail(CSX_ERR_INVALID, "%s:%u: unreadable line", path, lineno);
            break;
        }
    }
    fclose(f);
    if (st) return st;
    for (const TuneCell& c : cells) record_cell(c.w, 1ull << c.log2n, c.sel, c.cfg, c.ms);
    TuneTable& t = tune_table();
    std::lock_guard<std::mutex> g(t.mu);
    for (auto& b : batch) t.batch[b.first] = b.second;
    return ok();
}

csx_status csx_tune_clear(void)
{
    TuneTable& t = tune_table();
    std::lock_guard<std::mutex> g(t.mu);
    t.cells.clear();
    for (csx_config& c : t.batch) c = csx_config{0, 0, 0, 0, 0};
    return ok();
}

} // extern "C"
