// csx_server.cpp -- continuous-batching compaction server (host side, C++).
//
// Requests (device buffers in, mask, out of n elements) are queued on the
// host; every csx_server_step() groups the queued requests by element width
// and launches one batched compaction per group (csx_compact_batch) on the
// next stream of a ring, so a batch formed from new arrivals overlaps with
// batches still running.  csx_server_poll() retires finished batches (event
// query, no blocking) and reports each request's selected count and its
// latency from submission to the poll that saw it finish.
//
// The arena (csx_arena_*) is the HBM budget: request buffers are carved from
// one device allocation with a first-fit free list, so admission is simply
// "does the request fit in what is free".
#include "csx.h"

#include <hip/hip_runtime.h>

#include "csx_internal.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <vector>

namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

} // namespace

// ------------------------------------------------------------------- arena

struct csx_arena {
    uint8_t* base = nullptr;
    uint64_t bytes = 0;
    uint64_t used = 0;
    std::map<uint64_t, uint64_t> free_list;  // offset -> size, coalesced
    std::map<uint64_t, uint64_t> live;       // offset -> size
    std::mutex mu;
};

static constexpr uint64_t kArenaAlign = 256;

extern "C" csx_status csx_arena_create(uint64_t bytes, csx_arena** out)
{
    if (!out || bytes == 0) return csx::set_error(CSX_ERR_INVALID, "arena: null result pointer or zero bytes");
    auto* a = new (std::nothrow) csx_arena();
    if (!a) return csx::set_error(CSX_ERR_INVALID, "arena: out of host memory");
    a->bytes = (bytes + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
    const hipError_t e = hipMalloc(&a->base, a->bytes);
    if (e != hipSuccess) {
        delete a;
        return csx::set_error(CSX_ERR_HIP, "arena: hipMalloc of %llu bytes: %s", (unsigned long long)bytes, hipGetErrorString(e));
    }
    a->free_list[0] = a->bytes;
    *out = a;
    return csx::clear_error();
}

extern "C" csx_status csx_arena_destroy(csx_arena* a)
{
    if (!a) return csx::set_error(CSX_ERR_INVALID, "arena: null handle");
    hipError_t e = hipFree(a->base);
    delete a;
    if (e != hipSuccess) return csx::set_error(CSX_ERR_HIP, "arena: hipFree: %s", hipGetErrorString(e));
    return csx::clear_error();
}

// first fit; CSX_ERR_WORKSPACE when no free range is large enough
extern "C" csx_status csx_arena_alloc(csx_arena* a, uint64_t bytes, void** ptr)
{
    if (!a || !ptr) return csx::set_error(CSX_ERR_INVALID, "arena: null handle or result pointer");
    const uint64_t need = std::max<uint64_t>(kArenaAlign, (bytes + kArenaAlign - 1) / kArenaAlign * kArenaAlign);
    std::lock_guard<std::mutex> lk(a->mu);
    for (auto it = a->free_list.begin(); it != a->free_list.end(); ++it) {
        if (it->second >= need) {
            const uint64_t off = it->first, rest = it->second - need;
            a->free_list.erase(it);
            if (rest) a->free_list[off + need] = rest;
            a->live[off] = need;
            a->used += need;
            *ptr = a->base + off;
            return CSX_OK;
        }
    }
    return csx::set_error(CSX_ERR_WORKSPACE, "arena: no free range of %llu bytes (%llu of %llu in use)", (unsigned long long)need,
                          (unsigned long long)a->used, (unsigned long long)a->bytes);
}

extern "C" csx_status csx_arena_free(csx_arena* a, void* ptr)
{
    if (!a || !ptr) return csx::set_error(CSX_ERR_INVALID, "arena: null handle or pointer");
    std::lock_guard<std::mutex> lk(a->mu);
    const uint8_t* p = static_cast<uint8_t*>(ptr);
    if (p < a->base || p >= a->base + a->bytes) return csx::set_error(CSX_ERR_INVALID, "arena: pointer outside the arena");
    const uint64_t off = (uint64_t)(p - a->base);
    auto it = a->live.find(off);
    if (it == a->live.end()) return csx::set_error(CSX_ERR_INVALID, "arena: pointer is not a live allocation");
    uint64_t start = off, size = it->second;
    a->live.erase(it);
    a->used -= size;
    auto next = a->free_list.lower_bound(start);
    if (next != a->free_list.end() && next->first == start + size) {
        size += next->second;
        next = a->free_list.erase(next);
    }
    if (next != a->free_list.begin()) {
        auto prev = std::prev(next);
        if (prev->first + prev->second == start) {
            start = prev->first;
            size += prev->second;
            a->free_list.erase(prev);
        }
    }
    a->free_list[start] = size;
    return CSX_OK;
}

extern "C" uint64_t csx_arena_used(const csx_arena* a) { return a ? a->used : 0; }

// ------------------------------------------------------------------- server

namespace {

struct Pending {
    csx_request req;
    double submit_ms;
    uint64_t seq;  // submission order
};

// one in-flight batch: its stream slot, requests and count read-back
struct Batch {
    int slot = -1;
    uint32_t width = 0;
    std::vector<Pending> reqs;
};

// A stream with its own table, counts and workspace; at most one batch in
// flight, so one completion event per slot.
struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    csx_segment* h_table = nullptr;  // pinned
    csx_segment* d_table = nullptr;
    uint64_t* d_counts = nullptr;
    uint64_t* h_counts = nullptr;    // pinned
    void* d_ws = nullptr;
    size_t ws_bytes = 0;
    bool busy = false;
};

} // namespace

struct csx_server {
    csx_server_config cfg{};
    std::vector<Slot> slots;
    std::deque<Pending> queue;
    std::deque<Batch> inflight;
    uint32_t next_slot = 0;
    uint64_t launched_batches = 0;
    uint64_t next_seq = 0;  // submission order, kept when a failed batch is re-queued
};

extern "C" csx_status csx_server_create(const csx_server_config* cfg, csx_server** out)
{
    if (!cfg || !out) return csx::set_error(CSX_ERR_INVALID, "server: null configuration or result pointer");
    if (cfg->streams < 1 || cfg->streams > 16 || cfg->max_batch < 1 || cfg->max_batch_elements == 0)
        return csx::set_error(CSX_ERR_INVALID, "server: streams must be 1..16, max_batch and max_batch_elements positive");
    auto* s = new (std::nothrow) csx_server();
    if (!s) return csx::set_error(CSX_ERR_INVALID, "server: out of host memory");
    s->cfg = *cfg;
    s->slots.resize(cfg->streams);
    size_t ws = 0;
    for (uint32_t w : {1u, 2u, 4u, 8u}) {
        size_t b = 0;
        const csx_status st = csx_batch_workspace_size(cfg->max_batch, cfg->max_batch_elements, w, &b);
        if (st != CSX_OK) {
            delete s;
            return st;
        }
        ws = std::max(ws, b);
    }
    for (auto& sl : s->slots) {
        bool ok = hipStreamCreate(&sl.stream) == hipSuccess && hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) == hipSuccess &&
                  hipHostMalloc(&sl.h_table, sizeof(csx_segment) * cfg->max_batch) == hipSuccess &&
                  hipMalloc(&sl.d_table, sizeof(csx_segment) * cfg->max_batch) == hipSuccess &&
                  hipMalloc(&sl.d_counts, sizeof(uint64_t) * cfg->max_batch) == hipSuccess &&
                  hipHostMalloc(&sl.h_counts, sizeof(uint64_t) * cfg->max_batch) == hipSuccess && hipMalloc(&sl.d_ws, ws) == hipSuccess;
        sl.ws_bytes = ws;
        if (!ok) {
            csx_server_destroy(s);
            return csx::set_error(CSX_ERR_HIP, "server: creating stream %zu's resources failed", (size_t)(&sl - s->slots.data()));
        }
    }
    *out = s;
    return csx::clear_error();
}

// Waits for in-flight batches; their completions are dropped.
extern "C" csx_status csx_server_destroy(csx_server* s)
{
    if (!s) return csx::set_error(CSX_ERR_INVALID, "server: null handle");
    for (auto& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.h_table) (void)hipHostFree(sl.h_table);
        if (sl.d_table) (void)hipFree(sl.d_table);
        if (sl.d_counts) (void)hipFree(sl.d_counts);
        if (sl.h_counts) (void)hipHostFree(sl.h_counts);
        if (sl.d_ws) (void)hipFree(sl.d_ws);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    delete s;
    return csx::clear_error();
}

extern "C" csx_status csx_server_submit(csx_server* s, const csx_request* r)
{
    if (!s || !r) return csx::set_error(CSX_ERR_INVALID, "server: null handle or request");
    if (r->elem_size != 1 && r->elem_size != 2 && r->elem_size != 4 && r->elem_size != 8)
        return csx::set_error(CSX_ERR_INVALID, "server: element size %u must be 1, 2, 4 or 8", r->elem_size);
    if (r->n > s->cfg.max_batch_elements)
        return csx::set_error(CSX_ERR_INVALID, "server: request of %llu elements exceeds max_batch_elements %llu", (unsigned long long)r->n,
                              (unsigned long long)s->cfg.max_batch_elements);
    if (r->n > 0 && (!r->in || !r->mask || !r->out)) return csx::set_error(CSX_ERR_INVALID, "server: null input, mask or output");
    s->queue.push_back(Pending{*r, now_ms(), s->next_seq++});
    return csx::clear_error();
}

extern "C" uint32_t csx_server_queued(const csx_server* s) { return s ? (uint32_t)s->queue.size() : 0; }
extern "C" uint32_t csx_server_inflight(const csx_server* s) { return s ? (uint32_t)s->inflight.size() : 0; }

// Launch batches from the queue while stream slots are free: each batch takes
// the oldest queued request's width and every queued request of that width,
// up to max_batch requests and max_batch_elements elements.  A failed launch
// drains the slot's stream and puts the batch's requests back in submission
// order.
static csx_status launch_batch(csx_server* s, Slot& sl, const Batch& b, uint64_t elems)
{
    hipError_t e = hipSuccess;
    hipEvent_t waited = nullptr;  // requests admitted together usually share one event
    for (size_t i = 0; i < b.reqs.size() && e == hipSuccess; i++) {
        const csx_request& r = b.reqs[i].req;
        sl.h_table[i] = csx_segment{r.in, r.mask, r.out, r.n};
        if (r.ready && r.ready != waited) {
            e = hipStreamWaitEvent(sl.stream, r.ready, 0);
            waited = r.ready;
        }
    }
    const uint32_t nseg = (uint32_t)b.reqs.size();
    if (e == hipSuccess) e = hipMemcpyAsync(sl.d_table, sl.h_table, sizeof(csx_segment) * nseg, hipMemcpyHostToDevice, sl.stream);
    if (e != hipSuccess) return csx::set_error(CSX_ERR_HIP, "server: batch table upload: %s", hipGetErrorString(e));
    // the batch's exact element total bounds the grids (workgroups past the
    // last tile only exit, but each is still dispatched)
    const csx_status st = csx_compact_batch(sl.d_table, nseg, elems, b.width, CSX_MSB_FIRST, sl.d_counts, sl.d_ws, sl.ws_bytes, sl.stream);
    if (st != CSX_OK) return st;
    e = hipMemcpyAsync(sl.h_counts, sl.d_counts, sizeof(uint64_t) * nseg, hipMemcpyDeviceToHost, sl.stream);
    if (e == hipSuccess) e = hipEventRecord(sl.done, sl.stream);
    if (e != hipSuccess) return csx::set_error(CSX_ERR_HIP, "server: batch completion: %s", hipGetErrorString(e));
    return CSX_OK;
}

extern "C" csx_status csx_server_step(csx_server* s, uint32_t* launched)
{
    if (!s) return csx::set_error(CSX_ERR_INVALID, "server: null handle");
    uint32_t nl = 0;
    while (!s->queue.empty()) {
        int slot = -1;
        for (uint32_t k = 0; k < s->slots.size(); k++) {
            const uint32_t i = (s->next_slot + k) % s->slots.size();
            if (!s->slots[i].busy) {
                slot = (int)i;
                break;
            }
        }
        if (slot < 0) break;
        Slot& sl = s->slots[slot];
        Batch b;
        b.slot = slot;
        b.width = s->queue.front().req.elem_size;
        uint64_t elems = 0;
        for (auto it = s->queue.begin(); it != s->queue.end() && b.reqs.size() < s->cfg.max_batch;) {
            if (it->req.elem_size == b.width && elems + it->req.n <= s->cfg.max_batch_elements) {
                elems += it->req.n;
                b.reqs.push_back(*it);
                it = s->queue.erase(it);
            }
            else {
                ++it;
            }
        }
        const csx_status st = launch_batch(s, sl, b, elems);
        if (st != CSX_OK) {
            // Part of the batch may already be queued on the slot's stream (the
            // table upload reads the pinned h_table), so the slot is drained
            // before any later step rewrites it.  The requests go back in
            // submission order: merged by sequence number with the requests
            // still queued, which stay in order themselves.
            (void)hipStreamSynchronize(sl.stream);
            std::deque<Pending> merged;
            auto q = s->queue.begin();
            for (const Pending& p : b.reqs) {
                while (q != s->queue.end() && q->seq < p.seq) merged.push_back(*q++);
                merged.push_back(p);
            }
            merged.insert(merged.end(), q, s->queue.end());
            s->queue.swap(merged);
            if (launched) *launched = nl;
            return st;
        }
        sl.busy = true;
        s->next_slot = (uint32_t)(slot + 1) % s->slots.size();
        s->inflight.push_back(std::move(b));
        s->launched_batches++;
        nl++;
    }
    if (launched) *launched = nl;
    return csx::clear_error();
}

// Retire finished batches (any order) without blocking.  `cap` must hold a
// whole batch (at least max_batch completions).
extern "C" csx_status csx_server_poll(csx_server* s, csx_completion* out, uint32_t cap, uint32_t* n)
{
    if (!s || !n || !out) return csx::set_error(CSX_ERR_INVALID, "server: null handle, completion buffer or count pointer");
    if (cap < s->cfg.max_batch) return csx::set_error(CSX_ERR_INVALID, "server: capacity %u is below max_batch %u", cap, s->cfg.max_batch);
    uint32_t k = 0;
    for (auto it = s->inflight.begin(); it != s->inflight.end();) {
        if (k + it->reqs.size() > cap) break;
        Slot& sl = s->slots[it->slot];
        const hipError_t q = hipEventQuery(sl.done);
        if (q == hipErrorNotReady) {
            ++it;
            continue;
        }
        if (q != hipSuccess) {
            *n = k;
            return csx::set_error(CSX_ERR_HIP, "server: batch on stream %d: %s", it->slot, hipGetErrorString(q));
        }
        const double t = now_ms();
        for (size_t i = 0; i < it->reqs.size(); i++) {
            out[k].id = it->reqs[i].req.id;
            out[k].count = sl.h_counts[i];
            out[k].latency_ms = t - it->reqs[i].submit_ms;
            k++;
        }
        sl.busy = false;
        it = s->inflight.erase(it);
    }
    *n = k;
    return csx::clear_error();
}
