// csx_compat.hpp -- the reference's C++ launcher and bench interface over the
// csx C ABI, for code written against RememberOfLife/gpu_compressstore2's
// src/kernels/kernel_3pass.cuh, src/kernels/kernel_pattern.cuh,
// src/cub_wraps.cuh and src/benchmarks.cuh.
//
// Same function names, argument order and meaning, same error behaviour
// (print and exit, src/cuda_try.cuh:8-20), same timing convention (every
// launcher synchronises, times its phase with an event pair on the legacy
// stream and returns milliseconds, src/cuda_time.cuh:11-21).  Differences:
//   * the block count of the LDS-staged writeout (optimized_writeout) and of
//     the engine's three-phase pipeline becomes the tile's workgroup size,
//     rounded to the nearest size built for gfx950 (128, 256, 512); the other
//     kernels run 256-thread workgroups.  Grid counts are accepted and
//     ignored: the tile kernels launch one workgroup per tile.
//     csx_last_launch() tells what a call launched;
//   * the selected count written to d_out_count is exact;
//   * prepare_buffers does not modify the mask (the kernels ignore bits past n);
//   * launchers whose reference signature has no workspace argument use a
//     per-thread scratch buffer that grows on demand (hipMalloc outside the
//     timed region of the call that grows it).
#pragma once

#include <csx.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <vector>

namespace csx_compat {

inline void fail_exit(const char* what, const char* detail)
{
    fprintf(stderr, "%s failed: %s\n", what, detail);
    exit(EXIT_FAILURE);
}

inline void check(csx_status s, const char* what)
{
    if (s != CSX_OK) fail_exit(what, csx_last_error());
}

inline void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) fail_exit(what, hipGetErrorString(e));
}

// CUDA_TIME on the legacy stream: synchronise, record, run, record, wait
template <class F> float time_phase(hipEvent_t start, hipEvent_t stop, F&& body, hipStream_t stream = 0)
{
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    hip_check(hipEventRecord(start, stream), "hipEventRecord");
    body();
    hip_check(hipEventRecord(stop, stream), "hipEventRecord");
    hip_check(hipEventSynchronize(stop), "hipEventSynchronize");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, start, stop), "hipEventElapsedTime");
    return ms;
}

// per-thread device scratch for launchers without a workspace argument
inline void* scratch(size_t bytes, int slot = 0)
{
    static thread_local void* ptr[4] = {nullptr, nullptr, nullptr, nullptr};
    static thread_local size_t cap[4] = {0, 0, 0, 0};
    if (cap[slot] < bytes) {
        if (ptr[slot]) hip_check(hipFree(ptr[slot]), "hipFree");
        hip_check(hipMalloc(&ptr[slot], bytes), "hipMalloc");
        cap[slot] = bytes;
    }
    return ptr[slot];
}

template <class T> constexpr uint32_t width() { return (uint32_t)sizeof(T); }

// a reference block size (threads per block, src/main.cu:38-46) as the
// engine's tile workgroup size: 0 (unset) is 256, anything else the nearest
// built size at or above it, clamped to [128, 512]
inline uint32_t engine_block(uint32_t threadcount)
{
    if (threadcount == 0) return 256;
    if (threadcount <= 128) return 128;
    return threadcount <= 256 ? 256 : 512;
}

// ------------------------------------------------------------ phase launchers

// src/kernels/kernel_3pass.cuh:40-57
inline float launch_3pass_popc_none(hipEvent_t ce_start, hipEvent_t ce_stop, uint32_t /*blockcount*/, uint32_t /*threadcount*/, uint8_t* d_mask,
                                    uint32_t* d_pss, uint32_t chunk_length, uint32_t element_count)
{
    return time_phase(ce_start, ce_stop,
                      [&] { check(csx_popc(d_mask, d_pss, element_count, chunk_length, CSX_MSB_FIRST, 0), "launch_3pass_popc_none"); });
}

// src/kernels/kernel_3pass.cuh:107-153: counts in d_pss become the up-swept
// tree; *d_out_count = the total (exact, modulo 2^32 as the reference's u32)
inline float launch_3pass_pss_gmem(hipEvent_t ce_start, hipEvent_t ce_stop, uint32_t /*blockcount*/, uint32_t /*threadcount*/, uint32_t* d_pss,
                                   uint32_t chunk_count, uint32_t* d_out_count)
{
    size_t ws_bytes = 0;
    check(csx_scan_workspace_size(chunk_count, &ws_bytes), "launch_3pass_pss_gmem");
    void* ws = scratch(ws_bytes + 256, 0);
    auto total = reinterpret_cast<uint64_t*>(static_cast<uint8_t*>(ws) + ws_bytes);
    return time_phase(ce_start, ce_stop, [&] {
        check(csx_scan_tree(d_pss, chunk_count, total, ws, ws_bytes, 0), "launch_3pass_pss_gmem");
        if (d_out_count) hip_check(hipMemcpyAsync(d_out_count, total, sizeof(uint32_t), hipMemcpyDeviceToDevice, 0), "hipMemcpyAsync");
    });
}

// src/kernels/kernel_3pass.cuh:188-214: exclusive offsets from the tree
inline float launch_3pass_pss2_gmem(hipEvent_t ce_start, hipEvent_t ce_stop, uint32_t /*blockcount*/, uint32_t /*threadcount*/, uint32_t* d_pss_in,
                                    uint32_t* d_pss_out, uint32_t chunk_count)
{
    return time_phase(ce_start, ce_stop, [&] { check(csx_tree_offsets(d_pss_in, d_pss_out, chunk_count, 0), "launch_3pass_pss2_gmem"); });
}

// src/cub_wraps.cuh:52-79: in-place exclusive scan of d_pss on `stream`;
// *d_pss_total = the exact total (u32)
inline float launch_cub_pss(hipStream_t stream, hipEvent_t ce_start, hipEvent_t ce_stop, uint32_t* d_pss, uint32_t* d_pss_total, uint32_t chunk_count)
{
    if (!chunk_count) return 0;
    size_t ws_bytes = 0;
    check(csx_scan_workspace_size(chunk_count, &ws_bytes), "launch_cub_pss");
    void* ws = scratch(ws_bytes + 256, 1);
    auto total = reinterpret_cast<uint64_t*>(static_cast<uint8_t*>(ws) + ws_bytes);
    auto body = [&] {
        check(csx_scan_u32(d_pss, d_pss, chunk_count, total, ws, ws_bytes, stream), "launch_cub_pss");
        if (d_pss_total) hip_check(hipMemcpyAsync(d_pss_total, total, sizeof(uint32_t), hipMemcpyDeviceToDevice, stream), "hipMemcpyAsync");
    };
    if (stream == 0) return time_phase(ce_start, ce_stop, body);
    body();
    return 0;
}

inline csx_writeout_args writeout_args(const void* in, const uint8_t* mask, void* out, const uint32_t* offsets, const uint32_t* counts,
                                       uint64_t n, uint32_t w, uint32_t chunk_length, int variant)
{
    csx_writeout_args a{};
    a.in = in;
    a.mask = mask;
    a.out = out;
    a.offsets = offsets;
    a.offsets_u32 = 1;
    a.counts = counts;
    a.out_base = nullptr;
    a.n = n;
    a.elem_size = w;
    a.chunk_len = chunk_length;
    a.bit_order = CSX_MSB_FIRST;
    a.variant = variant;
    return a;
}

// offsets for the writeout: the full exclusive scan, or (full_pss == false)
// the up-swept tree walked into a scratch array first
inline const uint32_t* full_offsets(uint32_t* d_pss, bool full_pss, uint32_t chunk_count)
{
    if (full_pss) return d_pss;
    auto tmp = static_cast<uint32_t*>(scratch(sizeof(uint32_t) * (chunk_count + 1), 2));
    check(csx_tree_offsets(d_pss, tmp, chunk_count, 0), "tree walk");
    return tmp;
}

// src/kernels/kernel_3pass.cuh:536-576 (optimized_writeout: LDS-staged stores)
template <typename T, bool optimized_writeout>
float launch_3pass_proc_true(hipEvent_t ce_start, hipEvent_t ce_stop, uint32_t /*blockcount*/, uint32_t threadcount, T* d_input, T* d_output,
                             uint8_t* d_mask, uint32_t* d_pss, bool full_pss, uint32_t* d_popc, uint32_t chunk_length, uint32_t element_count)
{
    const uint32_t chunks = (uint32_t)csx_chunk_count(element_count, chunk_length);
    return time_phase(ce_start, ce_stop, [&] {
        csx_writeout_args a = writeout_args(d_input, d_mask, d_output, full_offsets(d_pss, full_pss, chunks), d_popc, element_count, width<T>(),
                                            chunk_length, optimized_writeout ? CSX_WRITEOUT_STAGED : CSX_WRITEOUT_DIRECT);
        a.block_size = optimized_writeout ? engine_block(threadcount) : 256;
        check(csx_writeout_ex(&a, 0), "launch_3pass_proc_true");
    });
}

// src/kernels/kernel_3pass.cuh:642-693 (thread per chunk, bit-serial)
template <typename T>
float launch_3pass_proc_none(hipEvent_t ce_start, hipEvent_t ce_stop, uint32_t /*blockcount*/, uint32_t /*threadcount*/, T* d_input, T* d_output,
                             uint8_t* d_mask, uint32_t* d_pss, bool full_pss, uint32_t* d_popc, uint32_t chunk_length, uint32_t element_count)
{
    const uint32_t chunks = (uint32_t)csx_chunk_count(element_count, chunk_length);
    return time_phase(ce_start, ce_stop, [&] {
        const csx_writeout_args a = writeout_args(d_input, d_mask, d_output, full_offsets(d_pss, full_pss, chunks), d_popc, element_count,
                                                  width<T>(), chunk_length, CSX_WRITEOUT_THREAD_PER_CHUNK);
        check(csx_writeout_ex(&a, 0), "launch_3pass_proc_none");
    });
}

// src/kernels/kernel_pattern.cuh:71-131
template <typename T>
float launch_pattern_proc(hipEvent_t ce_start, hipEvent_t ce_stop, uint32_t /*blockcount*/, uint32_t /*threadcount*/, T* d_input, T* d_output,
                          uint64_t N, uint32_t pattern, int pattern_length, uint32_t /*chunk_length*/)
{
    return time_phase(ce_start, ce_stop, [&] {
        check(csx_pattern_compact(d_input, d_output, N, width<T>(), pattern, (uint32_t)pattern_length, nullptr, nullptr, 0), "launch_pattern_proc");
    });
}

// ------------------------------------------------------------ timings

// src/benchmarks.cuh:10-52: -1 marks a phase the bench does not have
struct timings {
    float popc, pss1, pss2, proc, total;
    timings() : popc(-1), pss1(-1), pss2(-1), proc(-1), total(-1) {}
    static float add_timings(float a, float b) { return a < 0 ? b : (b < 0 ? a : a + b); }
    timings& operator/(float r)
    {
        popc /= r;
        pss1 /= r;
        pss2 /= r;
        proc /= r;
        total /= r;
        return *this;
    }
    void operator+=(timings r)
    {
        popc = add_timings(popc, r.popc);
        pss1 = add_timings(pss1, r.pss1);
        pss2 = add_timings(pss2, r.pss2);
        proc = add_timings(proc, r.proc);
        total = add_timings(total, r.total);
    }
};

inline std::ostream& operator<<(std::ostream& os, timings t)
{
    os << std::setprecision(6) << t.popc << ";" << t.pss1 << ";" << t.pss2 << ";" << t.proc << ";" << t.total;
    return os;
}

// ------------------------------------------------------------ intermediate_data

// src/benchmarks.cuh:53-151: everything a bench needs besides its inputs
struct intermediate_data {
    uint32_t* d_pss;
    uint32_t* d_pss2;
    uint32_t* d_popc;
    uint64_t* d_out_count;  // [max_stream_count + 1]; [0] holds the exact selected count
    uint64_t* d_failure_count;
    void* d_workspace;      // csx_compact / csx_compact_streams workspace
    size_t workspace_size;
    size_t element_count;
    size_t chunk_length_min;
    size_t max_stream_count;
    size_t elem_size;
    hipEvent_t dummy_event_1, dummy_event_2, start, stop;
    std::vector<hipStream_t> streams;

    template <typename T> intermediate_data(size_t element_count, int chunk_length_min, int max_stream_count, T* /*data_type_dummy*/)
    {
        this->element_count = element_count;
        // the streaming bench always runs at chunk length 1024 (src/main.cu:141-143)
        this->chunk_length_min = (size_t)(chunk_length_min < 1024 ? chunk_length_min : 1024);
        chunk_length_min = (int)this->chunk_length_min;
        this->max_stream_count = (size_t)max_stream_count;
        this->elem_size = sizeof(T);
        const size_t chunks = chunk_count(chunk_length_min) + 1;
        hip_check(hipMalloc(&d_pss, chunks * sizeof(uint32_t)), "hipMalloc");
        hip_check(hipMalloc(&d_pss2, chunks * sizeof(uint32_t)), "hipMalloc");
        hip_check(hipMalloc(&d_popc, chunks * sizeof(uint32_t)), "hipMalloc");
        hip_check(hipMalloc(&d_out_count, (max_stream_count + 1) * sizeof(uint64_t)), "hipMalloc");
        hip_check(hipMalloc(&d_failure_count, sizeof(uint64_t)), "hipMalloc");
        size_t a = 0, b = 0;
        check(csx_workspace_size(element_count, sizeof(T), chunk_length_min, &a), "csx_workspace_size");
        check(csx_streams_workspace_size(element_count, sizeof(T), chunk_length_min, max_stream_count, &b), "csx_streams_workspace_size");
        workspace_size = a > b ? a : b;
        hip_check(hipMalloc(&d_workspace, workspace_size), "hipMalloc");
        hip_check(hipEventCreate(&dummy_event_1), "hipEventCreate");
        hip_check(hipEventCreate(&dummy_event_2), "hipEventCreate");
        hip_check(hipEventCreate(&start), "hipEventCreate");
        hip_check(hipEventCreate(&stop), "hipEventCreate");
        streams.resize(max_stream_count);
        for (auto& s : streams) hip_check(hipStreamCreate(&s), "hipStreamCreate");
    }
    intermediate_data(const intermediate_data&) = delete;
    intermediate_data& operator=(const intermediate_data&) = delete;
    ~intermediate_data()
    {
        (void)hipFree(d_pss);
        (void)hipFree(d_pss2);
        (void)hipFree(d_popc);
        (void)hipFree(d_out_count);
        (void)hipFree(d_failure_count);
        (void)hipFree(d_workspace);
        (void)hipEventDestroy(dummy_event_1);
        (void)hipEventDestroy(dummy_event_2);
        (void)hipEventDestroy(start);
        (void)hipEventDestroy(stop);
        for (auto s : streams) (void)hipStreamDestroy(s);
    }
    // src/benchmarks.cuh:147-150 (element_count rounded up to whole mask bytes)
    size_t chunk_count(size_t chunk_length) const { return ((element_count + 7) / 8 * 8 + chunk_length - 1) / chunk_length; }

    // src/benchmarks.cuh:109-128, without touching the mask
    template <typename T> void prepare_buffers(size_t n, size_t chunk_length, T* d_output, uint8_t* /*d_mask*/)
    {
        if (n > element_count || (chunk_length && chunk_length < chunk_length_min))
            fail_exit("prepare_buffers", "sizes in intermediate data are smaller than the ones submitted to the algorithm");
        hip_check(hipMemset(d_out_count, 0, (max_stream_count + 1) * sizeof(uint64_t)), "hipMemset");
        hip_check(hipMemset(d_output, 0xFF, n * sizeof(T)), "hipMemset");
    }
};

// ------------------------------------------------------------ bench compositions
// Each returns the reference's timings; [0] of d_out_count is the exact count.

// bench1 (src/benchmarks.cuh:153-172): popc -> up-sweep tree -> tree walk -> thread per chunk
template <class T>
timings bench1_base_variant(intermediate_data* id, T* d_input, uint8_t* d_mask, T* d_output, size_t element_count, size_t chunk_length, int block_size,
                            int grid_size)
{
    id->prepare_buffers(element_count, chunk_length, d_output, d_mask);
    timings t{};
    const uint32_t chunks = (uint32_t)csx_chunk_count(element_count, chunk_length);
    t.total = time_phase(id->start, id->stop, [&] {
        t.popc = launch_3pass_popc_none(id->dummy_event_1, id->dummy_event_2, grid_size, block_size, d_mask, id->d_pss, chunk_length, element_count);
        t.pss1 = launch_3pass_pss_gmem(id->dummy_event_1, id->dummy_event_2, grid_size, block_size, id->d_pss, chunks,
                                       reinterpret_cast<uint32_t*>(id->d_out_count));
        t.pss2 = launch_3pass_pss2_gmem(id->dummy_event_1, id->dummy_event_2, grid_size, block_size, id->d_pss, id->d_pss2, chunks);
        t.proc = launch_3pass_proc_none<T>(id->dummy_event_1, id->dummy_event_2, grid_size, block_size, d_input, d_output, d_mask, id->d_pss2, true,
                                           nullptr, chunk_length, element_count);
    });
    return t;
}

// bench2 (src/benchmarks.cuh:174-193): as bench1, skipping chunks whose tree entry is 0
template <class T>
timings bench2_base_variant_skipping(intermediate_data* id, T* d_input, uint8_t* d_mask, T* d_output, size_t element_count, size_t chunk_length,
                                     int block_size, int grid_size)
{
    id->prepare_buffers(element_count, chunk_length, d_output, d_mask);
    timings t{};
    const uint32_t chunks = (uint32_t)csx_chunk_count(element_count, chunk_length);
    t.total = time_phase(id->start, id->stop, [&] {
        t.popc = launch_3pass_popc_none(id->dummy_event_1, id->dummy_event_2, grid_size, block_size, d_mask, id->d_pss, chunk_length, element_count);
        t.pss1 = launch_3pass_pss_gmem(id->dummy_event_1, id->dummy_event_2, grid_size, block_size, id->d_pss, chunks,
                                       reinterpret_cast<uint32_t*>(id->d_out_count));
        t.pss2 = launch_3pass_pss2_gmem(id->dummy_event_1, id->dummy_event_2, grid_size, block_size, id->d_pss, id->d_pss2, chunks);
        t.proc = launch_3pass_proc_none<T>(id->dummy_event_1, id->dummy_event_2, grid_size, block_size, d_input, d_output, d_mask, id->d_pss2, true,
                                           id->d_pss, chunk_length, element_count);
    });
    return t;
}

// bench3 (src/benchmarks.cuh:195-293): split across streams with a carry
template <class T>
timings bench3_3pass_streaming(intermediate_data* id, T* d_input, uint8_t* d_mask, T* d_output, size_t element_count, size_t chunk_length,
                               int /*block_size*/, int /*grid_size*/, int stream_count = 2)
{
    id->prepare_buffers(element_count, chunk_length, d_output, d_mask);
    if (stream_count < 1 || (size_t)stream_count > id->max_stream_count) fail_exit("bench3_3pass_streaming", "stream_count out of range");
    timings t{};
    t.total = time_phase(id->start, id->stop, [&] {
        check(csx_compact_streams(d_input, d_mask, d_output, element_count, sizeof(T), chunk_length, CSX_MSB_FIRST, id->streams.data(),
                                  (uint32_t)stream_count, id->d_out_count, nullptr, id->d_workspace, id->workspace_size),
              "bench3_3pass_streaming");
        hip_check(hipStreamSynchronize(id->streams[0]), "hipStreamSynchronize");
    });
    return t;
}

// the CUB-scan benches' common prefix: popc -> (copy) -> in-place u32 scan
template <class T>
void popc_and_scan(intermediate_data* id, timings& t, uint8_t* d_mask, size_t element_count, size_t chunk_length, bool keep_popc, int bs, int gs)
{
    const uint32_t chunks = (uint32_t)csx_chunk_count(element_count, chunk_length);
    uint32_t* counts = keep_popc ? id->d_popc : id->d_pss;
    t.popc = launch_3pass_popc_none(id->dummy_event_1, id->dummy_event_2, gs, bs, d_mask, counts, chunk_length, element_count);
    if (keep_popc) hip_check(hipMemcpy(id->d_pss, id->d_popc, chunks * sizeof(uint32_t), hipMemcpyDeviceToDevice), "hipMemcpy");
    size_t ws_bytes = 0;
    check(csx_scan_workspace_size(chunks, &ws_bytes), "scan");
    void* ws = scratch(ws_bytes, 3);
    check(csx_scan_u32(id->d_pss, id->d_pss2, chunks, id->d_out_count, ws, ws_bytes, 0), "scan");
}

// bench4 (src/benchmarks.cuh:295-316): scan, naive writeout, no skipping
template <class T>
timings bench4_3pass_optimized_read_non_skipping_cub_pss(intermediate_data* id, T* d_input, uint8_t* d_mask, T* d_output, size_t element_count,
                                                         size_t chunk_length, int block_size, int grid_size)
{
    id->prepare_buffers(element_count, chunk_length, d_output, d_mask);
    timings t{};
    t.total = time_phase(id->start, id->stop, [&] {
        popc_and_scan<T>(id, t, d_mask, element_count, chunk_length, false, block_size, grid_size);
        t.proc = launch_3pass_proc_true<T, false>(id->dummy_event_1, id->dummy_event_2, grid_size, block_size, d_input, d_output, d_mask, id->d_pss2,
                                                  true, nullptr, chunk_length, element_count);
    });
    return t;
}

// bench5 (src/benchmarks.cuh:318-336): up-swept tree consumed by the writeout (partial pss)
template <class T>
timings bench5_3pass_optimized_read_skipping_partial_pss(intermediate_data* id, T* d_input, uint8_t* d_mask, T* d_output, size_t element_count,
                                                         size_t chunk_length, int block_size, int grid_size)
{
    id->prepare_buffers(element_count, chunk_length, d_output, d_mask);
    timings t{};
    const uint32_t chunks = (uint32_t)csx_chunk_count(element_count, chunk_length);
    t.total = time_phase(id->start, id->stop, [&] {
        t.popc = launch_3pass_popc_none(id->dummy_event_1, id->dummy_event_2, grid_size, block_size, d_mask, id->d_popc, chunk_length, element_count);
        hip_check(hipMemcpy(id->d_pss, id->d_popc, chunks * sizeof(uint32_t), hipMemcpyDeviceToDevice), "hipMemcpy");
        t.pss1 = launch_3pass_pss_gmem(id->dummy_event_1, id->dummy_event_2, grid_size, block_size, id->d_pss, chunks,
                                       reinterpret_cast<uint32_t*>(id->d_out_count));
        t.proc = launch_3pass_proc_true<T, false>(id->dummy_event_1, id->dummy_event_2, grid_size, block_size, d_input, d_output, d_mask, id->d_pss,
                                                  false, id->d_popc, chunk_length, element_count);
    });
    return t;
}

// bench6 (src/benchmarks.cuh:338-358): tree + tree walk, naive writeout with skipping
template <class T>
timings bench6_3pass_optimized_read_skipping_two_phase_pss(intermediate_data* id, T* d_input, uint8_t* d_mask, T* d_output, size_t element_count,
                                                           size_t chunk_length, int block_size, int grid_size)
{
    id->prepare_buffers(element_count, chunk_length, d_output, d_mask);
    timings t{};
    const uint32_t chunks = (uint32_t)csx_chunk_count(element_count, chunk_length);
    t.total = time_phase(id->start, id->stop, [&] {
        t.popc = launch_3pass_popc_none(id->dummy_event_1, id->dummy_event_2, grid_size, block_size, d_mask, id->d_popc, chunk_length, element_count);
        hip_check(hipMemcpy(id->d_pss, id->d_popc, chunks * sizeof(uint32_t), hipMemcpyDeviceToDevice), "hipMemcpy");
        t.pss1 = launch_3pass_pss_gmem(id->dummy_event_1, id->dummy_event_2, grid_size, block_size, id->d_pss, chunks,
                                       reinterpret_cast<uint32_t*>(id->d_out_count));
        t.pss2 = launch_3pass_pss2_gmem(id->dummy_event_1, id->dummy_event_2, grid_size, block_size, id->d_pss, id->d_pss2, chunks);
        t.proc = launch_3pass_proc_true<T, false>(id->dummy_event_1, id->dummy_event_2, grid_size, block_size, d_input, d_output, d_mask, id->d_pss2,
                                                  true, id->d_popc, chunk_length, element_count);
    });
    return t;
}

// bench7 (src/benchmarks.cuh:360-382): scan, naive writeout with skipping
template <class T>
timings bench7_3pass_optimized_read_skipping_cub_pss(intermediate_data* id, T* d_input, uint8_t* d_mask, T* d_output, size_t element_count,
                                                     size_t chunk_length, int block_size, int grid_size)
{
    id->prepare_buffers(element_count, chunk_length, d_output, d_mask);
    timings t{};
    t.total = time_phase(id->start, id->stop, [&] {
        popc_and_scan<T>(id, t, d_mask, element_count, chunk_length, true, block_size, grid_size);
        t.proc = launch_3pass_proc_true<T, false>(id->dummy_event_1, id->dummy_event_2, grid_size, block_size, d_input, d_output, d_mask, id->d_pss2,
                                                  true, id->d_popc, chunk_length, element_count);
    });
    return t;
}

// bench9 (src/benchmarks.cuh:396-416): periodic pattern, no mask
template <class T>
timings bench9_pattern(intermediate_data* id, T* d_input, uint32_t pattern, int pattern_length, T* d_output, size_t element_count, size_t chunk_length,
                       int block_size, int grid_size)
{
    id->prepare_buffers(element_count, 0, d_output, (uint8_t*)nullptr);
    timings t{};
    t.total = launch_pattern_proc(id->dummy_event_1, id->dummy_event_2, grid_size, block_size, d_input, d_output, element_count, pattern, pattern_length,
                                  chunk_length);
    uint64_t k = 0;
    check(csx_pattern_count(element_count, pattern, (uint32_t)pattern_length, &k), "bench9_pattern");
    hip_check(hipMemcpy(id->d_out_count, &k, sizeof(k), hipMemcpyHostToDevice), "hipMemcpy");
    return t;
}

// bench10 (src/benchmarks.cuh:418-440): scan, LDS-staged writeout with skipping
template <class T>
timings bench10_3pass_optimized_read_skipping_optimized_writeout_cub_pss(intermediate_data* id, T* d_input, uint8_t* d_mask, T* d_output,
                                                                          size_t element_count, size_t chunk_length, int block_size, int grid_size)
{
    id->prepare_buffers(element_count, chunk_length, d_output, d_mask);
    timings t{};
    t.total = time_phase(id->start, id->stop, [&] {
        popc_and_scan<T>(id, t, d_mask, element_count, chunk_length, true, block_size, grid_size);
        t.proc = launch_3pass_proc_true<T, true>(id->dummy_event_1, id->dummy_event_2, grid_size, block_size, d_input, d_output, d_mask, id->d_pss2,
                                                 true, id->d_popc, chunk_length, element_count);
    });
    return t;
}

// the engine's own pipelines behind one call (no reference counterpart); the
// three-phase pipeline's writeout tile runs engine_block(block_size) threads
template <class T>
timings bench_csx_pipeline(intermediate_data* id, int pipeline, T* d_input, uint8_t* d_mask, T* d_output, size_t element_count, size_t chunk_length,
                           int block_size = 0)
{
    id->prepare_buffers(element_count, chunk_length, d_output, d_mask);
    csx_config cfg{};
    cfg.pipeline = pipeline;
    cfg.chunk_len = (uint32_t)chunk_length;
    cfg.vpt = 8;
    cfg.block_size = pipeline == CSX_PIPELINE_THREE_PHASE ? engine_block((uint32_t)block_size) : 256;
    timings t{};
    t.total = time_phase(id->start, id->stop, [&] {
        check(csx_compact_config(d_input, d_mask, d_output, element_count, sizeof(T), CSX_MSB_FIRST, &cfg, id->d_out_count, nullptr, id->d_workspace,
                                 id->workspace_size, 0),
              "csx_compact");
    });
    return t;
}

// Mismatch report of the reference's validation with report_failures
// (src/kernels/data_generator.cuh:136-147): one line per differing element,
// "failure: index: I: expected: E, got: G" with E and G the elements' bits
// in hex, here in index order and at most `limit` lines, then the number of
// further mismatches.  Host arrays; returns the number of mismatches.
template <typename T>
uint64_t report_mismatches(const T* expected, const T* got, uint64_t len, uint64_t limit, FILE* f)
{
    uint64_t bad = 0;
    for (uint64_t i = 0; i < len; i++) {
        uint64_t e = 0, g = 0;
        memcpy(&e, &expected[i], sizeof(T));
        memcpy(&g, &got[i], sizeof(T));
        if (e == g) continue;
        if (bad < limit) fprintf(f, "failure: index: %llu: expected: %llx, got: %llx\n", (unsigned long long)i, (unsigned long long)e, (unsigned long long)g);
        bad++;
    }
    if (bad > limit) fprintf(f, "... %llu more mismatching elements\n", (unsigned long long)(bad - limit));
    return bad;
}

// Lines of mismatch report per failed validation (the reference prints all).
constexpr uint64_t kReportLimit = 64;

// src/benchmarks.cuh:442-450: the device counts the mismatches; with
// report_failures a failed validation copies both arrays back and reports
// them element by element (report_mismatches).
template <typename T>
bool validate(intermediate_data* id, T* d_validation, T* d_output, uint64_t out_length, bool report_failures, uint64_t* failure_count = nullptr)
{
    hip_check(hipMemset(id->d_failure_count, 0, sizeof(uint64_t)), "hipMemset");
    check(csx_validate(d_validation, d_output, out_length, sizeof(T), id->d_failure_count, 0), "validate");
    uint64_t fc = 0;
    hip_check(hipMemcpy(&fc, id->d_failure_count, sizeof(fc), hipMemcpyDeviceToHost), "hipMemcpy");
    if (report_failures && fc) {
        std::vector<T> e(out_length), g(out_length);
        hip_check(hipMemcpy(e.data(), d_validation, out_length * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
        hip_check(hipMemcpy(g.data(), d_output, out_length * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
        report_mismatches(e.data(), g.data(), out_length, kReportLimit, stderr);
    }
    if (failure_count) *failure_count = fc;
    return fc == 0;
}

} // namespace csx_compat
