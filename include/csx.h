/*
 * csx.h -- C ABI of the MI355X (gfx950) bitmask stream-compaction engine.
 *
 * Replaces the launch functions of RememberOfLife/gpu_compressstore2's
 * three-phase pipeline (header-only CUDA templates, no library of its own):
 *   phase 1  launch_3pass_popc_none        src/kernels/kernel_3pass.cuh:40-57
 *   phase 2  launch_3pass_pss_gmem         src/kernels/kernel_3pass.cuh:107-153
 *            launch_3pass_pss2_gmem        src/kernels/kernel_3pass.cuh:188-214
 *            launch_cub_pss / CUB ExclusiveSum  src/cub_wraps.cuh:52-79,
 *                                          src/benchmarks.cuh:308,374,432
 *   phase 3  launch_3pass_proc_true<T,opt> src/kernels/kernel_3pass.cuh:536-576
 *            launch_3pass_proc_none<T>     src/kernels/kernel_3pass.cuh:642-693
 *   pipeline benchN(...)                   src/benchmarks.cuh:153-440
 *
 * Conventions (differences from the reference are deliberate):
 *   - Every call is asynchronous on the caller's stream and returns a status
 *     code instead of exiting (reference: CUDA_TRY -> exit,
 *     src/cuda_try.cuh:8-20).
 *   - The mask is const: bits at element index >= n are ignored in-kernel
 *     instead of being cleared in the caller's buffer
 *     (reference: prepare_buffers, src/benchmarks.cuh:109-121).
 *   - Counts are 32-bit per chunk, offsets and totals 64-bit (the reference
 *     uses u32 throughout, capping N below 2^32).  The returned selected
 *     count is exact (the reference's CUB benches report twice the last
 *     chunk's count, src/kernels/kernel_3pass.cuh:59-67).
 *   - No allocation on any compute call: scratch lives in a caller-owned
 *     workspace sized by the *_workspace_size queries (the two-call pattern
 *     the reference uses for CUB, src/benchmarks.cuh:83-90).
 *   - Element i of the mask is bit 7-(i%8) of byte i/8 with CSX_MSB_FIRST
 *     (the reference's order, src/cub_wraps.cuh:20-24), bit i%8 with
 *     CSX_LSB_FIRST.
 *   - Elements are opaque 1/2/4/8-byte words (float32/float64 included): the
 *     output is a bit-exact copy of the selected inputs in input order.
 *   - Pointers to elements must be aligned to the element size.
 */
#ifndef CSX_H
#define CSX_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CSX_OK = 0,
    CSX_ERR_INVALID = 1,     /* bad argument (null pointer, size, chunk length, width, order) */
    CSX_ERR_WORKSPACE = 2,   /* workspace missing or smaller than the size query returned */
    CSX_ERR_HIP = 3,         /* a HIP runtime call failed; see csx_last_error() */
    CSX_ERR_UNSUPPORTED = 4, /* valid request this build does not implement */
    CSX_ERR_DEVICE = 5       /* a kernel reported an internal failure (csx_workspace_error) */
} csx_status;

typedef enum { CSX_MSB_FIRST = 0, CSX_LSB_FIRST = 1 } csx_bit_order;

/* phase-3 writeout variants */
typedef enum {
    CSX_WRITEOUT_STAGED = 0,           /* ranks from mask words, LDS staging, 16-byte coalesced stores
                                          (replaces the optimized writeout, kernel_3pass.cuh:329-467) */
    CSX_WRITEOUT_DIRECT = 1,           /* same ranks, each element stored straight from registers
                                          (replaces the naive writeout, kernel_3pass.cuh:216-327) */
    CSX_WRITEOUT_THREAD_PER_CHUNK = 2, /* one thread walks one chunk bit by bit
                                          (replaces proc_none, kernel_3pass.cuh:578-639) */
    CSX_WRITEOUT_SPARSE = 3            /* no LDS: a lane per 128 elements gathers its selected elements one by one,
                                          ranks from a wave scan; the fastest below a few percent selectivity.
                                          No counterpart in the reference. */
} csx_writeout_variant;

/* whole-pipeline forms of csx_compact */
typedef enum {
    CSX_PIPELINE_FUSED = 0,       /* one pass: mask popcount, decoupled look-back and staged writeout per tile */
    CSX_PIPELINE_THREE_PHASE = 1, /* csx_popc -> csx_scan -> csx_writeout (STAGED), the reference's bench10 shape;
                                     the fastest form at low and medium selectivity (default of the bindings) */
    CSX_PIPELINE_TWO_PASS = 2,    /* pass 1 counts every tile and group of tiles from the mask; pass 2 is the staged
                                     writeout, each tile reducing its own offset from those counts */
    CSX_PIPELINE_SPARSE = 3,      /* csx_popc -> csx_scan -> csx_writeout (SPARSE): for low selectivity */
    CSX_PIPELINE_SINGLE = 4       /* one launch: each tile counts the whole mask before it for its own output offset,
                                     so no other kernel runs first; quadratic in n, for columns of at most
                                     CSX_SINGLE_MAX_ELEMENTS (mid-size calls, where launches, not bytes, cost time).
                                     No counterpart in the reference. */
} csx_pipeline;
#define CSX_SINGLE_MAX_ELEMENTS (UINT64_C(1) << 24)

/* Largest element count accepted (offsets are 62-bit in the look-back status words). */
#define CSX_MAX_ELEMENTS (UINT64_C(1) << 40)

const char* csx_status_string(csx_status s);
/* Message of the last failing call on this host thread ("" if none). */
const char* csx_last_error(void);
/* Version as major*10000 + minor*100 + patch. */
int csx_version(void);

/* Chunk lengths must be a power of two >= 32 (src/main.cu:236-245). */
uint64_t csx_chunk_count(uint64_t n, uint32_t chunk_len);

/* ---------------------------------------------------------------- phase 1
 * d_counts[c] = number of selected elements in [c*chunk_len, min((c+1)*chunk_len, n)),
 * for c < csx_chunk_count(n, chunk_len).
 * Replaces launch_3pass_popc_none (src/kernels/kernel_3pass.cuh:40-57). */
csx_status csx_popc(const uint8_t* d_mask, uint32_t* d_counts, uint64_t n, uint32_t chunk_len, int bit_order, hipStream_t stream);

/* ---------------------------------------------------------------- phase 2
 * d_offsets[c] = sum of d_counts[0..c) (exclusive), *d_total = sum of all
 * (d_total nullable).  Single-pass decoupled look-back scan.
 * Replaces launch_3pass_pss_gmem + launch_3pass_pss2_gmem
 * (src/kernels/kernel_3pass.cuh:107-214) and the CUB ExclusiveSum +
 * pssskip sequence (src/benchmarks.cuh:306-309). */
csx_status csx_scan_workspace_size(uint64_t chunks, size_t* bytes);
csx_status csx_scan(const uint32_t* d_counts, uint64_t* d_offsets, uint64_t chunks, uint64_t* d_total, void* d_ws, size_t ws_bytes,
                    hipStream_t stream);

/* The same scan with the reference's 32-bit offsets (its d_pss2 arrays,
 * src/benchmarks.cuh:91-93); d_offsets may equal d_counts (in place, as the
 * CUB ExclusiveSum + copy-back of launch_cub_pss, src/cub_wraps.cuh:52-79).
 * Offsets wrap modulo 2^32 past 2^32 - 1 selected elements. */
csx_status csx_scan_u32(const uint32_t* d_counts, uint32_t* d_offsets, uint64_t chunks, uint64_t* d_total, void* d_ws, size_t ws_bytes,
                        hipStream_t stream);

/* The reference's in-place up-sweep tree (src/kernels/kernel_3pass.cuh:69-153):
 * d_pss[j] becomes the sum of counts (j - lowbit(j+1), j]; *d_total (nullable)
 * the exact sum.  d_pss is read as counts and overwritten.  Workspace as csx_scan. */
csx_status csx_scan_tree(uint32_t* d_pss, uint64_t chunks, uint64_t* d_total, void* d_ws, size_t ws_bytes, hipStream_t stream);
/* Exclusive offsets from that tree, the tree walk of d_3pass_pproc_pssidx
 * (src/kernels/kernel_3pass.cuh:155-185). */
csx_status csx_tree_offsets(const uint32_t* d_tree, uint32_t* d_offsets, uint64_t chunks, hipStream_t stream);

/* ---------------------------------------------------------------- phase 3
 * Writes the selected elements of d_in to d_out in input order, element
 * d_offsets[c] + r for the r-th selected element of chunk c.
 * d_counts (nullable) enables zero-chunk skipping: chunks whose count is 0
 * are read neither from the mask nor from d_in (src/kernels/kernel_3pass.cuh:268-288).
 * Replaces launch_3pass_proc_true<T, optimized> (kernel_3pass.cuh:536-576)
 * and launch_3pass_proc_none<T> (kernel_3pass.cuh:642-693). */
csx_status csx_writeout(const void* d_in, const uint8_t* d_mask, const uint64_t* d_offsets, const uint32_t* d_counts, void* d_out, uint64_t n,
                        uint32_t elem_size, uint32_t chunk_len, int bit_order, int variant, hipStream_t stream);
/* All writeout options in one argument block:
 *   offsets_u32  the offsets are the reference's uint32_t d_pss values
 *   out_base     nullable device pointer; its value is added to every output
 *                index, the `offset` argument of the reference's
 *                switch_3pass_proc_true_striding (kernel_3pass.cuh:491-504,
 *                used by its multi-stream split, src/benchmarks.cuh:284-287) */
typedef struct {
    const void* in;
    const uint8_t* mask;
    void* out;
    const void* offsets;
    int offsets_u32;
    const uint32_t* counts;
    const uint64_t* out_base;
    uint64_t n;
    uint32_t elem_size;
    uint32_t chunk_len;
    int bit_order;
    int variant;
    uint32_t block_size; /* workgroup size of the STAGED variant's tile: 128, 256 or 512 (0: 256); the
                            reference's block size (switch_3pass_proc_true_striding, kernel_3pass.cuh:491-533) */
} csx_writeout_args;
csx_status csx_writeout_ex(const csx_writeout_args* args, hipStream_t stream);
csx_status csx_writeout_w1(const uint8_t* d_in, const uint8_t* d_mask, const uint64_t* d_offsets, const uint32_t* d_counts, uint8_t* d_out,
                           uint64_t n, uint32_t chunk_len, int bit_order, int variant, hipStream_t stream);
csx_status csx_writeout_w2(const uint16_t* d_in, const uint8_t* d_mask, const uint64_t* d_offsets, const uint32_t* d_counts, uint16_t* d_out,
                           uint64_t n, uint32_t chunk_len, int bit_order, int variant, hipStream_t stream);
csx_status csx_writeout_w4(const uint32_t* d_in, const uint8_t* d_mask, const uint64_t* d_offsets, const uint32_t* d_counts, uint32_t* d_out,
                           uint64_t n, uint32_t chunk_len, int bit_order, int variant, hipStream_t stream);
csx_status csx_writeout_w8(const uint64_t* d_in, const uint8_t* d_mask, const uint64_t* d_offsets, const uint32_t* d_counts, uint64_t* d_out,
                           uint64_t n, uint32_t chunk_len, int bit_order, int variant, hipStream_t stream);

/* ---------------------------------------------------------------- pipeline
 * Full compaction: d_out[0..k) = selected elements of d_in[0..n) in order,
 * k written to *d_count (device, nullable) and, if h_count is not null, to
 * *h_count after synchronising the stream.  d_out must hold k elements
 * (n is always enough); nothing at or past d_out[k] is written.  Not in
 * place: d_out's n elements must not overlap d_in or d_mask (checked, as
 * for every entry point below that takes both; CSX_ERR_INVALID).
 * chunk_len only matters for CSX_PIPELINE_THREE_PHASE and CSX_PIPELINE_SPARSE,
 * whose popcount counts units of min(chunk_len, writeout tile) elements.
 * CSX_PIPELINE_THREE_PHASE is the default of the language bindings.
 * Replaces the benchN compositions (src/benchmarks.cuh:153-440) -- bench10
 * is CSX_PIPELINE_THREE_PHASE with the staged writeout. */
csx_status csx_workspace_size(uint64_t n, uint32_t elem_size, uint32_t chunk_len, size_t* bytes);
csx_status csx_compact(const void* d_in, const uint8_t* d_mask, void* d_out, uint64_t n, uint32_t elem_size, uint32_t chunk_len, int bit_order,
                       int pipeline, uint64_t* d_count, uint64_t* h_count, void* d_ws, size_t ws_bytes, hipStream_t stream);
csx_status csx_compact_w1(const uint8_t* d_in, const uint8_t* d_mask, uint8_t* d_out, uint64_t n, uint32_t chunk_len, int bit_order,
                          int pipeline, uint64_t* d_count, uint64_t* h_count, void* d_ws, size_t ws_bytes, hipStream_t stream);
csx_status csx_compact_w2(const uint16_t* d_in, const uint8_t* d_mask, uint16_t* d_out, uint64_t n, uint32_t chunk_len, int bit_order,
                          int pipeline, uint64_t* d_count, uint64_t* h_count, void* d_ws, size_t ws_bytes, hipStream_t stream);
csx_status csx_compact_w4(const uint32_t* d_in, const uint8_t* d_mask, uint32_t* d_out, uint64_t n, uint32_t chunk_len, int bit_order,
                          int pipeline, uint64_t* d_count, uint64_t* h_count, void* d_ws, size_t ws_bytes, hipStream_t stream);
csx_status csx_compact_w8(const uint64_t* d_in, const uint8_t* d_mask, uint64_t* d_out, uint64_t n, uint32_t chunk_len, int bit_order,
                          int pipeline, uint64_t* d_count, uint64_t* h_count, void* d_ws, size_t ws_bytes, hipStream_t stream);

/* Multi-stream split: the column is cut into stream_count slices of whole
 * max(tile, chunk_len) units; slice i runs popcount, scan and writeout on
 * streams[i], concurrently with the others, and its output offset is the
 * carry base[i] = base[i-1] + total[i-1] computed on the device after an
 * event from slice i-1 (one writer per word, ordered by events).  streams[0]
 * waits for every slice at the end: synchronise on it.  The result equals
 * csx_compact's.  Replaces bench3_3pass_streaming + its carry kernel
 * (src/benchmarks.cuh:195-293, src/kernels/kernel_streaming_add.cuh:8-16). */
csx_status csx_streams_workspace_size(uint64_t n, uint32_t elem_size, uint32_t chunk_len, uint32_t stream_count, size_t* bytes);
csx_status csx_compact_streams(const void* d_in, const uint8_t* d_mask, void* d_out, uint64_t n, uint32_t elem_size, uint32_t chunk_len,
                               int bit_order, const hipStream_t* streams, uint32_t stream_count, uint64_t* d_count, uint64_t* h_count, void* d_ws,
                               size_t ws_bytes);

/* ---------------------------------------------------------------- batched columns
 * Many independent columns in one call: d_segments is a DEVICE array of
 * nseg {in, mask, out, n} descriptors; column s is compacted into its own
 * out with d_counts[s] selected elements (device u64 array of nseg).
 * Sizes are read on the device, so the call launches four kernels with no
 * host round trip and can be captured into a graph once and replayed after
 * rewriting the table.  max_total_elements bounds the sum of n (it sizes the
 * grids and the workspace, so keep it tight: the write kernel dispatches one
 * workgroup per tile of the bound, and those past the real tiles only exit);
 * a table exceeding it sets the workspace error word (csx_workspace_error)
 * and leaves the excess uncompacted.  A workspace sized for a larger bound
 * serves any smaller one.  A column's out must not overlap any column's in
 * or mask (the table is in device memory, so this is not checked).  No
 * counterpart in the reference beyond its two-stream split
 * (src/benchmarks.cuh:195-293); MSB- or LSB-first masks, 1/2/4/8-byte
 * elements, all columns of one width. */
typedef struct {
    const void* in;
    const uint8_t* mask;
    void* out;
    uint64_t n;
} csx_segment;
csx_status csx_batch_workspace_size(uint32_t nseg, uint64_t max_total_elements, uint32_t elem_size, size_t* bytes);
csx_status csx_compact_batch(const csx_segment* d_segments, uint32_t nseg, uint64_t max_total_elements, uint32_t elem_size, int bit_order,
                             uint64_t* d_counts, void* d_ws, size_t ws_bytes, hipStream_t stream);

/* ---------------------------------------------------------------- serving
 * Device memory arena: one allocation carved first-fit; the HBM budget of
 * the serving loop.  csx_arena_alloc returns CSX_ERR_WORKSPACE when no free
 * range fits (the caller keeps the request waiting).  Not thread-safe
 * against csx_arena_destroy. */
typedef struct csx_arena csx_arena;
csx_status csx_arena_create(uint64_t bytes, csx_arena** arena);
csx_status csx_arena_destroy(csx_arena* arena);
csx_status csx_arena_alloc(csx_arena* arena, uint64_t bytes, void** d_ptr);
csx_status csx_arena_free(csx_arena* arena, void* d_ptr);
uint64_t csx_arena_used(const csx_arena* arena);

/* Continuous-batching server.  csx_server_submit queues a request (device
 * buffers, MSB-first mask) on the host; csx_server_step launches, while one
 * of `streams` slots is free, one batched compaction (csx_compact_batch) of
 * queued requests of one element width (at most max_batch requests and
 * max_batch_elements elements), so new arrivals run while earlier batches are
 * still on the device; csx_server_poll retires finished batches without
 * blocking and reports, per request, the caller's id, the selected count and
 * the latency from submission to that poll (`capacity` must hold a whole
 * batch: at least max_batch completions).  A step that fails waits for the
 * failed batch's stream and puts its requests back in submission order.  The results
 * equal csx_compact's on each request alone.  One host thread drives a
 * server; csx_server_destroy waits for in-flight batches.  No counterpart in
 * the reference. */
typedef struct csx_server csx_server;
typedef struct {
    uint32_t streams;            /* concurrent batches (1..16) */
    uint32_t max_batch;          /* requests per batch */
    uint64_t max_batch_elements; /* elements per batch (also the largest request) */
} csx_server_config;
typedef struct {
    uint64_t id;
    const void* in;
    const uint8_t* mask;
    void* out;
    uint64_t n;
    uint32_t elem_size;
    hipEvent_t ready; /* nullable: the batch stream waits for it (inputs produced on another stream) */
} csx_request;
typedef struct {
    uint64_t id;
    uint64_t count;
    double latency_ms;
} csx_completion;
csx_status csx_server_create(const csx_server_config* config, csx_server** server);
csx_status csx_server_destroy(csx_server* server);
csx_status csx_server_submit(csx_server* server, const csx_request* request);
csx_status csx_server_step(csx_server* server, uint32_t* launched);
csx_status csx_server_poll(csx_server* server, csx_completion* done, uint32_t capacity, uint32_t* count);
uint32_t csx_server_queued(const csx_server* server);
uint32_t csx_server_inflight(const csx_server* server);

/* ---------------------------------------------------------------- tuning
 * A compaction configuration.  pipeline: CSX_PIPELINE_*; chunk_len: the
 * three-phase chunk length; vpt: tile size in 16-byte vectors per thread of
 * the three-phase writeout (2, 4, 8, or 16 when elem_size >= 2), of the
 * single launch (2, 4 or 8) or of the batched form (2, 4 or 8); TWO_PASS,
 * FUSED and SPARSE take 8, which they do not use; count_blocks: grid bound of the batched form's per-tile
 * count kernel (0: 4096); block_size: the workgroup size of those tiles
 * (128, 256 or 512; 0 means 256; the other pipelines take 0 or 256) -- the
 * reference's block-size argument (src/kernels/kernel_3pass.cuh:491-533).
 * The tuning table maps (elem_size, log2 n, selectivity) to the fastest
 * measured configuration; csx_tune_lookup answers with the measured cell of
 * that width nearest in log2 n, then in selectivity, or with the default
 * (THREE_PHASE, 1024, 8) when none was measured.  No counterpart in the
 * reference, whose block size and chunk length are benchmark arguments
 * (src/main.cu:38-46, src/kernels/kernel_3pass.cuh:491-533). */
typedef struct {
    int pipeline;
    uint32_t chunk_len;
    uint32_t vpt;
    uint32_t count_blocks;
    uint32_t block_size; /* threads per workgroup of the staged writeout tile (THREE_PHASE) or of the batched
                            write kernel: 128, 256 or 512 (0: 256); the tile is block_size * vpt 16-byte vectors */
} csx_config;
/* csx_compact with an explicit configuration; workspace csx_workspace_size(n, elem_size, cfg->chunk_len) */
csx_status csx_compact_config(const void* d_in, const uint8_t* d_mask, void* d_out, uint64_t n, uint32_t elem_size, int bit_order,
                              const csx_config* cfg, uint64_t* d_count, uint64_t* h_count, void* d_ws, size_t ws_bytes, hipStream_t stream);
/* workspace that fits every configuration the tuner may choose */
csx_status csx_auto_workspace_size(uint64_t n, uint32_t elem_size, size_t* bytes);
/* selectivity in [0, 1], or < 0 when unknown (then 0.5) */
csx_status csx_tune_lookup(uint32_t elem_size, uint64_t n, double selectivity, csx_config* cfg);
csx_status csx_compact_auto(const void* d_in, const uint8_t* d_mask, void* d_out, uint64_t n, uint32_t elem_size, int bit_order, double selectivity,
                            uint64_t* d_count, uint64_t* h_count, void* d_ws, size_t ws_bytes, hipStream_t stream);
/* Autotuner for one cell: generates n elements and a Bernoulli(selectivity)
 * mask on the device, times every candidate (THREE_PHASE x chunk {256, 1024,
 * 4096} x vpt {2, 4, 8, 16} in 256-thread tiles, THREE_PHASE at chunk 1024 x
 * vpt in 128- and 512-thread tiles, TWO_PASS, FUSED, SPARSE at chunks 1024
 * and 4096, and SINGLE x vpt {2, 4, 8} up to its cap) with HIP events -- `reps` calls back to back between two events, three
 * rounds after one warm-up, the median round's per-call time -- checks that
 * all give the same count,
 * records the fastest in the table and returns it.  Allocates and frees its
 * own device buffers and synchronises the stream: an offline step. */
csx_status csx_autotune(uint32_t elem_size, uint64_t n, double selectivity, uint32_t reps, csx_config* best, double* best_ms, hipStream_t stream);
/* The same candidates timed on the caller's own column: d_in / d_mask (n
 * elements, bit_order) into d_out (n elements), with the caller's workspace
 * of csx_auto_workspace_size bytes.  The fastest is recorded under
 * (elem_size, n, k/n), k the selected count (returned in *h_count when not
 * null), and csx_compact_auto uses it from then on for that cell.  The table
 * is measured on Bernoulli masks; on masks whose selected elements come in
 * runs (sorted or correlated columns) a larger tile can be faster at the
 * same selectivity, and tuning on such a column records it.  Runs the call
 * 3 reps + 1 times per candidate (d_out and the workspace are rewritten each
 * time) and synchronises the stream: an offline step.  No counterpart in the
 * reference. */
csx_status csx_autotune_inputs(const void* d_in, const uint8_t* d_mask, void* d_out, uint64_t n, uint32_t elem_size, int bit_order, uint32_t reps,
                               void* d_ws, size_t ws_bytes, csx_config* best, double* best_ms, uint64_t* h_count, hipStream_t stream);
/* The batched form tuned on the caller's own table (vpt {2, 4, 8} x
 * count_blocks {1024, 2048, 4096, 8192}); the result is recorded for
 * elem_size and used by csx_compact_batch from then on.  Runs the batch
 * 3 reps + 1 times per candidate (outputs and counts are rewritten each time). */
csx_status csx_autotune_batch(const csx_segment* d_segments, uint32_t nseg, uint64_t max_total_elements, uint32_t elem_size, int bit_order,
                              uint64_t* d_counts, void* d_ws, size_t ws_bytes, uint32_t reps, csx_config* best, double* best_ms,
                              hipStream_t stream);
csx_status csx_compact_batch_config(const csx_segment* d_segments, uint32_t nseg, uint64_t max_total_elements, uint32_t elem_size, int bit_order,
                                    const csx_config* cfg, uint64_t* d_counts, void* d_ws, size_t ws_bytes, hipStream_t stream);
/* csx_compact_config with no workspace argument.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * scratch is a device array inside the library, CSX_NOALLOC_SCRATCH_BYTES on
 * each GPU, whose address is looked up on the first call on that device; it
 * is held until the process ends.  The call allocates nothing.  A request
 * whose workspace (csx_workspace_size(n, elem_size, cfg->chunk_len)) exceeds
 * the scratch returns CSX_ERR_UNSUPPORTED before any device work: 32 MiB
 * covers about 2^31 four-byte elements at chunk 1024.  Every call on one GPU
 * shares that GPU's scratch, so calls that overlap on one GPU (other streams
 * or host threads) race: the caller serialises them. */
#define CSX_NOALLOC_SCRATCH_BYTES (UINT64_C(32) << 20)
csx_status csx_compact_noalloc(const void* d_in, const uint8_t* d_mask, void* d_out, uint64_t n, uint32_t elem_size, int bit_order,
                               const csx_config* cfg, uint64_t* d_count, uint64_t* h_count, hipStream_t stream);
/* the table as text (one measured cell a line); load adds to the table */
csx_status csx_tune_save(const char* path);
csx_status csx_tune_load(const char* path);
csx_status csx_tune_clear(void);

/* ---------------------------------------------------------------- predicate
 * The step before the path: a column and a comparison give the packed mask.
 * csx_predicate_mask replaces gen_predicate (src/utils.cuh:102-122) with the
 * driver's predicate_function `value > threshold` (src/main.cu:24-28), on the
 * device, for any element type and comparison; element i is selected iff
 * `in[i] OP a` (BETWEEN: a <= in[i] <= b), compared in the column's own type
 * (IEEE: every comparison with NaN is false except NE).  Mask bytes past
 * ceil(n/8) are not written; bits at index >= n are zero.  d_count (nullable)
 * receives the number selected.
 * csx_compact_predicate fuses the predicate into the compaction: one kernel
 * evaluates it, writes the mask and the per-chunk counts (the popcount phase),
 * then the look-back scan and the staged writeout run as in csx_compact.
 * d_mask may be null (the mask then lives in the workspace).  No counterpart
 * in the reference beyond the two steps it replaces. */
typedef enum {
    CSX_TYPE_U8 = 0,
    CSX_TYPE_U16 = 1,
    CSX_TYPE_U32 = 2,
    CSX_TYPE_U64 = 3,
    CSX_TYPE_I8 = 4,
    CSX_TYPE_I16 = 5,
    CSX_TYPE_I32 = 6,
    CSX_TYPE_I64 = 7,
    CSX_TYPE_F32 = 8,
    CSX_TYPE_F64 = 9
} csx_type;
typedef enum {
    CSX_PRED_GT = 0,
    CSX_PRED_GE = 1,
    CSX_PRED_LT = 2,
    CSX_PRED_LE = 3,
    CSX_PRED_EQ = 4,
    CSX_PRED_NE = 5,
    CSX_PRED_BETWEEN = 6
} csx_pred_op;
typedef union {
    uint64_t u64;
    int64_t i64;
    double f64;
    float f32;
    uint32_t u32;
    int32_t i32;
    uint16_t u16;
    int16_t i16;
    uint8_t u8;
    int8_t i8;
} csx_scalar; /* the member of the column's type is read */
typedef struct {
    int type; /* csx_type */
    int op;   /* csx_pred_op */
    csx_scalar a, b;
} csx_predicate;
csx_status csx_predicate_mask(const void* d_in, uint64_t n, const csx_predicate* pred, uint8_t* d_mask, int bit_order, uint64_t* d_count,
                              hipStream_t stream);
csx_status csx_predicate_workspace_size(uint64_t n, int type, size_t* bytes);
csx_status csx_compact_predicate(const void* d_in, uint64_t n, const csx_predicate* pred, void* d_out, uint8_t* d_mask, int bit_order,
                                 uint64_t* d_count, uint64_t* h_count, void* d_ws, size_t ws_bytes, hipStream_t stream);

/* Non-zero if a single-pass kernel that used this workspace gave up waiting
 * (never expected; reported as CSX_ERR_DEVICE by synchronous calls).
 * Synchronises the stream. */
csx_status csx_workspace_error(const void* d_ws, uint32_t* h_error, hipStream_t stream);

/* ---------------------------------------------------------------- checks
 * Number of positions i < len where a[i] and b[i] differ bytewise, added
 * to *d_failures (device u64).  Replaces kernel_check_validation / validate
 * (src/kernels/data_generator.cuh:122-161, src/benchmarks.cuh:442-450). */
csx_status csx_validate(const void* d_expected, const void* d_actual, uint64_t len, uint32_t elem_size, uint64_t* d_failures,
                        hipStream_t stream);

/* ---------------------------------------------------------------- pattern
 * Compaction by a periodic pattern, with no mask in memory: element i is
 * selected iff bit 31-(i % pattern_length) of `pattern` is set (the pattern
 * starts at its most significant bit, src/data_generator.cuh:113-128); bits
 * below the top pattern_length are ignored.  pattern_length in [1, 32].
 * The count has a closed form, so *h_count (nullable) is set without a
 * synchronisation; *d_count (nullable) is written on the stream.
 * Replaces launch_pattern_proc<T> (src/kernels/kernel_pattern.cuh:71-131). */
csx_status csx_pattern_count(uint64_t n, uint32_t pattern, uint32_t pattern_length, uint64_t* count);
csx_status csx_pattern_compact(const void* d_in, void* d_out, uint64_t n, uint32_t elem_size, uint32_t pattern, uint32_t pattern_length,
                               uint64_t* d_count, uint64_t* h_count, hipStream_t stream);

/* ---------------------------------------------------------------- generators
 * Device-side seeded inputs.  All draw from the reference's generator
 * (PCG32 XSH-RR, increment 1, seeded by one splitmix64 step,
 * src/fast_prng.cuh:7-36) with element k at fixed draw positions, so the
 * device output equals a sequential CPU generator's bit for bit:
 *   bernoulli: element k selected iff draw k < (uint64)(p * UINT32_MAX) --
 *              with seed 42 and MSB-first this is generate_mask_uniform
 *              (src/data_generator.cuh:14-30) for n a multiple of 8;
 *   pattern:   element k is bit 31-(k % pattern_length) of pattern
 *              (generate_mask_pattern, src/data_generator.cuh:113-128);
 *   zipf blocks: block b of block_len elements (a multiple of 32) uses draw
 *              b*(block_len+1) to pick rank r in [0, levels) with
 *              P(r) ~ (r+1)^-s and selectivity r/(levels-1), then one draw
 *              per element;
 *   runs:      element k flips the running bit iff draw k < 2^32/mean_run
 *              (alternating runs with geometric lengths of mean mean_run);
 *   values:    element k <- draw k (1/2/4-byte, low bits) or draws 2k, 2k+1
 *              (8-byte, high word first); CSX_VALUES_FLOAT gives float32
 *              (draw>>8)*2^-24 or float64 (u64>>11)*2^-53 in [0, 1).
 * Bits at index >= n are zero; bytes past ceil(n/8) are not written. */
#define CSX_VALUES_INTEGER 0
#define CSX_VALUES_FLOAT 1
csx_status csx_gen_mask_bernoulli(uint8_t* d_mask, uint64_t n, double p, uint64_t seed, int bit_order, hipStream_t stream);
csx_status csx_gen_mask_pattern(uint8_t* d_mask, uint64_t n, uint32_t pattern, uint32_t pattern_length, int bit_order, hipStream_t stream);
/* host-side rank table of the zipf-block generator (levels in [2, 128]) */
csx_status csx_zipf_table(double s, uint32_t levels, uint32_t* cdf, uint64_t* thresholds);
csx_status csx_gen_mask_zipf_blocks(uint8_t* d_mask, uint64_t n, uint32_t block_len, double s, uint32_t levels, uint64_t seed, int bit_order,
                                    hipStream_t stream);
csx_status csx_gen_runs_workspace_size(uint64_t n, size_t* bytes);
csx_status csx_gen_mask_runs(uint8_t* d_mask, uint64_t n, uint32_t mean_run, uint64_t seed, int bit_order, void* d_ws, size_t ws_bytes,
                             hipStream_t stream);
csx_status csx_gen_values(void* d_out, uint64_t n, uint32_t elem_size, int kind, uint64_t seed, hipStream_t stream);

/* Plain device copy of `bytes` bytes (16-byte aligned pointers): the
 * bandwidth reference the benchmark drivers report beside the compaction and
 * the known-size kernel that calibrates the FETCH_SIZE/WRITE_SIZE counters.
 * No counterpart in the reference. */
csx_status csx_copy(void* d_dst, const void* d_src, uint64_t bytes, hipStream_t stream);

/* What HBM sustains for plain access mixes in the writeout's access shape:
 * one block of 256 threads per 256 * vectors_per_thread 16-byte vectors (a
 * grid covering the buffer, no loop), non-temporal loads and stores, blocks
 * in runs of xcd_run per XCD (0: plain order).  kind CSX_PROBE_READ reads
 * d_src (d_dst: 16 bytes that are written only on a hash match);
 * CSX_PROBE_WRITE writes d_dst; CSX_PROBE_COPY copies; CSX_PROBE_MIX reads
 * `bytes` of d_src and writes bytes / 2 to d_dst (output vector o = input
 * vectors 2o xor 2o+1): the byte mix of a 50 % compaction.  bytes a multiple
 * of 32, pointers 16-byte aligned, vectors_per_thread 1 or 4.  The roofline
 * context the benchmark reports beside the compaction; no counterpart in the
 * reference. */
#define CSX_PROBE_READ 0
#define CSX_PROBE_WRITE 1
#define CSX_PROBE_COPY 2
#define CSX_PROBE_MIX 3
csx_status csx_stream_probe(int kind, void* d_dst, const void* d_src, uint64_t bytes, uint32_t vectors_per_thread, uint32_t xcd_run,
                            hipStream_t stream);

/* ---------------------------------------------------------------- profiling
 * Kernel timing for benchmark drivers: after csx_profile_begin(), every
 * csx_compact / csx_writeout call on this host thread records a HIP event
 * pair on its stream immediately around its dominant kernel (the fused tile
 * kernel, or the writeout kernel of the three-phase pipeline).
 * csx_profile_end() synchronises those events and returns the summed
 * milliseconds and the number of timed launches.  (The reference times each
 * phase with CUDA_TIME event pairs, src/cuda_time.cuh:11-21.) */
csx_status csx_profile_begin(void);
csx_status csx_profile_end(double* total_ms, uint64_t* launches);
/* The workgroup size and grid (workgroups) of the dominant kernel -- the one
 * csx_profile times -- of the calling thread's last call that reached it:
 * what the call actually launched, for drivers that print the launch shape
 * beside the block and grid arguments they were given (the reference's CSV
 * columns block_size;grid_size, src/main.cu:166).  0 and 0 when that call
 * launched nothing (empty input). */
csx_status csx_last_launch(uint32_t* block_size, uint64_t* grid);

#ifdef __cplusplus
}
#endif

#endif /* CSX_H */
