/*
 * csx_oracle.h -- CPU oracle for bitmask stream compaction.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may use it, and only as the checker (or, for the CPU
 * baseline, as the thing timed beside the GPU).
 *
 * Plain C restatement of the reference's CPU-side algorithm for the
 * compaction path of RememberOfLife/gpu_compressstore2.  Each function cites
 * the reference file:line it follows.  Bit order:
 *   CSXO_MSB_FIRST (0): element i is bit 7-(i%8) of byte i/8 -- the
 *                       reference's order (src/cub_wraps.cuh:20-24,
 *                       src/data_generator.cuh:177-180).
 *   CSXO_LSB_FIRST (1): element i is bit i%8 of byte i/8.
 */
#ifndef CSX_ORACLE_H
#define CSX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { CSXO_MSB_FIRST = 0, CSXO_LSB_FIRST = 1 };

/* ---- PRNG: PCG32 XSH-RR with inc=1, seeded by one splitmix64 step ---- */
typedef struct {
    uint64_t state;
    uint64_t inc;
} csxo_prng;

void csxo_prng_seed(csxo_prng* r, uint64_t seed);
uint32_t csxo_prng_next(csxo_prng* r);
/* n outputs of a freshly seeded generator */
void csxo_prng_seq(uint64_t seed, uint32_t* out, size_t n);
/* advance the generator by `delta` draws (LCG jump-ahead, O(log delta)) */
void csxo_prng_advance(csxo_prng* r, uint64_t delta);

/* ---- mask bit helpers ---- */
int csxo_get_bit(const uint8_t* mask, uint64_t i, int bit_order);
void csxo_set_bit(uint8_t* mask, uint64_t i, int bit_order);
/* zero every mask bit at index >= n inside the last byte */
void csxo_clear_tail(uint8_t* mask, uint64_t n, int bit_order);

/* ---- reference mask generators (src/data_generator.cuh), MSB-first ---- */
void csxo_mask_uniform(uint8_t* buf, uint64_t offset, uint64_t length, double selectivity, uint64_t seed, uint64_t* out_count);
void csxo_mask_zipf(uint8_t* buf, uint64_t count, uint64_t offset, uint64_t length, uint64_t* out_count);
void csxo_mask_burst(uint8_t* buf, uint64_t count, uint64_t offset, uint64_t length, double segment_sizer);
void csxo_mask_offset(uint8_t* buf, uint64_t offset, uint64_t length, int64_t spacing);
void csxo_mask_pattern(uint8_t* buf, uint64_t offset, uint64_t length, uint32_t pattern, uint32_t pattern_length, uint64_t* out_count);
/* writes total_elements/8 bytes */
void csxo_mask_clustering(uint8_t* buf, float selectivity, uint64_t cluster_count, uint64_t total_elements, uint64_t* out_count);
/* pattern from selectivity and length, src/main.cu:54-64 */
uint32_t csxo_pattern_from_selectivity(int pattern_length, float selectivity, uint64_t seed);

/* ---- seeded workload generators (element k consumes draws in order) ---- */
/* Bernoulli bits: element k selected iff draw k < (uint64)(p*UINT32_MAX). */
void csxo_mask_bernoulli(uint8_t* buf, uint64_t n, double p, uint64_t seed, int bit_order);
/* Alternating runs with geometric lengths: bit k flips w.p. 1/mean_run
 * (draw k < floor(2^32/mean_run)); the first element is 0 before any flip. */
void csxo_mask_geometric_runs(uint8_t* buf, uint64_t n, uint32_t mean_run, uint64_t seed, int bit_order);
/* Zipf-skewed blocks: block b (block_len elements) uses draws
 * [b*(block_len+1), (b+1)*(block_len+1)): the first picks rank r in 1..levels
 * with P(r) ~ r^-s through a 32-bit CDF table, selectivity = (r-1)/(levels-1);
 * then one Bernoulli draw per element. */
void csxo_zipf_thresholds(double s, uint32_t levels, uint32_t* cdf_u32 /* levels entries */);
void csxo_mask_zipf_blocks(uint8_t* buf, uint64_t n, uint32_t block_len, double s, uint32_t levels, uint64_t seed, int bit_order);
/* Periodic: element k is bit 31-(k%32) of `pattern`. */
void csxo_mask_periodic(uint8_t* buf, uint64_t n, uint32_t pattern, int bit_order);
/* Values: width 1/2/4 -> low bits of draw k; width 8 -> (draw 2k << 32) | draw 2k+1.
 * kind 0 = integer, kind 1 = float in [0,1): f32 (draw>>8)*2^-24, f64 (u64>>11)*2^-53. */
void csxo_fill_values(void* buf, uint64_t n, uint32_t width, int kind, uint64_t seed);
/* predicate -> mask (src/utils.cuh:102-122): bit set iff value > threshold */
uint64_t csxo_predicate_mask_f32(const float* col, uint64_t n, float threshold, uint8_t* mask, int bit_order);
uint64_t csxo_predicate_mask_u32(const uint32_t* col, uint64_t n, uint32_t threshold, uint8_t* mask, int bit_order);
/* any type (0..3 u8..u64, 4..7 i8..i64, 8 f32, 9 f64) and comparison (0 >, 1 >=, 2 <, 3 <=, 4 ==, 5 !=, 6 a<=v<=b) */
uint64_t csxo_predicate_mask_typed(const void* col, uint64_t n, int type, int op, uint64_t a_bits, uint64_t b_bits, uint8_t* mask, int bit_order);

/* ---- the compaction path ---- */
/* Exact restatement of generate_validation<T> (src/data_generator.cuh:171-188):
 * walks n/8 whole bytes only, zero-fills out[count..n).  Returns count. */
uint64_t csxo_generate_validation(const void* in, const uint8_t* mask, void* out, uint64_t n, uint32_t width);
/* Stable compaction over all i < n (the GPU kernels' semantics).  Writes only
 * out[0..count).  Returns count. */
uint64_t csxo_compact(const void* in, const uint8_t* mask, void* out, uint64_t n, uint32_t width, int bit_order);
/* counts[c] = popcount of bits [c*C, min((c+1)*C, n)) (src/kernels/kernel_3pass.cuh:12-38) */
void csxo_popc(const uint8_t* mask, uint32_t* counts, uint64_t n, uint32_t chunk_len, int bit_order);
/* exclusive prefix sum u32 -> u64, returns total */
uint64_t csxo_exclusive_scan(const uint32_t* counts, uint64_t* offsets, uint64_t chunks);
/* in-place up-sweep tree of src/kernels/kernel_3pass.cuh:69-153 */
void csxo_upsweep_tree(uint32_t* pss, uint64_t chunks);
/* exclusive prefix of chunk idx from the up-swept tree (kernel_3pass.cuh:155-169) */
uint32_t csxo_tree_prefix(const uint32_t* tree, uint64_t idx, uint64_t chunks);
/* count of selected elements with index < n */
uint64_t csxo_mask_count(const uint8_t* mask, uint64_t n, int bit_order);

#ifdef __cplusplus
}
#endif

#endif
