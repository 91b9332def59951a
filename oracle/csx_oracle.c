/*
 * csx_oracle.c -- CPU oracle for bitmask stream compaction.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Never linked
 * into, loaded by or called from the product library.  See csx_oracle.h.
 *
 * Reference citations are to RememberOfLife/gpu_compressstore2 (paths
 * relative to the repository root).
 */
#include "csx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ PRNG */

/* splitmix64 pre-seed, then PCG32 state; src/fast_prng.cuh:11-19 */
void csxo_prng_seed(csxo_prng* r, uint64_t seed)
{
    uint64_t z = seed + UINT64_C(0x9E3779B97F4A7C15);
    z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
    z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
    r->state = z ^ (z >> 31);
    r->inc = 1;
}

/* PCG32 XSH-RR output of the old state, LCG step; src/fast_prng.cuh:21-30 */
uint32_t csxo_prng_next(csxo_prng* r)
{
    uint64_t old = r->state;
    r->state = old * UINT64_C(6364136223846793005) + (r->inc | 1);
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
}

void csxo_prng_seq(uint64_t seed, uint32_t* out, size_t n)
{
    csxo_prng r;
    csxo_prng_seed(&r, seed);
    for (size_t i = 0; i < n; i++) out[i] = csxo_prng_next(&r);
}

/* LCG jump-ahead by square-and-multiply of the affine map x -> a*x + c
 * (F. Brown, "Random Number Generation with Arbitrary Stride", 1994). */
void csxo_prng_advance(csxo_prng* r, uint64_t delta)
{
    uint64_t cur_mult = UINT64_C(6364136223846793005);
    uint64_t cur_plus = r->inc | 1;
    uint64_t acc_mult = 1, acc_plus = 0;
    while (delta) {
        if (delta & 1) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1;
    }
    r->state = acc_mult * r->state + acc_plus;
}

/* -------------------------------------------------------------- bit order */

int csxo_get_bit(const uint8_t* mask, uint64_t i, int bit_order)
{
    unsigned sh = bit_order == CSXO_LSB_FIRST ? (unsigned)(i % 8) : 7u - (unsigned)(i % 8);
    return (mask[i / 8] >> sh) & 1;
}

void csxo_set_bit(uint8_t* mask, uint64_t i, int bit_order)
{
    unsigned sh = bit_order == CSXO_LSB_FIRST ? (unsigned)(i % 8) : 7u - (unsigned)(i % 8);
    mask[i / 8] |= (uint8_t)(1u << sh);
}

/* tail hygiene of src/main.cu:110-116 / src/benchmarks.cuh:111-121 */
void csxo_clear_tail(uint8_t* mask, uint64_t n, int bit_order)
{
    if (n % 8 == 0) return;
    for (uint64_t i = n; i < (n / 8 + 1) * 8; i++) {
        unsigned sh = bit_order == CSXO_LSB_FIRST ? (unsigned)(i % 8) : 7u - (unsigned)(i % 8);
        mask[n / 8] &= (uint8_t)~(1u << sh);
    }
}

/* ------------------------------------------- reference mask generators */

/* src/data_generator.cuh:14-30 (the reference always seeds 42) */
void csxo_mask_uniform(uint8_t* buf, uint64_t offset, uint64_t length, double selectivity, uint64_t seed, uint64_t* out_count)
{
    uint64_t p_adjusted = (uint64_t)(selectivity * UINT32_MAX);
    csxo_prng rng;
    csxo_prng_seed(&rng, seed);
    for (uint64_t i = offset; i < offset + length; i++) {
        uint8_t acc = 0;
        for (int j = 7; j >= 0; j--) {
            if ((uint64_t)csxo_prng_next(&rng) < p_adjusted) {
                acc |= (uint8_t)(1u << j);
                if (out_count) (*out_count)++;
            }
        }
        buf[i] = acc;
    }
}

/* src/data_generator.cuh:35-59 */
void csxo_mask_zipf(uint8_t* buf, uint64_t count, uint64_t offset, uint64_t length, uint64_t* out_count)
{
    csxo_prng rng;
    csxo_prng_seed(&rng, 42);
    uint64_t n = count * 8;
    double a = 1.2;
    double c = log10((double)n) / (double)n;
    double k = 1.43;
    for (uint64_t i = offset; i < offset + length; i++) {
        uint8_t acc = 0;
        for (int j = 7; j >= 0; j--) {
            double ev = a * (1 / (pow((c * (double)(i * 8 + (uint64_t)(7 - j))), k)));
            double rv = (double)csxo_prng_next(&rng) / (double)UINT32_MAX;
            if (rv < ev) {
                acc |= (uint8_t)(1u << j);
                if (out_count) (*out_count)++;
            }
        }
        buf[i] = acc;
    }
}

/* src/data_generator.cuh:65-88 (unsigned `--len <= 0` is `== 0`) */
void csxo_mask_burst(uint8_t* buf, uint64_t count, uint64_t offset, uint64_t length, double segment_sizer)
{
    csxo_prng rng;
    csxo_prng_seed(&rng, 42);
    double segment = (double)count * segment_sizer;
    double rv = (double)csxo_prng_next(&rng) / (double)UINT32_MAX;
    uint64_t current_length = (uint64_t)(segment * (rv + 0.5));
    int is_one = 0;
    for (uint64_t i = offset; i < offset + length; i++) {
        uint8_t acc = 0;
        for (int j = 7; j >= 0; j--) {
            if (is_one) acc |= (uint8_t)(1u << j);
            if (--current_length == 0) {
                rv = (double)csxo_prng_next(&rng) / (double)UINT32_MAX;
                current_length = (uint64_t)(segment * (rv + 0.5));
                is_one = !is_one;
            }
        }
        buf[i] = acc;
    }
}

/* src/data_generator.cuh:93-107 (the modulus is taken in uint64) */
void csxo_mask_offset(uint8_t* buf, uint64_t offset, uint64_t length, int64_t spacing)
{
    int invert = spacing < 0;
    spacing = (spacing == 0) ? 1 : spacing;
    for (uint64_t i = offset; i < offset + length; i++) {
        uint8_t acc = 0;
        for (int j = 7; j >= 0; j--) {
            if ((i * 8 + (uint64_t)(7 - j)) % (uint64_t)spacing == 0) acc |= (uint8_t)(1u << j);
        }
        buf[i] = invert ? (uint8_t)~acc : acc;
    }
}

/* src/data_generator.cuh:113-128 */
void csxo_mask_pattern(uint8_t* buf, uint64_t offset, uint64_t length, uint32_t pattern, uint32_t pattern_length, uint64_t* out_count)
{
    for (uint64_t i = offset; i < offset + length; i++) {
        uint8_t acc = 0;
        for (int j = 7; j >= 0; j--) {
            if ((pattern >> (31 - ((i * 8 + (uint64_t)(7 - j)) % pattern_length))) & 1u) {
                acc |= (uint8_t)(1u << j);
                if (out_count) (*out_count)++;
            }
        }
        buf[i] = acc;
    }
}

/* src/data_generator.cuh:130-167.  `selectivity * total_elements` is a float
 * product in the reference (float times size_t). */
void csxo_mask_clustering(uint8_t* buf, float selectivity, uint64_t cluster_count, uint64_t total_elements, uint64_t* out_count)
{
    uint64_t total_set_one = (uint64_t)(selectivity * (float)total_elements);
    uint64_t cluster_size = total_set_one / cluster_count;
    uint64_t slice = total_elements / cluster_count;
    uint64_t bytes = total_elements / 8;
    memset(buf, 0, bytes);
    for (uint64_t i = 0; i < cluster_count; i++) {
        for (uint64_t k = 0; k < cluster_size; k++) {
            uint64_t bit = k + i * slice;
            if (bit / 8 < bytes) {
                if (!((buf[bit / 8] >> (7 - bit % 8)) & 1) && out_count) (*out_count)++;
                buf[bit / 8] |= (uint8_t)(1u << (7 - bit % 8));
            }
        }
    }
}

/* src/main.cu:54-64: draw distinct positions rand()%L until L*s ones are set */
uint32_t csxo_pattern_from_selectivity(int pattern_length, float selectivity, uint64_t seed)
{
    int ones = (int)((float)pattern_length * selectivity);
    csxo_prng rng;
    csxo_prng_seed(&rng, seed);
    uint32_t pattern = 0;
    while (ones > 0) {
        uint32_t i = csxo_prng_next(&rng) % (uint32_t)pattern_length;
        if (((pattern >> (31 - i)) & 1u) == 0) {
            ones--;
            pattern |= 1u << (31 - i);
        }
    }
    return pattern;
}

/* ----------------------------------------------- seeded workload generators */

static uint64_t bernoulli_threshold(double p)
{
    /* the reference's rule, src/data_generator.cuh:16,21 */
    return (uint64_t)(p * UINT32_MAX);
}

void csxo_mask_bernoulli(uint8_t* buf, uint64_t n, double p, uint64_t seed, int bit_order)
{
    uint64_t thr = bernoulli_threshold(p);
    csxo_prng rng;
    csxo_prng_seed(&rng, seed);
    memset(buf, 0, (n + 7) / 8);
    for (uint64_t k = 0; k < n; k++) {
        if ((uint64_t)csxo_prng_next(&rng) < thr) csxo_set_bit(buf, k, bit_order);
    }
}

void csxo_mask_geometric_runs(uint8_t* buf, uint64_t n, uint32_t mean_run, uint64_t seed, int bit_order)
{
    uint32_t thr = (uint32_t)((UINT64_C(1) << 32) / mean_run);
    csxo_prng rng;
    csxo_prng_seed(&rng, seed);
    memset(buf, 0, (n + 7) / 8);
    int v = 0;
    for (uint64_t k = 0; k < n; k++) {
        if (csxo_prng_next(&rng) < thr) v ^= 1;
        if (v) csxo_set_bit(buf, k, bit_order);
    }
}

void csxo_zipf_thresholds(double s, uint32_t levels, uint32_t* cdf_u32)
{
    double norm = 0.0;
    for (uint32_t r = 1; r <= levels; r++) norm += pow((double)r, -s);
    double acc = 0.0;
    for (uint32_t r = 1; r <= levels; r++) {
        acc += pow((double)r, -s) / norm;
        double t = acc * 4294967296.0;
        cdf_u32[r - 1] = (r == levels || t >= 4294967295.0) ? UINT32_MAX : (uint32_t)t;
    }
}

void csxo_mask_zipf_blocks(uint8_t* buf, uint64_t n, uint32_t block_len, double s, uint32_t levels, uint64_t seed, int bit_order)
{
    uint32_t* cdf = (uint32_t*)malloc(sizeof(uint32_t) * levels);
    csxo_zipf_thresholds(s, levels, cdf);
    csxo_prng rng;
    csxo_prng_seed(&rng, seed);
    memset(buf, 0, (n + 7) / 8);
    for (uint64_t base = 0; base < n; base += block_len) {
        uint32_t u = csxo_prng_next(&rng);
        uint32_t r = 0;
        while (r + 1 < levels && u >= cdf[r]) r++;
        double sel = (double)r / (double)(levels - 1);
        uint64_t thr = bernoulli_threshold(sel);
        for (uint64_t k = base; k < base + block_len; k++) {
            uint32_t d = csxo_prng_next(&rng);
            if (k < n && (uint64_t)d < thr) csxo_set_bit(buf, k, bit_order);
        }
    }
    free(cdf);
}

void csxo_mask_periodic(uint8_t* buf, uint64_t n, uint32_t pattern, int bit_order)
{
    memset(buf, 0, (n + 7) / 8);
    for (uint64_t k = 0; k < n; k++) {
        if ((pattern >> (31 - (k % 32))) & 1u) csxo_set_bit(buf, k, bit_order);
    }
}

void csxo_fill_values(void* buf, uint64_t n, uint32_t width, int kind, uint64_t seed)
{
    csxo_prng rng;
    csxo_prng_seed(&rng, seed);
    for (uint64_t k = 0; k < n; k++) {
        if (width == 8) {
            uint64_t hi = csxo_prng_next(&rng);
            uint64_t lo = csxo_prng_next(&rng);
            uint64_t v = (hi << 32) | lo;
            if (kind == 1) {
                double d = (double)(v >> 11) * (1.0 / 9007199254740992.0);
                memcpy((uint8_t*)buf + k * 8, &d, 8);
            }
            else {
                memcpy((uint8_t*)buf + k * 8, &v, 8);
            }
        }
        else {
            uint32_t v = csxo_prng_next(&rng);
            if (width == 4 && kind == 1) {
                float f = (float)(v >> 8) * (1.0f / 16777216.0f);
                memcpy((uint8_t*)buf + k * 4, &f, 4);
            }
            else {
                memcpy((uint8_t*)buf + k * width, &v, width); /* little-endian low bytes */
            }
        }
    }
}

/* src/utils.cuh:102-122 with predicate `value > threshold` (src/main.cu:24-28) */
uint64_t csxo_predicate_mask_f32(const float* col, uint64_t n, float threshold, uint8_t* mask, int bit_order)
{
    uint64_t ones = 0;
    memset(mask, 0, (n + 7) / 8);
    for (uint64_t i = 0; i < n; i++) {
        if (col[i] > threshold) {
            csxo_set_bit(mask, i, bit_order);
            ones++;
        }
    }
    return ones;
}

uint64_t csxo_predicate_mask_u32(const uint32_t* col, uint64_t n, uint32_t threshold, uint8_t* mask, int bit_order)
{
    uint64_t ones = 0;
    memset(mask, 0, (n + 7) / 8);
    for (uint64_t i = 0; i < n; i++) {
        if (col[i] > threshold) {
            csxo_set_bit(mask, i, bit_order);
            ones++;
        }
    }
    return ones;
}

/* Any column type and comparison (the device predicate of csx_predicate_mask):
 * element i is selected iff col[i] OP a, or a <= col[i] <= b for BETWEEN, with
 * C's own comparison in the column's type.  type: 0..3 u8..u64, 4..7 i8..i64,
 * 8 f32, 9 f64; op: 0 >, 1 >=, 2 <, 3 <=, 4 ==, 5 !=, 6 between.  a and b are
 * the little-endian low bytes of the bounds.  Returns the number selected, or
 * UINT64_MAX for an unknown type or op. */
#define CSXO_PRED_TYPED(T)                                                      \
    do {                                                                        \
        T a, b;                                                                 \
        memcpy(&a, &a_bits, sizeof(T));                                         \
        memcpy(&b, &b_bits, sizeof(T));                                         \
        const T* c = (const T*)col;                                             \
        for (uint64_t i = 0; i < n; i++) {                                      \
            int sel;                                                            \
            switch (op) {                                                       \
                case 0: sel = c[i] > a; break;                                  \
                case 1: sel = c[i] >= a; break;                                 \
                case 2: sel = c[i] < a; break;                                  \
                case 3: sel = c[i] <= a; break;                                 \
                case 4: sel = c[i] == a; break;                                 \
                case 5: sel = c[i] != a; break;                                 \
                default: sel = a <= c[i] && c[i] <= b; break;                   \
            }                                                                   \
            if (sel) {                                                          \
                csxo_set_bit(mask, i, bit_order);                               \
                ones++;                                                         \
            }                                                                   \
        }                                                                       \
    } while (0)

uint64_t csxo_predicate_mask_typed(const void* col, uint64_t n, int type, int op, uint64_t a_bits, uint64_t b_bits, uint8_t* mask, int bit_order)
{
    uint64_t ones = 0;
    if (op < 0 || op > 6) return UINT64_MAX;
    memset(mask, 0, (n + 7) / 8);
    switch (type) {
        case 0: CSXO_PRED_TYPED(uint8_t); break;
        case 1: CSXO_PRED_TYPED(uint16_t); break;
        case 2: CSXO_PRED_TYPED(uint32_t); break;
        case 3: CSXO_PRED_TYPED(uint64_t); break;
        case 4: CSXO_PRED_TYPED(int8_t); break;
        case 5: CSXO_PRED_TYPED(int16_t); break;
        case 6: CSXO_PRED_TYPED(int32_t); break;
        case 7: CSXO_PRED_TYPED(int64_t); break;
        case 8: CSXO_PRED_TYPED(float); break;
        case 9: CSXO_PRED_TYPED(double); break;
        default: return UINT64_MAX;
    }
    return ones;
}

/* -------------------------------------------------------- compaction path */

/* src/data_generator.cuh:171-188, including its u32 count and its walk over
 * n/8 whole bytes (elements of a trailing partial byte are not selected). */
uint64_t csxo_generate_validation(const void* in, const uint8_t* mask, void* out, uint64_t n, uint32_t width)
{
    const uint8_t* src = (const uint8_t*)in;
    uint8_t* dst = (uint8_t*)out;
    uint32_t onecount = 0;
    uint64_t val_idx = 0;
    for (uint64_t i = 0; i < n / 8; i++) {
        uint32_t acc = mask[i];
        for (int j = 7; j >= 0; j--) {
            uint64_t idx = i * 8 + (uint64_t)(7 - j);
            if (1u & (acc >> j)) {
                onecount++;
                memcpy(dst + val_idx * width, src + idx * width, width);
                val_idx++;
            }
        }
    }
    memset(dst + val_idx * width, 0, (n - val_idx) * width);
    return onecount;
}

uint64_t csxo_compact(const void* in, const uint8_t* mask, void* out, uint64_t n, uint32_t width, int bit_order)
{
    const uint8_t* src = (const uint8_t*)in;
    uint8_t* dst = (uint8_t*)out;
    uint64_t k = 0;
    for (uint64_t i = 0; i < n; i++) {
        if (csxo_get_bit(mask, i, bit_order)) {
            memcpy(dst + k * width, src + i * width, width);
            k++;
        }
    }
    return k;
}

void csxo_popc(const uint8_t* mask, uint32_t* counts, uint64_t n, uint32_t chunk_len, int bit_order)
{
    uint64_t chunks = (n + chunk_len - 1) / chunk_len;
    for (uint64_t c = 0; c < chunks; c++) {
        uint64_t lo = c * chunk_len, hi = lo + chunk_len < n ? lo + chunk_len : n;
        uint32_t acc = 0;
        for (uint64_t i = lo; i < hi; i++) acc += (uint32_t)csxo_get_bit(mask, i, bit_order);
        counts[c] = acc;
    }
}

uint64_t csxo_exclusive_scan(const uint32_t* counts, uint64_t* offsets, uint64_t chunks)
{
    uint64_t acc = 0;
    for (uint64_t c = 0; c < chunks; c++) {
        offsets[c] = acc;
        acc += counts[c];
    }
    return acc;
}

/* Level d (stride 2^d) adds pss[cid] into pss[cid+stride] for
 * cid = 2*t*stride + stride - 1 while cid + stride < chunks;
 * src/kernels/kernel_3pass.cuh:89-105, levels per :119-121,137-145. */
void csxo_upsweep_tree(uint32_t* pss, uint64_t chunks)
{
    uint32_t max_depth = 0;
    for (uint64_t p2 = 1; p2 < chunks; max_depth++) p2 *= 2;
    for (uint32_t d = 0; d < max_depth; d++) {
        uint64_t stride = UINT64_C(1) << d;
        for (uint64_t cid = stride - 1; cid < chunks; cid += 2 * stride) {
            if (cid + stride < chunks) pss[cid + stride] += pss[cid];
        }
    }
}

/* d_3pass_pproc_pssidx, src/kernels/kernel_3pass.cuh:155-169 */
uint32_t csxo_tree_prefix(const uint32_t* tree, uint64_t idx, uint64_t chunks)
{
    uint64_t p2 = 1;
    while (p2 < chunks) p2 *= 2;
    p2 /= 2;
    uint64_t consumed = 0;
    uint32_t acc = 0;
    while (p2 >= 1) {
        if (idx >= consumed + p2) {
            acc += tree[consumed + p2 - 1];
            consumed += p2;
        }
        p2 /= 2;
    }
    return acc;
}

uint64_t csxo_mask_count(const uint8_t* mask, uint64_t n, int bit_order)
{
    uint64_t k = 0;
    for (uint64_t i = 0; i < n; i++) k += (uint64_t)csxo_get_bit(mask, i, bit_order);
    return k;
}
