// ref_shim.cpp -- C entry points over the reference's own CPU code.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile (target `ref`)
// against the reference sources where they lie (src/fast_prng.cuh,
// src/data_generator.cuh of RememberOfLife/gpu_compressstore2); nothing of
// the reference is copied here.  hipcc in HIP host-only mode supplies the
// __host__/__device__ qualifiers those headers use.  The resulting
// oracle/_ref/libcsx_ref.so pins the C restatement in csx_oracle.c and serves
// as the "reference" CPU baseline.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "src/data_generator.cuh"

extern "C" {

void ref_prng_seq(uint64_t seed, uint32_t* out, size_t n)
{
    fast_prng rng(seed);
    for (size_t i = 0; i < n; i++) out[i] = rng.rand();
}

void ref_mask_uniform(uint8_t* buf, uint64_t offset, uint64_t length, double selectivity, uint64_t* out_count)
{
    size_t c = 0;
    generate_mask_uniform(buf, offset, length, selectivity, &c);
    if (out_count) *out_count = c;
}

void ref_mask_zipf(uint8_t* buf, uint64_t count, uint64_t offset, uint64_t length, uint64_t* out_count)
{
    size_t c = 0;
    generate_mask_zipf(buf, count, offset, length, &c);
    if (out_count) *out_count = c;
}

void ref_mask_burst(uint8_t* buf, uint64_t count, uint64_t offset, uint64_t length, double segment_sizer)
{
    generate_mask_burst(buf, count, offset, length, segment_sizer);
}

void ref_mask_offset(uint8_t* buf, uint64_t offset, uint64_t length, int64_t spacing)
{
    generate_mask_offset(buf, offset, length, spacing);
}

void ref_mask_pattern(uint8_t* buf, uint64_t offset, uint64_t length, uint32_t pattern, uint32_t pattern_length, uint64_t* out_count)
{
    size_t c = 0;
    generate_mask_pattern(buf, offset, length, pattern, pattern_length, &c);
    if (out_count) *out_count = c;
}

uint64_t ref_mask_clustering(uint8_t* buf, float selectivity, uint64_t cluster_count, uint64_t total_elements, uint64_t* out_count)
{
    size_t c = 0;
    std::vector<uint8_t> m = generate_mask_clustering(selectivity, cluster_count, total_elements, &c);
    if (!m.empty()) std::memcpy(buf, m.data(), m.size());
    if (out_count) *out_count = c;
    return m.size();
}

uint64_t ref_generate_validation(const void* in, uint8_t* mask, void* out, uint64_t n, uint32_t width)
{
    switch (width) {
        case 1: return generate_validation((uint8_t*)in, mask, (uint8_t*)out, n);
        case 2: return generate_validation((uint16_t*)in, mask, (uint16_t*)out, n);
        case 4: return generate_validation((uint32_t*)in, mask, (uint32_t*)out, n);
        case 8: return generate_validation((uint64_t*)in, mask, (uint64_t*)out, n);
        default: return UINT64_MAX;
    }
}

} // extern "C"
