// baseline_rocprim.hip -- rocPRIM flagged select over the packed mask, the
// vendor baseline the engine is measured against (the analogue of the
// reference's cub::DeviceSelect::Flagged over its bitstream_iterator,
// src/benchmarks.cuh:384-394, src/cub_wraps.cuh:12-50).  Built into its own
// library, libcsx_baseline.so; it is a baseline, not part of the engine.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>
#include <stdint.h>

namespace {

// element i of an MSB-first packed mask (the reference's bit order)
struct MaskBit {
    const uint8_t* mask;
    __host__ __device__ bool operator()(uint64_t i) const { return (mask[i >> 3] >> (7 - (i & 7))) & 1u; }
};

template <typename T>
hipError_t run(void* temp, size_t& temp_bytes, const void* in, const uint8_t* mask, void* out, unsigned long long* count, uint64_t n, hipStream_t s)
{
    auto flags = rocprim::make_transform_iterator(rocprim::make_counting_iterator<uint64_t>(0), MaskBit{mask});
    return rocprim::select(temp, temp_bytes, static_cast<const T*>(in), flags, static_cast<T*>(out), count, (size_t)n, s);
}

hipError_t dispatch(uint32_t w, void* temp, size_t& temp_bytes, const void* in, const uint8_t* mask, void* out, unsigned long long* count, uint64_t n,
                    hipStream_t s)
{
    switch (w) {
        case 1: return run<uint8_t>(temp, temp_bytes, in, mask, out, count, n, s);
        case 2: return run<uint16_t>(temp, temp_bytes, in, mask, out, count, n, s);
        case 4: return run<uint32_t>(temp, temp_bytes, in, mask, out, count, n, s);
        case 8: return run<uint64_t>(temp, temp_bytes, in, mask, out, count, n, s);
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t run_bytes(void* temp, size_t& temp_bytes, const void* in, const uint8_t* flags, void* out, unsigned long long* count, uint64_t n,
                     hipStream_t s)
{
    return rocprim::select(temp, temp_bytes, static_cast<const T*>(in), flags, static_cast<T*>(out), count, (size_t)n, s);
}

hipError_t dispatch_bytes(uint32_t w, void* temp, size_t& temp_bytes, const void* in, const uint8_t* flags, void* out, unsigned long long* count,
                          uint64_t n, hipStream_t s)
{
    switch (w) {
        case 1: return run_bytes<uint8_t>(temp, temp_bytes, in, flags, out, count, n, s);
        case 2: return run_bytes<uint16_t>(temp, temp_bytes, in, flags, out, count, n, s);
        case 4: return run_bytes<uint32_t>(temp, temp_bytes, in, flags, out, count, n, s);
        case 8: return run_bytes<uint64_t>(temp, temp_bytes, in, flags, out, count, n, s);
        default: return hipErrorInvalidValue;
    }
}

} // namespace

extern "C" {

// the same select with one flag byte per element (the reference's unused
// launch_cub_flagged_bytemask, src/cub_wraps.cuh:101-136); the byte flags are
// expanded by the caller, outside any timing
int csxb_rocprim_select_bytes_temp_size(uint32_t w, uint64_t n, size_t* bytes)
{
    size_t b = 0;
    hipError_t e = dispatch_bytes(w, nullptr, b, nullptr, nullptr, nullptr, nullptr, n, 0);
    *bytes = b;
    return (int)e;
}

int csxb_rocprim_select_bytes(uint32_t w, const void* d_in, const uint8_t* d_flags, void* d_out, uint64_t* d_count, uint64_t n, void* d_temp,
                              size_t temp_bytes, hipStream_t stream)
{
    size_t b = temp_bytes;
    return (int)dispatch_bytes(w, d_temp, b, d_in, d_flags, d_out, reinterpret_cast<unsigned long long*>(d_count), n, stream);
}

// bytes of temporary storage rocprim::select needs for n elements of width w
int csxb_rocprim_select_temp_size(uint32_t w, uint64_t n, size_t* bytes)
{
    size_t b = 0;
    hipError_t e = dispatch(w, nullptr, b, nullptr, nullptr, nullptr, nullptr, n, 0);
    *bytes = b;
    return (int)e;
}

// out[0..k) = selected elements (MSB-first mask), *d_count = k (device u64)
int csxb_rocprim_select(uint32_t w, const void* d_in, const uint8_t* d_mask, void* d_out, uint64_t* d_count, uint64_t n, void* d_temp,
                        size_t temp_bytes, hipStream_t stream)
{
    size_t b = temp_bytes;
    return (int)dispatch(w, d_temp, b, d_in, d_mask, d_out, reinterpret_cast<unsigned long long*>(d_count), n, stream);
}

} // extern "C"
