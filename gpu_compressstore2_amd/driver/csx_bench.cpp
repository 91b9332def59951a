// csx_bench -- the benchmark driver of the engine, a drop-in for the
// reference's src/main.cu: the same flags (getopt ":zuer8d:l:i:f:p:s:t:g:m:c:k:b:n:",
// src/main.cu:197), the same inputs (seed-42 PRNG data and masks), the same
// sweep and the same CSV columns (src/main.cu:166) followed by
// ;bytes;gbps;pct_peak;launch_block;launch_grid -- the workgroup size and grid
// of the dominant kernel that actually ran (csx_last_launch): -b/-n set the
// tile workgroup size of the LDS-staged writeout (bench10 and csx_three_phase;
// nearest of 128, 256, 512), -g/-m are accepted and change nothing, the tile
// kernels launching one workgroup per tile.  Every run is validated against a CPU compaction of
// the same inputs, as the reference does (src/main.cu:173-182).
//
// Extra flags: -w <1|2|4|8> element width (default 4; -8 keeps its meaning),
//              -x <list> comma-separated bench names to run (default all).
#include <csx.h>
#include <csx_compat.hpp>

#include <getopt.h>
#include <unistd.h>

#include <algorithm>
#include <bitset>
#include <cmath>
#include <cstring>
#include <fstream>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

extern "C" {
int csxb_rocprim_select_temp_size(uint32_t w, uint64_t n, size_t* bytes);
int csxb_rocprim_select(uint32_t w, const void* d_in, const uint8_t* d_mask, void* d_out, uint64_t* d_count, uint64_t n, void* d_temp,
                        size_t temp_bytes, hipStream_t stream);
}

using namespace csx_compat;

namespace {

constexpr double kHbmPeakGBs = 8000.0;

// PCG32 XSH-RR seeded through splitmix64 (the generator of src/fast_prng.cuh)
struct Pcg32 {
    uint64_t state;
    explicit Pcg32(uint64_t seed)
    {
        uint64_t z = seed + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        state = z ^ (z >> 31);
    }
    uint32_t next()
    {
        const uint64_t old = state;
        state = old * 6364136223846793005ull + 1ull;
        const uint32_t x = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (x >> rot) | (x << ((-rot) & 31));
    }
};

struct Opts {
    int lines = 1000;
    const char* csv_path = "../res/Arade_1.csv";
    int iterations = 100;
    bool report_failures = false;
    int chunk_length_max = 8192;
    int chunk_length_min = 32;
    int grid_size_max = 0;
    int grid_size_min = 0;
    int block_size_max = 256;
    int block_size_min = 256;
    bool use_csv = false;
    bool use_pattern_mask = true;
    int pattern_length = 8;
    uint32_t pattern = 0;
    float selectivity = 0.5f;
    bool use_uniform = false;
    bool use_zipf = false;
    bool use_clustering = false;
    int width = 4;
    float threshold = 200.f;
    std::string only;
};

[[noreturn]] void error(const char* msg)
{
    fputs(msg, stderr);
    fputs("\n", stderr);
    exit(EXIT_FAILURE);
}

bool pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// ---- host-side mask generators with sequential state (src/data_generator.cuh:35-59, 130-167)
void mask_zipf(std::vector<uint8_t>& buf)
{
    Pcg32 rng(42);
    const uint64_t count = buf.size();
    const uint64_t n = count * 8;
    const double a = 1.2, k = 1.43;
    const double c = log10((double)n) / (double)n;
    for (uint64_t i = 0; i < count; i++) {
        uint8_t acc = 0;
        for (int j = 7; j >= 0; j--) {
            const double ev = a * (1 / (pow((c * (double)(i * 8 + (uint64_t)(7 - j))), k)));
            const double rv = (double)rng.next() / (double)UINT32_MAX;
            if (rv < ev) acc |= (uint8_t)(1u << j);
        }
        buf[i] = acc;
    }
}

std::vector<uint8_t> mask_clustering(float selectivity, uint64_t cluster_count, uint64_t total_elements)
{
    std::vector<uint8_t> m(total_elements / 8, 0);
    const uint64_t total_set_one = (uint64_t)(selectivity * (float)total_elements);
    const uint64_t cluster_size = total_set_one / cluster_count;
    const uint64_t slice = total_elements / cluster_count;
    for (uint64_t i = 0; i < cluster_count; i++)
        for (uint64_t k = 0; k < cluster_size; k++) {
            const uint64_t bit = k + i * slice;
            if (bit / 8 < m.size()) m[bit / 8] |= (uint8_t)(1u << (7 - bit % 8));
        }
    return m;
}

// src/main.cu:54-64
uint32_t pattern_from_selectivity(int length, float selectivity)
{
    int ones = (int)((float)length * selectivity);
    Pcg32 rng(42);
    uint32_t pattern = 0;
    while (ones > 0) {
        const uint32_t i = rng.next() % (uint32_t)length;
        if (((pattern >> (31 - i)) & 1u) == 0) {
            ones--;
            pattern |= 1u << (31 - i);
        }
    }
    return pattern;
}

// column 3 of a '|'-separated file (src/csv_loader.hpp:22-41)
template <typename T> std::vector<T> load_csv_column(const char* path, size_t column, size_t max_lines)
{
    std::ifstream f(path);
    if (!f) error("cannot open the csv file");
    std::vector<T> col;
    std::string line;
    while (std::getline(f, line) && (max_lines == 0 || col.size() < max_lines)) {
        std::stringstream ss(line);
        std::string cell;
        for (size_t c = 0; std::getline(ss, cell, '|'); c++) {
            if (c == column) {
                col.push_back((T)strtod(cell.c_str(), nullptr));
                break;
            }
        }
    }
    return col;
}

template <typename T> uint64_t cpu_compact(const std::vector<T>& col, const std::vector<uint8_t>& mask, std::vector<T>& out)
{
    uint64_t k = 0;
    out.assign(col.size(), T(0));
    for (size_t i = 0; i < col.size(); i++)
        if ((mask[i / 8] >> (7 - i % 8)) & 1u) out[k++] = col[i];
    return k;
}

template <typename T> void run(Opts o)
{
    const size_t w = sizeof(T);
    std::vector<T> col;
    if (o.use_pattern_mask) {
        o.pattern = pattern_from_selectivity(o.pattern_length, o.selectivity);
        std::string bits = std::bitset<32>(o.pattern).to_string();
        fprintf(stderr, "pattern: %s\n", bits.substr(0, o.pattern_length).c_str());
    }
    T* d_input = nullptr;
    if (!o.use_csv) {
        fprintf(stderr, "generating %i lines of input\n", o.lines);
        col.resize(o.lines);
        // the column's bytes are the seed-42 uniform mask at p = 0.5 (src/main.cu:76)
        hip_check(hipMalloc(&d_input, std::max<size_t>(col.size(), 1) * w), "hipMalloc");
        check(csx_gen_mask_bernoulli(reinterpret_cast<uint8_t*>(d_input), col.size() * w * 8, 0.5, 42, CSX_MSB_FIRST, 0), "generate data");
        hip_check(hipMemcpy(col.data(), d_input, col.size() * w, hipMemcpyDeviceToHost), "hipMemcpy");
    }
    else {
        fprintf(stderr, "parsing %s\n", o.csv_path);
        col = load_csv_column<T>(o.csv_path, 3, (size_t)std::max(o.lines, 0));
        hip_check(hipMalloc(&d_input, std::max<size_t>(col.size(), 1) * w), "hipMalloc");
        hip_check(hipMemcpy(d_input, col.data(), col.size() * w, hipMemcpyHostToDevice), "hipMemcpy");
    }
    const size_t n = col.size();
    T* d_output = nullptr;
    hip_check(hipMalloc(&d_output, (n + 1) * w), "hipMalloc");

    // mask
    const size_t mask_bytes = (n + 7) / 8;
    std::vector<uint8_t> pred(std::max<size_t>(mask_bytes, 1), 0);
    uint8_t* d_mask = nullptr;
    hip_check(hipMalloc(&d_mask, pred.size()), "hipMalloc");
    if (o.use_csv) {
        for (size_t i = 0; i < n; i++)
            if ((float)col[i] > o.threshold) pred[i / 8] |= (uint8_t)(1u << (7 - i % 8));
    }
    else if (o.use_uniform) {
        check(csx_gen_mask_bernoulli(d_mask, mask_bytes * 8, o.selectivity, 42, CSX_MSB_FIRST, 0), "generate mask");
        hip_check(hipMemcpy(pred.data(), d_mask, mask_bytes, hipMemcpyDeviceToHost), "hipMemcpy");
    }
    else if (o.use_zipf) {
        pred.resize(mask_bytes);
        mask_zipf(pred);
    }
    else if (o.use_clustering) {
        pred = mask_clustering(o.selectivity, 1, (n + 7) / 8 * 8);
    }
    else {
        check(csx_gen_mask_pattern(d_mask, mask_bytes * 8, o.pattern, o.pattern_length, CSX_MSB_FIRST, 0), "generate mask");
        hip_check(hipMemcpy(pred.data(), d_mask, mask_bytes, hipMemcpyDeviceToHost), "hipMemcpy");
    }
    pred.resize(std::max<size_t>(mask_bytes, 1));
    if (n % 8) pred[n / 8] &= (uint8_t)(0xFF << (8 - n % 8));  // tail hygiene (src/main.cu:110-116)
    hip_check(hipMemcpy(d_mask, pred.data(), pred.size(), hipMemcpyHostToDevice), "hipMemcpy");

    std::vector<T> validation;
    const uint64_t out_length = cpu_compact(col, pred, validation);
    fprintf(stderr, "line count: %zu, one count: %llu, percentage: %f\n", n, (unsigned long long)out_length, n ? (double)out_length / n : 0.0);
    T* d_validation = nullptr;
    hip_check(hipMalloc(&d_validation, std::max<size_t>(n, 1) * w), "hipMalloc");
    hip_check(hipMemcpy(d_validation, validation.data(), n * w, hipMemcpyHostToDevice), "hipMemcpy");

    fprintf(stderr, "starting benchmark\n");
    intermediate_data id{n, o.chunk_length_min, 8, (T*)nullptr};
    size_t temp_bytes = 0;
    if (csxb_rocprim_select_temp_size((uint32_t)w, n, &temp_bytes) != 0) error("rocprim temp size query failed");
    void* d_temp = nullptr;
    hip_check(hipMalloc(&d_temp, std::max<size_t>(temp_bytes, 16)), "hipMalloc");

    using Fn = std::function<timings(int, int, int)>;
    std::vector<std::pair<std::string, Fn>> benchs;
    benchs.emplace_back("bench1_base_variant", [&](int cs, int bs, int gs) { return bench1_base_variant(&id, d_input, d_mask, d_output, n, cs, bs, gs); });
    benchs.emplace_back("bench2_base_variant_skipping",
                        [&](int cs, int bs, int gs) { return bench2_base_variant_skipping(&id, d_input, d_mask, d_output, n, cs, bs, gs); });
    benchs.emplace_back("bench3_3pass_streaming",
                        [&](int, int bs, int gs) { return bench3_3pass_streaming(&id, d_input, d_mask, d_output, n, 1024, bs, gs); });
    benchs.emplace_back("bench4_3pass_optimized_read_non_skipping_cub_pss", [&](int cs, int bs, int gs) {
        return bench4_3pass_optimized_read_non_skipping_cub_pss(&id, d_input, d_mask, d_output, n, cs, bs, gs);
    });
    benchs.emplace_back("bench5_3pass_optimized_read_skipping_partial_pss", [&](int cs, int bs, int gs) {
        return bench5_3pass_optimized_read_skipping_partial_pss(&id, d_input, d_mask, d_output, n, cs, bs, gs);
    });
    benchs.emplace_back("bench6_3pass_optimized_read_skipping_two_phase_pss", [&](int cs, int bs, int gs) {
        return bench6_3pass_optimized_read_skipping_two_phase_pss(&id, d_input, d_mask, d_output, n, cs, bs, gs);
    });
    benchs.emplace_back("bench7_3pass_optimized_read_skipping_cub_pss", [&](int cs, int bs, int gs) {
        return bench7_3pass_optimized_read_skipping_cub_pss(&id, d_input, d_mask, d_output, n, cs, bs, gs);
    });
    benchs.emplace_back("bench8_rocprim_flagged", [&](int, int, int) {
        id.prepare_buffers(n, 0, d_output, d_mask);
        timings t{};
        t.total = time_phase(id.start, id.stop, [&] {
            if (csxb_rocprim_select((uint32_t)w, d_input, d_mask, d_output, id.d_out_count, n, d_temp, temp_bytes, 0) != 0)
                error("rocprim select failed");
        });
        return t;
    });
    if (o.use_pattern_mask) {
        benchs.emplace_back("bench9_pattern", [&](int cs, int bs, int gs) {
            return bench9_pattern(&id, d_input, o.pattern, o.pattern_length, d_output, n, cs, bs, gs);
        });
    }
    benchs.emplace_back("bench10_3pass_optimized_read_skipping_optimized_writeout_cub_pss", [&](int cs, int bs, int gs) {
        return bench10_3pass_optimized_read_skipping_optimized_writeout_cub_pss(&id, d_input, d_mask, d_output, n, cs, bs, gs);
    });
    benchs.emplace_back("csx_three_phase",
                        [&](int cs, int bs, int) { return bench_csx_pipeline(&id, CSX_PIPELINE_THREE_PHASE, d_input, d_mask, d_output, n, cs, bs); });
    benchs.emplace_back("csx_two_pass", [&](int cs, int bs, int) { return bench_csx_pipeline(&id, CSX_PIPELINE_TWO_PASS, d_input, d_mask, d_output, n, cs, bs); });
    benchs.emplace_back("csx_fused", [&](int cs, int bs, int) { return bench_csx_pipeline(&id, CSX_PIPELINE_FUSED, d_input, d_mask, d_output, n, cs, bs); });
    benchs.emplace_back("csx_sparse", [&](int cs, int bs, int) { return bench_csx_pipeline(&id, CSX_PIPELINE_SPARSE, d_input, d_mask, d_output, n, cs, bs); });
    if (!o.only.empty()) {
        std::vector<std::pair<std::string, Fn>> keep;
        std::stringstream ss(o.only);
        std::string name;
        while (std::getline(ss, name, ','))
            for (auto& b : benchs)
                if (b.first == name) keep.push_back(b);
        benchs = keep;
    }

    const double bytes = (double)n * w + (double)mask_bytes + (double)out_length * w;
    std::cout << "benchmark;chunk_length;block_size;grid_size;time_popc;time_pss1;time_pss2;time_proc;time_total;bytes;gbps;pct_peak;launch_block;"
                 "launch_grid"
              << std::endl;
    int failures_total = 0;
    for (int grid_size = o.grid_size_min; grid_size <= o.grid_size_max; grid_size = grid_size ? grid_size * 2 : o.grid_size_max + 1) {
        for (int block_size = o.block_size_min; block_size <= o.block_size_max; block_size *= 2) {
            for (int chunk_length = o.chunk_length_min; chunk_length <= o.chunk_length_max; chunk_length *= 2) {
                std::vector<timings> ts(benchs.size());
                std::vector<std::pair<uint32_t, uint64_t>> shapes(benchs.size());
                for (int it = 0; it < o.iterations; it++) {
                    for (size_t i = 0; i < benchs.size(); i++) {
                        ts[i] += benchs[i].second(chunk_length, block_size, grid_size);
                        // rocPRIM's select is no kernel of the engine: 0;0
                        if (benchs[i].first == "bench8_rocprim_flagged")
                            shapes[i] = {0, 0};
                        else
                            check(csx_last_launch(&shapes[i].first, &shapes[i].second), "csx_last_launch");
                        uint64_t fc = 0;
                        uint64_t count = 0;
                        hip_check(hipMemcpy(&count, id.d_out_count, sizeof(count), hipMemcpyDeviceToHost), "hipMemcpy");
                        const bool count_ok = count == out_length;
                        if (!validate(&id, d_validation, d_output, out_length, o.report_failures, &fc) || !count_ok) {
                            fprintf(stderr, "validation failure in bench %s (%d, %d, %d), run %i: %llu failures, count %llu of %llu\n",
                                    benchs[i].first.c_str(), chunk_length, block_size, grid_size, it, (unsigned long long)fc,
                                    (unsigned long long)count, (unsigned long long)out_length);
                            failures_total++;
                        }
                    }
                }
                for (size_t i = 0; i < benchs.size(); i++) {
                    timings t = ts[i] / (float)o.iterations;
                    const double gbps = t.total > 0 ? bytes / (t.total * 1e-3) / 1e9 : 0.0;
                    std::cout << benchs[i].first << ";" << chunk_length << ";" << block_size << ";" << grid_size << ";" << t << ";" << (uint64_t)bytes
                              << ";" << gbps << ";" << 100.0 * gbps / kHbmPeakGBs << ";" << shapes[i].first << ";" << shapes[i].second << std::endl;
                }
            }
        }
    }
    hip_check(hipFree(d_temp), "hipFree");
    hip_check(hipFree(d_validation), "hipFree");
    hip_check(hipFree(d_mask), "hipFree");
    hip_check(hipFree(d_output), "hipFree");
    hip_check(hipFree(d_input), "hipFree");
    if (failures_total) exit(EXIT_FAILURE);
}

} // namespace

int main(int argc, char** argv)
{
    Opts o{};
    int option;
    while ((option = getopt(argc, argv, ":zuer8d:l:i:f:p:s:t:g:m:c:k:b:n:w:x:")) != -1) {
        switch (option) {
            case '8': o.width = 8; break;
            case 'w': o.width = atoi(optarg); break;
            case 'x': o.only = optarg; break;
            case 'g': o.grid_size_max = atoi(optarg); break;
            case 'm':
                o.grid_size_min = atoi(optarg);
                if (o.grid_size_min && !pow2(o.grid_size_min)) error("min grid size has to be a power of two");
                break;
            case 'b':
                o.block_size_max = atoi(optarg);
                if (!pow2(o.block_size_max) || o.block_size_max > 1024 || o.block_size_max < 32)
                    error("max block size has to be a power of two >= 32 and <= 1024");
                break;
            case 'n':
                o.block_size_min = atoi(optarg);
                if (!pow2(o.block_size_min) || o.block_size_min > 1024 || o.block_size_min < 32)
                    error("min block size has to be a power of two >= 32 and <= 1024");
                break;
            case 'd': hip_check(hipSetDevice(atoi(optarg)), "hipSetDevice"); break;
            case 'c':
                o.chunk_length_max = atoi(optarg);
                if (!pow2(o.chunk_length_max) || o.chunk_length_max < 32) error("max chunk length has to be a power of two >= 32");
                break;
            case 'k':
                o.chunk_length_min = atoi(optarg);
                if (!pow2(o.chunk_length_min) || o.chunk_length_min < 32) error("min chunk length has to be a power of two >= 32");
                break;
            case 'l': o.lines = atoi(optarg); break;
            case 'i': o.iterations = atoi(optarg); break;
            case 'r': o.report_failures = true; break;
            case 'f':
                o.csv_path = optarg;
                o.lines = 0;
                break;
            case 'p':
                o.pattern_length = atoi(optarg);
                if (o.pattern_length > 32 || o.pattern_length < 1) o.pattern_length = 8;
                break;
            case 's': o.selectivity = (float)atof(optarg); break;
            case 't':
                o.threshold = (float)atof(optarg);
                o.use_csv = true;
                break;
            case 'z': o.use_zipf = true; break;
            case 'u': o.use_uniform = true; break;
            case 'e': o.use_clustering = true; break;
            case ':': fprintf(stderr, "-%c needs a value\n", optopt); exit(-1);
            case '?': fprintf(stderr, "unknown option: %c\n", optopt); exit(-1);
        }
    }
    if (o.use_zipf || o.use_uniform || o.use_csv || o.use_clustering) o.use_pattern_mask = false;
    if (o.use_zipf + o.use_uniform + o.use_csv + o.use_pattern_mask + o.use_clustering != 1) error("can only use one mask type");
    if (o.block_size_min > o.block_size_max) o.block_size_max = o.block_size_min;
    if (o.grid_size_min > o.grid_size_max) o.grid_size_max = o.grid_size_min;
    switch (o.width) {
        case 1: run<uint8_t>(o); break;
        case 2: run<uint16_t>(o); break;
        case 4: run<uint32_t>(o); break;
        case 8: run<uint64_t>(o); break;
        default: error("element width must be 1, 2, 4 or 8");
    }
    return 0;
}
