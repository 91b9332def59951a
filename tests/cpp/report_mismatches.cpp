// Host harness for csx_compat::report_mismatches (no device calls): plants
// mismatches in copies of a 1000-element array of each width and prints the
// report to stdout, then the returned count.
#include <csx_compat.hpp>

#include <vector>

template <typename T> static void run(uint64_t limit, std::vector<uint64_t> bad)
{
    std::vector<T> e(1000), g;
    for (size_t i = 0; i < e.size(); i++) e[i] = (T)(0x0102030405060708ull * (i + 1));
    g = e;
    for (uint64_t i : bad) reinterpret_cast<unsigned char*>(&g[i])[0] ^= 0x5A;  // the low byte of the element's bits
    const uint64_t n = csx_compat::report_mismatches(e.data(), g.data(), e.size(), limit, stdout);
    printf("count %llu\n", (unsigned long long)n);
}

int main()
{
    run<uint8_t>(64, {0, 999});
    run<uint32_t>(2, {5, 17, 400, 401});
    run<uint64_t>(64, {});
    run<double>(64, {3});
    return 0;
}
