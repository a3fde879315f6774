// Exhaustive comparison of the device expf (oceanlib_amd/csrc/device_expf.hpp,
// compiled here for the host) with the host C library's expf over every float
// in [lo, hi] (both <= 0), split across threads. Prints the mismatch count
// and the first mismatching input. Used by tests/test_device_expf.py.
#include "../../oceanlib_amd/csrc/device_expf.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

int main(int argc, char** argv)
{
    const float lo = std::strtof(argv[1], nullptr);   // most negative
    const float hi = std::strtof(argv[2], nullptr);   // closest to zero (e.g. -0.0)
    std::uint32_t ulo, uhi;
    std::memcpy(&ulo, &lo, 4);
    std::memcpy(&uhi, &hi, 4);
    // negative floats: larger bit pattern = more negative
    const std::uint64_t a = uhi, b = std::uint64_t(ulo) + 1;
    const unsigned T = std::max(1u, std::thread::hardware_concurrency());
    std::vector<long long> bad(T, 0);
    std::vector<std::uint32_t> first(T, 0);
    std::vector<std::thread> th;
    for (unsigned k = 0; k < T; ++k) {
        th.emplace_back([&, k] {
            const std::uint64_t s = a + (b - a) * k / T, e = a + (b - a) * (k + 1) / T;
            for (std::uint64_t u = s; u < e; ++u) {
                float x;
                const std::uint32_t u32 = static_cast<std::uint32_t>(u);
                std::memcpy(&x, &u32, 4);
                const float g = ::expf(x), m = oceanamd::expf_libm(x);
                if (std::memcmp(&g, &m, 4) != 0) {
                    if (bad[k] == 0) first[k] = u32;
                    ++bad[k];
                }
            }
        });
    }
    for (auto& t : th) t.join();
    long long total = 0;
    std::uint32_t f = 0;
    for (unsigned k = 0; k < T; ++k) {
        if (bad[k] && !total) f = first[k];
        total += bad[k];
    }
    std::printf("%llu %lld 0x%08x\n", (unsigned long long)(b - a), total, f);
    return 0;
}
