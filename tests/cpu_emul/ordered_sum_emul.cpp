// CPU emulation of the coverage kernel's ordered sum (oceanlib_amd/csrc/
// ordered_sum.hpp, coverage_kernel in wave_kernels.hip): one workgroup of
// 1024 lanes, 4 consecutive terms per lane, chunks of 4096; per step each
// lane rounds its terms to the running value's binade and keeps a running
// sum of them, a Hillis-Steele scan over each 64-lane wavefront (lane l adds
// lane l - off's value for off = 1, 2, ..., 32), the wavefronts' totals added
// in order, the first index that ends the run, then s += the run's prefix
// and s += the ending term. Used by tests/test_ordered_sum_emulation.py to
// compare with the plain sequential double sum bit for bit.
#include "../../oceanlib_amd/csrc/ordered_sum.hpp"

#include <algorithm>
#include <cstdint>
#include <vector>

using namespace oceanamd;

namespace {

constexpr uint32_t kLanes = 1024, kPer = 4, kChunk = kLanes * kPer, kWaves = kLanes / 64;

}  // namespace

extern "C" double emul_sum_plain(const float* v, uint32_t n)
{
    double s = 0.0;
    for (uint32_t i = 0; i < n; ++i) s += static_cast<double>(v[i]);
    return s;
}

// Returns the sum (not divided by n); *steps counts the workgroup steps taken.
extern "C" double emul_sum_ordered(const float* v, uint32_t n, uint32_t* steps)
{
    double s = 0.0;
    uint32_t st_count = 0;
    std::vector<double> p(kChunk), lane_run(kLanes), scan(kLanes), before(kLanes);
    std::vector<bool> stop(kChunk);
    for (uint32_t base = 0; base < n; base += kChunk) {
        float t[kChunk];
        for (uint32_t i = 0; i < kChunk; ++i) t[i] = base + i < n ? v[base + i] : 0.0f;
        uint32_t pos = 0;
        while (pos < kChunk) {
            ++st_count;
            const Binade bin = binade_of(s);
            if (bin.mode == 2) {
                for (uint32_t i = base + pos; i < n; ++i) s += static_cast<double>(v[i]);
                *steps = st_count;
                return s;
            }
            for (uint32_t l = 0; l < kLanes; ++l) {
                double run = 0.0;
                for (uint32_t j = 0; j < kPer; ++j) {
                    const uint32_t i = l * kPer + j;
                    bool st = false;
                    double r = 0.0;
                    if (i >= pos) r = rounded_term(t[i], bin, st);
                    run += r;
                    p[i] = run;
                    stop[i] = st;
                }
                lane_run[l] = run;
            }
            // wavefront inclusive scans
            for (uint32_t w = 0; w < kWaves; ++w) {
                double x[64];
                for (uint32_t l = 0; l < 64; ++l) x[l] = lane_run[w * 64 + l];
                for (uint32_t off = 1; off < 64; off <<= 1) {
                    double y[64];
                    for (uint32_t l = 0; l < 64; ++l) y[l] = l >= off ? x[l - off] : 0.0;
                    for (uint32_t l = off; l < 64; ++l) x[l] += y[l];
                }
                for (uint32_t l = 0; l < 64; ++l) scan[w * 64 + l] = x[l];
            }
            double wave_sum[kWaves];
            for (uint32_t w = 0; w < kWaves; ++w) wave_sum[w] = scan[w * 64 + 63];
            uint32_t e = kChunk;
            for (uint32_t l = 0; l < kLanes; ++l) {
                const uint32_t w = l / 64, ll = l % 64;
                double b = 0.0;
                for (uint32_t k = 0; k < w; ++k) b += wave_sum[k];
                b += ll == 0 ? 0.0 : scan[l - 1];
                before[l] = b;
                for (uint32_t j = 0; j < kPer; ++j) {
                    const uint32_t i = l * kPer + j;
                    if (i >= pos && ends_run(stop[i], b + p[i], bin)) {
                        e = std::min(e, i);
                        break;
                    }
                }
            }
            if (e == kChunk) {
                double total = 0.0;
                for (uint32_t w = 0; w < kWaves; ++w) total += wave_sum[w];
                s += total;
                pos = kChunk;
            } else {
                const uint32_t l = e / kPer, j = e % kPer;
                const double pre = j == 0 ? before[l] : before[l] + p[l * kPer + j - 1];
                s += pre;
                s += static_cast<double>(t[e]);
                pos = e + 1;
            }
        }
    }
    *steps = st_count;
    return s;
}
