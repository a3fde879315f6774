// CPU emulation of the GPU team FFT (oceanlib_amd/csrc/fft_team.hpp): the
// same fft_team<N, ..., R> code (R points per lane) runs with one host thread
// per lane and a
// std::barrier standing in for the block barrier. Used by
// tests/test_fft_team_emulation.py to compare the kernel's butterfly graph
// with the oracle's FFT bit for bit, without a GPU.
#include "../../oceanlib_amd/csrc/fft_team.hpp"

#include <barrier>
#include <cstdint>
#include <thread>
#include <vector>

using namespace oceanamd;

template <int N, int R = 8>
static void run(float* re, float* im, const float* twr, const float* twi)
{
    constexpr int T = FftShape<N, R>::T;
    std::vector<float2> tw(N / 2);
    for (int j = 0; j < N / 2; ++j) tw[j] = make_float2(twr[j], twi[j]);
    std::vector<float2> xch(lds_span<N>());
    std::barrier bar(T);
    std::vector<std::thread> lanes;
    std::vector<float> ore(N), oim(N);
    for (int J = 0; J < T; ++J) {
        lanes.emplace_back([&, J] {
            float r[R], i[R];
            uint32_t idx[R];
            for (int a = 0; a < R; ++a) {
                r[a] = re[J + a * T];
                i[a] = im[J + a * T];
            }
            auto sync = [&] { bar.arrive_and_wait(); };
            fft_team<N, decltype(sync), false, R>(r, i, idx, uint32_t(J), xch.data(), tw.data(),
                                                 sync);
            for (int a = 0; a < R; ++a) {
                ore[idx[a]] = r[a];
                oim[idx[a]] = i[a];
            }
        });
    }
    for (auto& t : lanes) t.join();
    for (int j = 0; j < N; ++j) {
        re[j] = ore[j];
        im[j] = oim[j];
    }
}

// K sequences transformed together by fft_team_k (re/im hold K * N values).
template <int N, int K>
static void run_k(float* re, float* im, const float* twr, const float* twi)
{
    constexpr int T = FftShape<N>::T;
    constexpr uint32_t kstride = team_stride<N>();
    std::vector<float2> tw(N / 2);
    for (int j = 0; j < N / 2; ++j) tw[j] = make_float2(twr[j], twi[j]);
    std::vector<float2> xch(K * kstride);
    std::barrier bar(T);
    std::vector<std::thread> lanes;
    std::vector<float> ore(K * N), oim(K * N);
    for (int J = 0; J < T; ++J) {
        lanes.emplace_back([&, J] {
            float r[K][8], i[K][8];
            uint32_t idx[8];
            for (int k = 0; k < K; ++k)
                for (int a = 0; a < 8; ++a) {
                    r[k][a] = re[k * N + J + a * T];
                    i[k][a] = im[k * N + J + a * T];
                }
            fft_team_k<N, K>(r, i, idx, uint32_t(J), xch.data(), kstride, tw.data(),
                             [&] { bar.arrive_and_wait(); });
            for (int k = 0; k < K; ++k)
                for (int a = 0; a < 8; ++a) {
                    ore[k * N + idx[a]] = r[k][a];
                    oim[k * N + idx[a]] = i[k][a];
                }
        });
    }
    for (auto& t : lanes) t.join();
    for (int j = 0; j < K * N; ++j) {
        re[j] = ore[j];
        im[j] = oim[j];
    }
}

template <int N>
static int run_k_n(uint32_t k, float* re, float* im, const float* twr, const float* twi)
{
    switch (k) {
        case 2: run_k<N, 2>(re, im, twr, twi); return 0;
        case 3: run_k<N, 3>(re, im, twr, twi); return 0;
    }
    return -1;
}

extern "C" int emul_fft_team_k(uint32_t n, uint32_t k, float* re, float* im, const float* twr,
                               const float* twi)
{
    switch (n) {
        case 16: return run_k_n<16>(k, re, im, twr, twi);
        case 64: return run_k_n<64>(k, re, im, twr, twi);
        case 512: return run_k_n<512>(k, re, im, twr, twi);
        case 4096: return run_k_n<4096>(k, re, im, twr, twi);
    }
    return -1;
}

extern "C" int emul_fft_team(uint32_t n, float* re, float* im, const float* twr,
                             const float* twi)
{
    switch (n) {
        case 16: run<16>(re, im, twr, twi); return 0;
        case 32: run<32>(re, im, twr, twi); return 0;
        case 64: run<64>(re, im, twr, twi); return 0;
        case 128: run<128>(re, im, twr, twi); return 0;
        case 256: run<256>(re, im, twr, twi); return 0;
        case 512: run<512>(re, im, twr, twi); return 0;
        case 1024: run<1024>(re, im, twr, twi); return 0;
        case 2048: run<2048>(re, im, twr, twi); return 0;
        case 4096: run<4096>(re, im, twr, twi); return 0;
    }
    return -1;
}

template <int R>
static int run_r(uint32_t n, float* re, float* im, const float* twr, const float* twi)
{
    switch (n) {
        case 32: run<32, R>(re, im, twr, twi); return 0;
        case 64: run<64, R>(re, im, twr, twi); return 0;
        case 128: run<128, R>(re, im, twr, twi); return 0;
        case 256: run<256, R>(re, im, twr, twi); return 0;
        case 512: run<512, R>(re, im, twr, twi); return 0;
        case 1024: run<1024, R>(re, im, twr, twi); return 0;
        case 2048: run<2048, R>(re, im, twr, twi); return 0;
    }
    return -1;
}

// Radix groupings other than 8 points per lane (R = 4 or 16), N in 32..2048.
extern "C" int emul_fft_team_r(uint32_t n, uint32_t r, float* re, float* im, const float* twr,
                               const float* twi)
{
    switch (r) {
        case 4: return run_r<4>(n, re, im, twr, twi);
        case 16: return run_r<16>(n, re, im, twr, twi);
    }
    return -1;
}
