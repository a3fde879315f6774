// A minimal case harness shared by the reference-case programs: a case is a
// named function that throws CaseFailed on the first broken expectation.
#pragma once

#include "ocean/ocean.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace cases {

struct CaseFailed {
    std::string why;
};

struct Case {
    const char* name;
    void (*run)();
};

inline std::string g_case;

[[noreturn]] inline void fail(const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw CaseFailed{buf};
}

#define EXPECT(cond)                                                                 \
    do {                                                                             \
        if (!(cond)) ::cases::fail("line %d: %s", __LINE__, #cond);                 \
    } while (0)

// A measured value, printed as "INFO <case>: ..." for the log.
inline void info(const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    std::printf("INFO %s: %s\n", g_case.c_str(), buf);
}

inline std::vector<float> copy_of(const float* p, std::size_t count) { return {p, p + count}; }

inline bool same_bits(const float* a, const float* b, std::size_t count)
{
    return std::memcmp(a, b, count * sizeof(float)) == 0;
}

inline double rel_close(double a, double b)
{
    return std::fabs(a - b) / std::max(std::fabs(b), 1e-300);
}

// |a - b| <= eps (1 + max(|a|, |b|)): the reference tests' Approx(b).epsilon(eps)
inline bool approx(double a, double b, double eps)
{
    return std::fabs(a - b) <= eps * (1.0 + std::max(std::fabs(a), std::fabs(b)));
}

// Host scheduler hooks that honour the reference's parallel_for contract (every
// index once, blocking) in three orders, counting their own calls: the
// library accepts such hooks and never calls them.
inline std::atomic<int> g_hook_calls{0};

inline void counting_hook(void*, ocean::TaskFn task, void* ctx, std::uint32_t count)
{
    g_hook_calls.fetch_add(1);
    for (std::uint32_t i = 0; i < count; ++i) task(ctx, i);
}

inline void reverse_hook(void*, ocean::TaskFn task, void* ctx, std::uint32_t count)
{
    g_hook_calls.fetch_add(1);
    for (std::uint32_t i = count; i-- > 0;) task(ctx, i);
}

inline void scattering_hook(void*, ocean::TaskFn task, void* ctx, std::uint32_t count)
{
    g_hook_calls.fetch_add(1);
    std::atomic<std::uint32_t> next{0};
    std::vector<std::thread> workers;
    for (int w = 0; w < 4; ++w)
        workers.emplace_back([&] {
            for (std::uint32_t i; (i = next.fetch_add(1)) < count;) task(ctx, i);
        });
    for (auto& t : workers) t.join();
}

// Heap allocations (operator new) made from this library's own code since
// the program started: the program replaces the global operator new and
// counts the calls whose caller lies in liboceanlib_amd (the HIP runtime's
// own allocations inside each launch are not this library's).
std::size_t library_allocations();

extern const Case kFrameCases[];
extern const std::size_t kFrameCaseCount;
extern const Case kInteractionCases[];
extern const std::size_t kInteractionCaseCount;

}  // namespace cases
