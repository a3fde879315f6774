// CPU emulation of the stamp-mean kernel's sum (oceanlib_amd/csrc/exact_sum.hpp):
// chunk statistics gathered the way a wavefront gathers them (lane l adds the
// chunk's terms l, l + 64, l + 128, l + 192 in order, then a butterfly over
// lane distance 32 .. 1, lane 0 keeping the result), then the chain over the
// chunks with the exactness test. Used by tests/test_exact_sum_emulation.py
// to compare with the plain sequential double sum bit for bit.
#include "../../oceanlib_amd/csrc/exact_sum.hpp"

#include <cstdint>

using namespace oceanamd;

namespace {

constexpr int kLanes = 64;

ChunkStats chunk_stats(const float* t, int n)
{
    ChunkAcc lane[kLanes];
    for (int l = 0; l < kLanes; ++l)
        for (int i = l; i < n; i += kLanes) lane[l].add(t[i]);
    for (int off = kLanes / 2; off >= 1; off /= 2) {
        ChunkAcc next[kLanes];
        for (int l = 0; l < kLanes; ++l) {
            next[l] = lane[l];
            const ChunkAcc& o = lane[l ^ off];
            next[l].merge(o.total, o.abs_sum, o.q, o.finite);
        }
        for (int l = 0; l < kLanes; ++l) lane[l] = next[l];
    }
    return lane[0].finish();
}

}  // namespace

extern "C" double emul_sum_plain(const float* t, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += static_cast<double>(t[i]);
    return s;
}

// chunk <= 256 (four terms per lane). *skipped: chunks taken in one step.
extern "C" double emul_sum_chunked(const float* t, int n, int chunk, int* skipped)
{
    double s = 0.0;
    int sk = 0;
    for (int c0 = 0; c0 < n; c0 += chunk) {
        const int len = n - c0 < chunk ? n - c0 : chunk;
        const ChunkStats st = chunk_stats(t + c0, len);
        if (chunk_is_exact(s, st)) {
            s += st.total;
            ++sk;
        } else {
            for (int i = 0; i < len; ++i) s += static_cast<double>(t[c0 + i]);
        }
    }
    if (skipped) *skipped = sk;
    return s;
}
