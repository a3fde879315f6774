// CPU emulation of the stamp-mean kernel's sum (oceanlib_amd/csrc/exact_sum.hpp):
// chunk statistics gathered the way a wavefront gathers them (lane l adds the
// chunk's terms l, l + 64, l + 128, l + 192 in order, then a butterfly over
// lane distance 32 .. 1, lane 0 keeping the result), the statistics of each
// group of consecutive chunks merged in order from theirs, then the chain
// over the groups and chunks with the exactness test. Used by
// tests/test_exact_sum_emulation.py to compare with the plain sequential
// double sum bit for bit.
#include "../../oceanlib_amd/csrc/exact_sum.hpp"

#include <cstdint>

using namespace oceanamd;

namespace {

constexpr int kLanes = 64;

ChunkAcc chunk_acc(const float* t, int n)
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
    return lane[0];
}

}  // namespace

extern "C" double emul_sum_plain(const float* t, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += static_cast<double>(t[i]);
    return s;
}

// chunk <= 256 (four terms per lane), 1 <= group <= 256 chunks per group.
// *skipped: groups and chunks taken in one step.
extern "C" double emul_sum_chunked(const float* t, int n, int chunk, int group, int* skipped)
{
    const int chunks = (n + chunk - 1) / chunk;
    double s = 0.0;
    int sk = 0;
    for (int g0 = 0; g0 < chunks; g0 += group) {
        const int g1 = chunks - g0 < group ? chunks : g0 + group;
        ChunkAcc acc[256];
        ChunkAcc whole;
        for (int c = g0; c < g1; ++c) {
            const int len = n - c * chunk < chunk ? n - c * chunk : chunk;
            acc[c - g0] = chunk_acc(t + c * chunk, len);
            whole.merge(acc[c - g0].total, acc[c - g0].abs_sum, acc[c - g0].q, acc[c - g0].finite);
        }
        const ChunkStats gs = whole.finish();
        if (chunk_is_exact(s, gs)) {
            s += gs.total;
            sk += g1 - g0;
            continue;
        }
        for (int c = g0; c < g1; ++c) {
            const ChunkStats st = acc[c - g0].finish();
            const int len = n - c * chunk < chunk ? n - c * chunk : chunk;
            if (chunk_is_exact(s, st)) {
                s += st.total;
                ++sk;
            } else {
                for (int i = 0; i < len; ++i) s += static_cast<double>(t[c * chunk + i]);
            }
        }
    }
    if (skipped) *skipped = sk;
    return s;
}
