// Device-side query helpers shared by the frame, foam, cascade and combined
// query kernels: the periodic bilinear tap and the fixed-point inversion of
// the reference's Ocean::sample_at (src/ocean.cpp:309-452), and the foam
// field's periodic bilinear fetch (src/foam.cpp:22-41).
#pragma once

#include "hip_launch.h"

#include <hip/hip_runtime.h>
#include <cstdint>

namespace oceanamd {

struct Tap {
    uint32_t i00, i10, i01, i11;
    float w00, w10, w01, w11;
};

// The reference's periodic wrap of an int index, ((i % n) + n) % n
// (src/ocean.cpp:322-352). Every grid here has a power-of-two side (16 ..
// 4096, checked when an ocean is created), and for a two's-complement int
// that wrap is i & (n - 1): the same index without an integer division (its
// remainder sequence costs quarter-rate multiplies on every tap).
__device__ __forceinline__ uint32_t wrap_index(int i, int n)
{
    return static_cast<uint32_t>(i) & static_cast<uint32_t>(n - 1);
}

__device__ __forceinline__ Tap make_tap(float u, float v, uint32_t n, float inv_cell)
{
    const float gu = u * inv_cell;
    const float gv = v * inv_cell;
    const float fu = floorf(gu);
    const float fv = floorf(gv);
    const int ni = static_cast<int>(n);
    const uint32_t mask = n - 1u, sh = static_cast<uint32_t>(__builtin_ctz(n));  // n = 2^sh
    const uint32_t x0 = wrap_index(static_cast<int>(fu), ni);
    const uint32_t z0 = wrap_index(static_cast<int>(fv), ni);
    const uint32_t x1 = (x0 + 1u) & mask;
    const uint32_t z1 = (z0 + 1u) & mask;
    const float tu = gu - fu;
    const float tv = gv - fv;
    Tap t;
    t.i00 = (z0 << sh) + x0;
    t.i10 = (z0 << sh) + x1;
    t.i01 = (z1 << sh) + x0;
    t.i11 = (z1 << sh) + x1;
    t.w00 = (1.0f - tu) * (1.0f - tv);
    t.w10 = tu * (1.0f - tv);
    t.w01 = (1.0f - tu) * tv;
    t.w11 = tu * tv;
    return t;
}

__device__ __forceinline__ float4 fetch4(const float4* __restrict__ b, const Tap& t)
{
    const float4 p = b[t.i00], q = b[t.i10], r = b[t.i01], s = b[t.i11];
    float4 o;
    o.x = p.x * t.w00 + q.x * t.w10 + r.x * t.w01 + s.x * t.w11;
    o.y = p.y * t.w00 + q.y * t.w10 + r.y * t.w01 + s.y * t.w11;
    o.z = p.z * t.w00 + q.z * t.w10 + r.z * t.w01 + s.z * t.w11;
    o.w = p.w * t.w00 + q.w * t.w10 + r.w * t.w01 + s.w * t.w11;
    return o;
}

// The four texels of the cell a lane tapped last. The inversion's later taps
// mostly land in the same cell (about 2.3 distinct cells in the 5 taps of a
// query at 512^2, heavy chop), so a tap into the cached cell blends the
// registers instead of loading again; the blend is the same expression on the
// same values, so results are unchanged.
struct CellCache {
    uint32_t i00 = 0xffffffffu;  // never a real cell: the first tap loads
    float4 p{}, q{}, r{}, s{};   // zeroed, so no path reads an indeterminate value
};

__device__ __forceinline__ float4 blend_cached(const Tap& t, const CellCache& c)
{
    float4 o;
    o.x = c.p.x * t.w00 + c.q.x * t.w10 + c.r.x * t.w01 + c.s.x * t.w11;
    o.y = c.p.y * t.w00 + c.q.y * t.w10 + c.r.y * t.w01 + c.s.y * t.w11;
    o.z = c.p.z * t.w00 + c.q.z * t.w10 + c.r.z * t.w01 + c.s.z * t.w11;
    o.w = c.p.w * t.w00 + c.q.w * t.w10 + c.r.w * t.w01 + c.s.w * t.w11;
    return o;
}

__device__ __forceinline__ float4 fetch4_cached(const float4* __restrict__ b, const Tap& t,
                                                CellCache& c)
{
    if (t.i00 != c.i00) {  // i00 fixes the cell, hence i10, i01 and i11
        c.p = b[t.i00];
        c.q = b[t.i10];
        c.r = b[t.i01];
        c.s = b[t.i11];
        c.i00 = t.i00;
    }
    float4 o;
    o.x = c.p.x * t.w00 + c.q.x * t.w10 + c.r.x * t.w01 + c.s.x * t.w11;
    o.y = c.p.y * t.w00 + c.q.y * t.w10 + c.r.y * t.w01 + c.s.y * t.w11;
    o.z = c.p.z * t.w00 + c.q.z * t.w10 + c.r.z * t.w01 + c.s.z * t.w11;
    o.w = c.p.w * t.w00 + c.q.w * t.w10 + c.r.w * t.w01 + c.s.w * t.w11;
    return o;
}

// One point of Ocean::sample_at on one level (reference src/ocean.cpp:375-452).
// Returns the final tap so callers can reuse the solved parameter position.
__device__ __forceinline__ ocean_hip_surface sample_level(const float4* __restrict__ disp,
                                                          const float4* __restrict__ normal,
                                                          const float4* __restrict__ vel,
                                                          uint32_t n, float inv_cell, float wx,
                                                          float wz)
{
    float u = wx, v = wz;
    CellCache cache;
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        const Tap t = make_tap(u, v, n, inv_cell);
        const float4 d = fetch4_cached(disp, t, cache);
        u = wx - d.x;
        v = wz - d.z;
    }
    const Tap t = make_tap(u, v, n, inv_cell);
    const float4 d = fetch4_cached(disp, t, cache);
    const float4 nr = fetch4(normal, t);
    float nx = nr.x, ny = nr.y, nz = nr.z;
    const float len2 = nx * nx + ny * ny + nz * nz;
    if (len2 > 0.0f) {
        const float inv = 1.0f / sqrtf(len2);
        nx *= inv;
        ny *= inv;
        nz *= inv;
    } else {
        nx = 0.0f;
        ny = 1.0f;
        nz = 0.0f;
    }
    ocean_hip_surface s;
    s.height = d.y;
    s.offset_x = d.x;
    s.offset_z = d.z;
    s.normal_x = nx;
    s.normal_y = ny;
    s.normal_z = nz;
    s.foam = d.w;
    if (vel != nullptr) {
        const float4 w = fetch4(vel, t);
        s.velocity_x = w.x;
        s.velocity_y = w.y;
        s.velocity_z = w.z;
    } else {
        s.velocity_x = 0.0f;
        s.velocity_y = 0.0f;
        s.velocity_z = 0.0f;
    }
    return s;
}

// The height alone of sample_level (the same taps and blend; no normal or
// velocity loads): Ocean::height_at on one level.
__device__ __forceinline__ float height_level(const float4* __restrict__ disp, uint32_t n,
                                              float inv_cell, float wx, float wz)
{
    float u = wx, v = wz;
    CellCache cache;
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        const Tap t = make_tap(u, v, n, inv_cell);
        const float4 d = fetch4_cached(disp, t, cache);
        u = wx - d.x;
        v = wz - d.z;
    }
    const Tap t = make_tap(u, v, n, inv_cell);
    return fetch4_cached(disp, t, cache).y;
}

// height_level with two lanes per point: lanes 2j and 2j + 1 of a
// wavefront follow the same point. When a tap leaves the cached cell, the
// even lane loads the cell's two x0 texels and the odd lane its two x1
// texels, so each row of the cell is one cache line for the pair instead of
// two lanes' worth, and they exchange them with a lane-pair shuffle (both
// lanes of a pair always take the same branch). The same texels and the same
// blend as height_level.
__device__ __forceinline__ float4 pair_swap(float4 a)
{
    return make_float4(__shfl_xor(a.x, 1, 64), __shfl_xor(a.y, 1, 64), __shfl_xor(a.z, 1, 64), a.w);
}
__device__ __forceinline__ float height_level_lane_pair(const float4* __restrict__ disp, uint32_t n,
                                                        float inv_cell, float wx, float wz, bool odd)
{
    float u = wx, v = wz;
    CellCache c;
#pragma unroll
    for (int it = 0; it < 5; ++it) {
        const Tap t = make_tap(u, v, n, inv_cell);
        if (t.i00 != c.i00) {
            const float4 m0 = disp[odd ? t.i10 : t.i00], m1 = disp[odd ? t.i11 : t.i01];
            const float4 o0 = pair_swap(m0), o1 = pair_swap(m1);
            c.p = odd ? o0 : m0;
            c.q = odd ? m0 : o0;
            c.r = odd ? o1 : m1;
            c.s = odd ? m1 : o1;
            c.i00 = t.i00;
        }
        const float4 d = blend_cached(t, c);
        if (it == 4) return d.y;
        u = wx - d.x;
        v = wz - d.z;
    }
    return 0.0f;
}

__device__ __forceinline__ float sample_wrap(const float* __restrict__ f, float u, float v, int n)
{
    const float fu = floorf(u);
    const float fv = floorf(v);
    const float tu = u - fu;
    const float tv = v - fv;
    const int x0 = static_cast<int>(wrap_index(static_cast<int>(fu), n));
    const int z0 = static_cast<int>(wrap_index(static_cast<int>(fv), n));
    const int x1 = (x0 + 1) & (n - 1);
    const int z1 = (z0 + 1) & (n - 1);
    const int sh = __builtin_ctz(n);  // n = 2^sh
    const float a = f[(z0 << sh) + x0];
    const float b = f[(z0 << sh) + x1];
    const float c = f[(z1 << sh) + x0];
    const float d = f[(z1 << sh) + x1];
    return a * (1.0f - tu) * (1.0f - tv) + b * tu * (1.0f - tv) + c * (1.0f - tu) * tv +
           d * tu * tv;
}


// Cascade query: reference src/cascade.cpp:82-153 (levels in order, height,
// offsets and velocity summed, normals summed then renormalised, foam
// screen-blended).
__device__ __forceinline__ ocean_hip_surface sample_cascade(const ocean_hip_level* __restrict__ levels,
                                                            uint32_t level_count, float wx, float wz)
{
    float height = 0.0f, ox = 0.0f, oz = 0.0f, nx = 0.0f, ny = 0.0f, nz = 0.0f;
    float transp = 1.0f, vx = 0.0f, vy = 0.0f, vz = 0.0f;
    for (uint32_t l = 0; l < level_count; ++l) {
        const ocean_hip_level L = levels[l];
        const float inv_cell = static_cast<float>(L.n) / L.patch_length;
        const ocean_hip_surface s = sample_level(
            reinterpret_cast<const float4*>(L.disp), reinterpret_cast<const float4*>(L.normal),
            reinterpret_cast<const float4*>(L.vel), L.n, inv_cell, wx, wz);
        height += s.height;
        ox += s.offset_x;
        oz += s.offset_z;
        nx += s.normal_x;
        ny += s.normal_y;
        nz += s.normal_z;
        vx += s.velocity_x;
        vy += s.velocity_y;
        vz += s.velocity_z;
        transp *= (1.0f - s.foam);
    }
    ocean_hip_surface o;
    o.height = height;
    o.offset_x = ox;
    o.offset_z = oz;
    o.foam = 1.0f - transp;
    o.velocity_x = vx;
    o.velocity_y = vy;
    o.velocity_z = vz;
    const float len2 = nx * nx + ny * ny + nz * nz;
    if (len2 > 0.0f) {
        const float inv = 1.0f / sqrtf(len2);
        o.normal_x = nx * inv;
        o.normal_y = ny * inv;
        o.normal_z = nz * inv;
    } else {
        o.normal_x = 0.0f;
        o.normal_y = 1.0f;
        o.normal_z = 0.0f;
    }
    return o;
}

}  // namespace oceanamd
