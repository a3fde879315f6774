// Converged height queries and the learned start-point proposal.
//
// The reference's height_at takes four fixed-point steps from the query point
// (src/ocean.cpp:375-400). The converged form iterates the same map,
// u <- X - Dx(u,v), v <- Z - Dz(u,v), on each level until the step falls to a
// tolerance, in (integer base cell + float offset) coordinates so that float
// resolution at kilometre coordinates does not limit it. The learned form
// starts each level from a small MLP's proposal instead of the query point;
// the plain iteration refines it to the same tolerance and restarts from the
// query point if the step does not fall.
//
// The MLP (per query, 3 levels): 60 features -> 64 -> 64 -> 6, ReLU, bf16
// weights and activations on MFMA (v_mfma_f32_16x16x32_bf16 for inference,
// v_mfma_f32_16x16x16_bf16 for training), fp32
// accumulate, fp32 biases. Features of level l: the (dx, dz) of the 3x3
// nodes at spacing s_l cells around the node nearest the query, divided by
// s_l * cell_l, and the query's offset from that node divided by s_l. Outputs:
// the start offset of each level, (u - X, v - Z) / (s_l * cell_l). The spacing
// s_l, the power of two in 1..16 nearest (in log scale) to 2 * rms_l / cell_l,
// follows each level's current displacement RMS, so the stencil spans the
// region the solution lies in; the nodes then sit at multiples of s_l, and
// each call gathers them once into a coarse grid per level.
#include "hip_launch.h"

#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <stdint.h>

namespace oceanamd {
namespace {

constexpr int kLevels = OCEAN_HIP_MLP_LEVELS;
constexpr int kIn = OCEAN_HIP_MLP_IN;    // 60 used + 4 zero
constexpr int kHid = OCEAN_HIP_MLP_HIDDEN;
constexpr int kOut = OCEAN_HIP_MLP_OUT;  // 6 used + 10 zero
constexpr int kOutUsed = 2 * kLevels;
constexpr int kFeatPerLevel = 20;
constexpr int oW1 = 0;
constexpr int ob1 = oW1 + kIn * kHid;
constexpr int oW2 = ob1 + kHid;
constexpr int ob2 = oW2 + kHid * kHid;
constexpr int oW3 = ob2 + kHid;
constexpr int ob3 = oW3 + kHid * kOut;
constexpr int kParams = ob3 + kOut;
static_assert(kParams == OCEAN_HIP_MLP_PARAMS, "parameter layout");
constexpr int kMaxIter = 64;
constexpr int kMaxDevices = 64;  // devices whose per-device launch facts are cached

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
typedef float f32x2v __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2v __attribute__((ext_vector_type(2)));

// Two fp32 -> a packed pair of bf16 (`lo` in the low half), round to nearest
// even, NaN stays NaN: one v_cvt_pk_bf16_f32.
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2v{lo, hi}, bf16x2v));
}
__device__ __forceinline__ uint16_t to_bf16(float f) { return uint16_t(pack_bf16(f, 0.0f)); }
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// v_mfma_f32_16x16x32_bf16 on the concatenated fragments of two 16-input
// blocks: A[i][k] from lane i + 16 * (k / 8), element k % 8 (so a lane's
// elements 0..3 are its 16x16x16 fragment of the first block and 4..7 of the
// second), the same for B, D as for 16x16x16. One instruction for what two
// chained mfma16 compute.
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ f32x4 mfma32(bf16x4 a0, bf16x4 a1, bf16x4 b0, bf16x4 b1, f32x4 c)
{
    const s16x8 a = __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7);
    const s16x8 b = __builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                   c, 0, 0, 0);
}

// 4 consecutive bf16 from LDS (8-byte aligned).
__device__ __forceinline__ bf16x4 ld4(const uint16_t* p)
{
    return *reinterpret_cast<const bf16x4*>(p);
}
// 4 bf16 at stride `s`.
__device__ __forceinline__ bf16x4 ld4s(const uint16_t* p, int s)
{
    bf16x4 r;
    r[0] = short(p[0]);
    r[1] = short(p[s]);
    r[2] = short(p[2 * s]);
    r[3] = short(p[3 * s]);
    return r;
}

// ---- geometry ---------------------------------------------------------------

struct Base {
    int32_t ix, iz;  // integer cell of the query point (wrapped to [0, N))
    float fx, fz;    // offset inside that cell, [0, 1)
};

// g = n / patch_length, cells per metre: a kernel whose lanes all query one
// level computes it once (base_cell_g) instead of per query.
__device__ __forceinline__ double cells_per_metre(const ocean_hip_level& L)
{
    return double(L.n) / double(L.patch_length);
}

__device__ __forceinline__ Base base_cell_g(float X, float Z, double g, uint32_t n)
{
    const double gx = double(X) * g, gz = double(Z) * g;
    const double flx = floor(gx), flz = floor(gz);
    const int64_t m = int64_t(n) - 1;
    Base b;
    b.ix = int32_t(int64_t(flx) & m);
    b.iz = int32_t(int64_t(flz) & m);
    b.fx = float(gx - flx);
    b.fz = float(gz - flz);
    return b;
}

__device__ __forceinline__ Base base_cell(float X, float Z, const ocean_hip_level& L)
{
    return base_cell_g(X, Z, cells_per_metre(L), L.n);
}

// Bilinear (dx, dz[, h]) at cell position (ix + px, iz + pz), periodic, as
// two lerps along x and one along z with fused multiply-adds (the converged
// queries are judged by tolerance, not bit for bit). The cell's texels are
// kept in `k` as the (x0, z0) texel and its difference to (x1, z0), and the
// (x0, z1) texel and its difference to (x1, z1): once the iteration settles
// into its final cell, later taps blend registers instead of loading again.
struct TapCache {
    uint32_t key = 0xffffffffu;  // z0 * n + x0 of the cached cell
    // Zero-initialised: the first tap always loads (key < n * n), but an
    // indeterminate member would let the compiler drop the components that
    // only the final tap reads (the height) from the loop's cached cell.
    float3 a{}, ca{}, d{}, ed{};
};

// A level's displacement as a global-memory pointer (the levels' pointers
// arrive through a struct, so the compiler cannot infer it): texel loads then
// take a scalar base plus a 32-bit byte offset (global_load with saddr)
// instead of flat loads with 64-bit address arithmetic per texel.
typedef const __attribute__((address_space(1))) float4 gfloat4;
__device__ __forceinline__ gfloat4* global_texels(const void* p) { return (gfloat4*)(p); }
// (dx, height, dz) of texel i: the foam channel is not read.
typedef float f32x3 __attribute__((ext_vector_type(3)));
__device__ __forceinline__ float3 texel(gfloat4* __restrict__ disp, uint32_t i)
{
    typedef const __attribute__((address_space(1))) char gchar;
    typedef const __attribute__((address_space(1))) f32x3 gf32x3;
    const f32x3 v = *(gf32x3*)((gchar*)disp + (i << 4));  // i < n^2 <= 4096^2: the offset fits 32 bits
    return make_float3(v[0], v[1], v[2]);
}

template <bool HEIGHT>
__device__ __forceinline__ void tap(gfloat4* __restrict__ disp, uint32_t n, const Base& b,
                                    float px, float pz, TapCache& k, float& dx, float& dz, float& h)
{
    const float flx = floorf(px), flz = floorf(pz);
    const float tu = px - flx, tv = pz - flz;
    const uint32_t mask = n - 1u;
    const uint32_t x0 = uint32_t(b.ix + int32_t(flx)) & mask;
    const uint32_t z0 = uint32_t(b.iz + int32_t(flz)) & mask;
    const uint32_t key = z0 * n + x0;
    if (key != k.key) {
        const uint32_t x1 = (x0 + 1u) & mask, z1 = (z0 + 1u) & mask;
        const float3 a = texel(disp, key), c = texel(disp, z0 * n + x1);
        const float3 d = texel(disp, z1 * n + x0), e = texel(disp, z1 * n + x1);
        k.a = make_float3(a.x, a.y, a.z);
        k.ca = make_float3(c.x - a.x, c.y - a.y, c.z - a.z);
        k.d = make_float3(d.x, d.y, d.z);
        k.ed = make_float3(e.x - d.x, e.y - d.y, e.z - d.z);
        k.key = key;
    }
    const float x_0 = __builtin_fmaf(tu, k.ca.x, k.a.x), x_1 = __builtin_fmaf(tu, k.ed.x, k.d.x);
    const float z_0 = __builtin_fmaf(tu, k.ca.z, k.a.z), z_1 = __builtin_fmaf(tu, k.ed.z, k.d.z);
    dx = __builtin_fmaf(tv, x_1 - x_0, x_0);
    dz = __builtin_fmaf(tv, z_1 - z_0, z_0);
    if (HEIGHT) {
        const float y_0 = __builtin_fmaf(tu, k.ca.y, k.a.y), y_1 = __builtin_fmaf(tu, k.ed.y, k.d.y);
        h = __builtin_fmaf(tv, y_1 - y_0, y_0);
    }
}

struct Solve {
    float du, dv;  // solved offset from the query point, cells
    int iters;     // taps taken by the inversion
    bool converged;
    bool fell_back;
};

// Fixed-point inversion of one level from the start offset (du0, dv0) cells.
__device__ __forceinline__ Solve solve_level(const ocean_hip_level& L, const Base& b, float du0,
                                             float dv0, float tol_cells, int max_iter, bool learned,
                                             TapCache& cache)
{
    gfloat4* __restrict__ disp = global_texels(L.disp);
    const float inv_cell = float(L.n) / L.patch_length;
    Solve s{du0, dv0, 0, false, false};
    float prev = 3.0e38f;
    float h;
    while (s.iters < max_iter) {
        float dx, dz;
        tap<false>(disp, L.n, b, b.fx + s.du, b.fz + s.dv, cache, dx, dz, h);
        ++s.iters;
        const float nu = -dx * inv_cell, nv = -dz * inv_cell;
        const float step = fmaxf(fabsf(nu - s.du), fabsf(nv - s.dv));
        s.du = nu;
        s.dv = nv;
        if (step <= tol_cells) {
            s.converged = true;
            break;
        }
        if (learned && !s.fell_back && s.iters >= 2 && !(step < prev)) {
            // The proposal's residual does not fall: restart from the query point.
            s.fell_back = true;
            s.du = 0.0f;
            s.dv = 0.0f;
            prev = 3.0e38f;
            continue;
        }
        prev = step;
    }
    return s;
}

// Features of one level (20 floats) at spacing s (a power of two): the nodes
// of the level's coarse grid, every s-th cell in x and z (NC = N / s per
// side, periodic), nearest to the query and its 8 neighbours, each node's
// (dx, dz) divided by s * cell; then the query's offset from the nearest node
// in units of s cells, in [-0.5, 0.5).
__device__ __forceinline__ void level_features(const ocean_hip_level& L, const Base& b, int s, float* f)
{
    const float4* __restrict__ disp = reinterpret_cast<const float4*>(L.disp);
    const uint32_t n = L.n;
    const uint32_t nc = n / uint32_t(s), cmask = nc - 1u;
    const float fs = float(s);
    const float inv_s = 1.0f / fs;  // s is a power of two: exact, and x * inv_s == x / s
    const float ux = (float(b.ix) + b.fx) * inv_s;
    const float uz = (float(b.iz) + b.fz) * inv_s;
    const float rx = floorf(ux + 0.5f), rz = floorf(uz + 0.5f);
    const int32_t cx = int32_t(rx), cz = int32_t(rz);
    const float inv_h = float(L.n) / (L.patch_length * fs);
    int k = 0;
#pragma unroll
    for (int bz = -1; bz <= 1; ++bz) {
        const uint32_t zc = uint32_t(cz + bz) & cmask;
#pragma unroll
        for (int ax = -1; ax <= 1; ++ax) {
            const uint32_t xc = uint32_t(cx + ax) & cmask;
            const float4 v = disp[(zc * uint32_t(s)) * n + xc * uint32_t(s)];
            f[k++] = v.x * inv_h;
            f[k++] = v.z * inv_h;
        }
    }
    f[k++] = ux - rx;
    f[k++] = uz - rz;
}

// ---- the MLP on one wavefront (64 rows) ---------------------------------------
//
// v_mfma_f32_16x16x16_bf16 lane layout: A[i][k] from lane i + 16 * (k / 4),
// element k % 4; B[k][j] from lane j + 16 * (k / 4), element k % 4; D[i][j]
// in lane j + 16 * (i / 4), element i % 4.

// The bf16 parameters in device memory, in the layouts the kernels read:
// W1^T [64][64], W2^T [64][64], W3^T [16][64] (rows are output units: the
// B operand takes 4 consecutive inputs), then W1, W2, W3 row-major (the
// backward pass's B = W^T operand). Biases are read in fp32 from the master.
constexpr int bT1 = 0, bT2 = bT1 + kHid * kIn, bT3 = bT2 + kHid * kHid;
constexpr int bR1 = bT3 + kOut * kHid, bR2 = bR1 + kIn * kHid, bR3 = bR2 + kHid * kHid;
constexpr int kBf16 = bR3 + kHid * kOut;
static_assert(kBf16 == OCEAN_HIP_MLP_BF16, "bf16 layout");

// bf16 position(s) of master parameter p; -1 for biases.
__device__ __forceinline__ void bf16_slots(int p, int& t, int& r)
{
    t = r = -1;
    if (p < ob1) {
        const int i = p - oW1, k = i / kHid, j = i % kHid;
        t = bT1 + j * kIn + k;
        r = bR1 + i;
    } else if (p >= oW2 && p < ob2) {
        const int i = p - oW2, k = i / kHid, j = i % kHid;
        t = bT2 + j * kHid + k;
        r = bR2 + i;
    } else if (p >= oW3 && p < ob3) {
        const int i = p - oW3, k = i / kOut, j = i % kOut;
        t = bT3 + j * kHid + k;
        r = bR3 + i;
    }
}

// Inference LDS: the transposed weights at a padded row pitch, so the
// 16 rows an MFMA operand load touches fall in distinct banks.
constexpr int kPW = 68;
struct alignas(16) InferLds {
    uint16_t wt[(kHid + kHid + kOut) * kPW];  // W1^T rows, then W2^T, then W3^T
    float b1[kHid], b2[kHid], b3[kOut];       // 16-byte aligned: read as 4-float fragments
};
static_assert(sizeof(uint16_t) * (kHid + kHid + kOut) * kPW % 16 == 0, "bias alignment");
constexpr int kWaveTile = 64 * kPW + 2 * 16 * kPW;  // X (64 rows) + two 16-row buffers

__device__ void load_infer(InferLds& w, const uint16_t* __restrict__ pb, const float* __restrict__ pf)
{
    const uint2* src = reinterpret_cast<const uint2*>(pb + bT1);
    for (int c = threadIdx.x; c < (kHid + kHid + kOut) * 16; c += blockDim.x) {
        const int row = c >> 4, part = c & 15;
        *reinterpret_cast<uint2*>(w.wt + row * kPW + part * 4) = src[c];
    }
    for (int i = threadIdx.x; i < kHid; i += blockDim.x) {
        w.b1[i] = pf[ob1 + i];
        w.b2[i] = pf[ob2 + i];
    }
    for (int i = threadIdx.x; i < kOut; i += blockDim.x) w.b3[i] = pf[ob3 + i];
}

// load_infer for a block of NT threads, with every global load issued
// before the first LDS store: one round trip instead of one per stride.
template <int NT>
__device__ __forceinline__ void load_infer_batched(InferLds& w, const uint16_t* __restrict__ pb,
                                                   const float* __restrict__ pf)
{
    constexpr int kParts = (kHid + kHid + kOut) * 16;  // 8-byte parts of the transposed weights
    static_assert(kParts % NT == 0 && NT >= kHid + kHid + kOut, "block size");
    const uint2* src = reinterpret_cast<const uint2*>(pb + bT1);
    const int t = int(threadIdx.x);
    uint2 v[kParts / NT];
#pragma unroll
    for (int k = 0; k < kParts / NT; ++k) v[k] = src[t + k * NT];
    // thread t < 144 brings bias t of b1, b2, b3 in turn
    const int bias_at = t < kHid ? ob1 + t : (t < 2 * kHid ? ob2 + t - kHid : ob3 + t - 2 * kHid);
    const float bias = t < 2 * kHid + kOut ? pf[bias_at] : 0.0f;
#pragma unroll
    for (int k = 0; k < kParts / NT; ++k) {
        const int c = t + k * NT, row = c >> 4, part = c & 15;
        *reinterpret_cast<uint2*>(w.wt + row * kPW + part * 4) = v[k];
    }
    if (t < kHid)
        w.b1[t] = bias;
    else if (t < 2 * kHid)
        w.b2[t - kHid] = bias;
    else if (t < 2 * kHid + kOut)
        w.b3[t - 2 * kHid] = bias;
}

// One 16-row tile: out = act(in * W + b). in: 16 rows at pitch PI; wt: W^T
// [NO][PWT]; RELU: bf16 out at pitch PO; else fp32 outf at pitch PO.
template <int NI, int NO, int PI, int PWT, int PO, bool RELU>
__device__ __forceinline__ void dense_tile(const uint16_t* __restrict__ in,
                                           const uint16_t* __restrict__ wt,
                                           const float* __restrict__ bias, uint16_t* __restrict__ out,
                                           float* __restrict__ outf)
{
    const int l = threadIdx.x & 63;
    const int li = l & 15, lk = (l >> 4) * 4;
#pragma unroll
    for (int ct = 0; ct < NO / 16; ++ct) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < NI / 16; ++kb) {
            const bf16x4 a = ld4(in + li * PI + kb * 16 + lk);
            const bf16x4 bb = ld4(wt + (ct * 16 + li) * PWT + kb * 16 + lk);
            acc = mfma16(a, bb, acc);
        }
        const int j = ct * 16 + li;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int i = lk + e;
            const float z = acc[e] + bias[j];
            if (RELU)
                out[i * PO + j] = to_bf16(z > 0.0f ? z : 0.0f);
            else
                outf[i * PO + j] = z;
        }
    }
}

// Writes lane l's 64 features as row l of X (bf16, 8-byte stores).
__device__ __forceinline__ void store_row(uint16_t* X, const float* f)
{
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < kIn / 4; ++c) {
        *reinterpret_cast<uint2*>(X + l * kPW + 4 * c) =
            make_uint2(pack_bf16(f[4 * c], f[4 * c + 1]), pack_bf16(f[4 * c + 2], f[4 * c + 3]));
    }
}

// The forward pass of one wavefront's 64 rows (X), 16 rows at a time through
// the two small buffers; lane l receives row l's outputs in y. Every wave of
// the block calls it (it contains block barriers).
__device__ void forward_wave(const InferLds& w, uint16_t* tile, float* y)
{
    const int l = threadIdx.x & 63;
    uint16_t* X = tile;
    uint16_t* Ha = X + 64 * kPW;
    uint16_t* Hb = Ha + 16 * kPW;
    float* Yf = reinterpret_cast<float*>(Ha);  // 16 x 16 fp32 fits in Ha
#pragma unroll 1
    for (int rt = 0; rt < 4; ++rt) {
        dense_tile<kIn, kHid, kPW, kPW, kPW, true>(X + rt * 16 * kPW, w.wt, w.b1, Ha, nullptr);
        __syncthreads();
        dense_tile<kHid, kHid, kPW, kPW, kPW, true>(Ha, w.wt + kHid * kPW, w.b2, Hb, nullptr);
        __syncthreads();
        dense_tile<kHid, kOut, kPW, kPW, kOut, false>(Hb, w.wt + 2 * kHid * kPW, w.b3, nullptr, Yf);
        __syncthreads();
        if ((l >> 4) == rt) {
#pragma unroll
            for (int k = 0; k < kOut; ++k) y[k] = Yf[(l & 15) * kOut + k];
        }
        __syncthreads();
    }
}

// ---- training layouts (unpadded; the trainer is not the hot path) --------------

struct TrainWeights {
    uint16_t w1t[kHid * kIn];    // W1^T [j][k]
    uint16_t w2t[kHid * kHid];
    uint16_t w3t[kOut * kHid];
    uint16_t w1[kIn * kHid];     // W1 [k][j]
    uint16_t w2[kHid * kHid];
    uint16_t w3[kHid * kOut];
    float b1[kHid], b2[kHid], b3[kOut];
};

__device__ void load_train(TrainWeights& w, const uint16_t* __restrict__ pb, const float* __restrict__ pf)
{
    const uint2* src = reinterpret_cast<const uint2*>(pb);
    uint2* dst = reinterpret_cast<uint2*>(w.w1t);  // the six arrays in the global order
    for (int c = threadIdx.x; c < kBf16 / 4; c += blockDim.x) dst[c] = src[c];
    for (int i = threadIdx.x; i < kHid; i += blockDim.x) {
        w.b1[i] = pf[ob1 + i];
        w.b2[i] = pf[ob2 + i];
    }
    for (int i = threadIdx.x; i < kOut; i += blockDim.x) w.b3[i] = pf[ob3 + i];
}

// out[64 x NO] = act(in[64 x NI] * W + b) over a whole wave's rows, unpadded.
template <int NI, int NO, bool RELU>
__device__ __forceinline__ void dense(const uint16_t* __restrict__ in, const uint16_t* __restrict__ wt,
                                      const float* __restrict__ bias, uint16_t* __restrict__ out,
                                      float* __restrict__ outf)
{
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
        dense_tile<NI, NO, NI, NI, NO, RELU>(in + rt * 16 * NI, wt, bias,
                                             RELU ? out + rt * 16 * NO : nullptr,
                                             RELU ? nullptr : outf + rt * 16 * NO);
}

// ---- kernels ------------------------------------------------------------------

// Displacement RMS of each level -> stencil spacing. One block per level
// (spacing_kernel), or kSpacingBlocks blocks per level writing partial sums
// that spacing_final_kernel adds in block order (spacing_partial_kernel).
constexpr int kSpacingBlocks = 64;

__device__ __forceinline__ int32_t spacing_of(double sum, const ocean_hip_level& L)
{
    // s = the power of two nearest 2 rms / cell (in log scale), 1 .. 16
    const double cells = double(L.n) * double(L.n);
    const double rms = sqrt(sum / (2.0 * cells));
    const double cell = double(L.patch_length) / double(L.n);
    const double r = 2.0 * rms / cell;
    int32_t sp = 1;
    while (sp < 16 && r >= double(sp) * 1.4142135623730951) sp *= 2;
    return sp;
}

__device__ __forceinline__ double block_sum(double acc)
{
    __shared__ double part[256];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    return part[0];
}

__global__ __launch_bounds__(256) void spacing_kernel(const ocean_hip_level* __restrict__ levels,
                                                      int32_t* __restrict__ spacing)
{
    const ocean_hip_level L = levels[blockIdx.x];
    const float4* __restrict__ disp = reinterpret_cast<const float4*>(L.disp);
    const uint32_t cells = L.n * L.n;
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < cells; i += 256u) {
        const float4 v = disp[i];
        acc += double(v.x) * double(v.x) + double(v.z) * double(v.z);
    }
    const double sum = block_sum(acc);
    if (threadIdx.x == 0) spacing[blockIdx.x] = spacing_of(sum, L);
}

__global__ __launch_bounds__(256) void spacing_partial_kernel(const ocean_hip_level* __restrict__ levels,
                                                              double* __restrict__ partial)
{
    const ocean_hip_level L = levels[blockIdx.y];
    const float4* __restrict__ disp = reinterpret_cast<const float4*>(L.disp);
    const uint32_t cells = L.n * L.n;
    double acc = 0.0;
#pragma unroll 8
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < cells; i += kSpacingBlocks * 256u) {
        const float4 v = disp[i];
        acc += double(v.x) * double(v.x) + double(v.z) * double(v.z);
    }
    const double sum = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.y * kSpacingBlocks + blockIdx.x] = sum;
}

__global__ void spacing_final_kernel(const ocean_hip_level* __restrict__ levels,
                                     const double* __restrict__ partial, uint32_t level_count,
                                     int32_t* __restrict__ spacing)
{
    const uint32_t lv = threadIdx.x;
    if (blockIdx.x != 0 || lv >= level_count) return;
    double sum = 0.0;
    for (int b = 0; b < kSpacingBlocks; ++b) sum += partial[lv * kSpacingBlocks + b];
    spacing[lv] = spacing_of(sum, levels[lv]);
}

// The levels' stencil spacings and coarse grids, for levels whose grid has at
// most OCEAN_HIP_MLP_COARSE_SIDE nodes a side. Node (xc, zc) is the
// displacement's (dx, dz) at cell (s xc, s zc) divided by s * cell, as
// level_features computes it, rounded to a bf16 pair (pack_bf16, as the
// network's input). The grid is stored as records of 16 bytes, NC rows of
// NC + 2: record (zc, r) holds the nodes (xc, zc - 1), (xc, zc), (xc, zc + 1)
// of column xc = r - 1 (mod NC), and a zero. A query's 3 x 3 stencil around
// node (cx, cz) is then records cx .. cx + 2 of row cz: 48 contiguous bytes.
// Every block adds its level's kSpacingBlocks partial sums
// (spacing_partial_kernel) in block order, as spacing_final_kernel does, and
// block 0 writes the spacing. Each thread then builds a run of kCoarseRun
// records down one record column (kCoarseRun + 2 node reads), the block
// striding over the level's runs. grid: (kCoarseBlocks, levels).
constexpr uint32_t kCoarseBlocks = 128;
constexpr uint32_t kCoarseRun = 8;
__global__ __launch_bounds__(256) void coarse_kernel(const ocean_hip_level* __restrict__ levels,
                                                     const double* __restrict__ partial,
                                                     int32_t* __restrict__ spacing,
                                                     uint4* __restrict__ coarse)
{
    __shared__ double s_part[kSpacingBlocks];
    __shared__ int32_t s_level;
    const ocean_hip_level L = levels[blockIdx.y];
    if (threadIdx.x < uint32_t(kSpacingBlocks)) s_part[threadIdx.x] = partial[blockIdx.y * kSpacingBlocks + threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int b = 0; b < kSpacingBlocks; ++b) sum += s_part[b];
        s_level = spacing_of(sum, L);
        if (blockIdx.x == 0) spacing[blockIdx.y] = s_level;
    }
    __syncthreads();
    const uint32_t s = uint32_t(s_level);
    const uint32_t nc = L.n / s, cmask = nc - 1u, pitch = nc + 2u;
    if (nc > OCEAN_HIP_MLP_COARSE_SIDE) return;
    const float inv_h = float(L.n) / (L.patch_length * float(s));
    const float4* __restrict__ disp = reinterpret_cast<const float4*>(L.disp);
    uint4* __restrict__ out = coarse + size_t(blockIdx.y) * OCEAN_HIP_MLP_COARSE_RECORDS;
    const uint32_t runs = (nc + kCoarseRun - 1u) / kCoarseRun;
    for (uint32_t t = blockIdx.x * 256u + threadIdx.x; t < runs * pitch; t += kCoarseBlocks * 256u) {
        const uint32_t zr = t / pitch, xr = t - zr * pitch;
        const uint32_t xc = (xr + cmask) & cmask, z0 = zr * kCoarseRun;  // column r - 1 (mod NC)
        uint32_t w[kCoarseRun + 2];
#pragma unroll
        for (uint32_t k = 0; k < kCoarseRun + 2; ++k) {
            const uint32_t z = (z0 + cmask + k) & cmask;  // rows z0 - 1 .. z0 + kCoarseRun (mod NC)
            const float4 v = disp[(z * s) * L.n + xc * s];
            w[k] = pack_bf16(v.x * inv_h, v.z * inv_h);
        }
#pragma unroll
        for (uint32_t k = 0; k < kCoarseRun; ++k)
            if (z0 + k < nc) out[(z0 + k) * pitch + xr] = make_uint4(w[k], w[k + 1], w[k + 2], 0u);
    }
}

// Proposals. Each wavefront takes 64 queries at a time and strides through
// the query list (a persistent grid: the network is staged into LDS once per
// block): lane l gathers query l's features level by level into row l of the
// wavefront's own LDS tile (bf16), then the network runs as W^T X^T on MFMA,
// the weight matrices as the A operand (from LDS) and the activations as the
// B operand: a layer's output tile lands in exactly the lanes and elements
// the next layer's B operand needs, so the hidden activations never leave
// registers and no wavefront waits for another (the block's one barrier is
// after the staging). The products and their summation per 16-input block are
// those of the row-major form (forward_kernel); the accumulators start from
// the biases instead of adding them last. Output: each level's start offset
// in cells, level-major float2: prop[lv * count + q].

// relu(acc + bias) rounded to bf16, as the next layer's B fragment; lane
// (i, g)'s D fragment of output tile `it` holds units 16 it + 4 g + e.
// The proposal kernel's accumulators start from the bias (z = b + sum), so
// here only the rectifier remains. It is applied after the rounding, to each
// bf16 of a packed pair as a 16-bit integer max with 0 (v_pk_max_i16): a
// bf16 with its sign bit set is a negative int16, so negative values and -0
// become +0 and the rest pass unchanged. Rounding to nearest is monotonic
// and keeps the sign, so this equals rounding z > 0 ? z : 0 for every
// non-NaN z (a NaN with a clear sign bit passes through, as numpy's maximum
// lets it).
typedef short s16x2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t relu_pair(uint32_t v)
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2v, v), s16x2v{0, 0}));
}
__device__ __forceinline__ bf16x4 relu_bf16(f32x4 z)
{
    const uint32_t lo = relu_pair(pack_bf16(z[0], z[1]));
    const uint32_t hi = relu_pair(pack_bf16(z[2], z[3]));
    return __builtin_bit_cast(bf16x4, u32x2v{lo, hi});
}
// Four consecutive fp32 from LDS (16-byte aligned): a bias fragment.
__device__ __forceinline__ f32x4 ld_bias(const float* p) { return *reinterpret_cast<const f32x4*>(p); }

// LDS ordering inside one wavefront: its LDS operations execute in program
// order, so only the compiler has to be kept from moving them.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The 20 features of one level (level_features' values and order) as 10
// bf16 pairs, read from the level's coarse-grid records (coarse_kernel): node
// k's (dx, dz) in pair k, the query's offset from the nearest node in pair 9.
// The three records of the stencil's columns are one 48-byte span; record a
// holds column cx + a - 1, rows cz - 1, cz, cz + 1. coarse_span gives the
// span's byte offset and pair 9; the records are then loaded
// (coarse_record) and unpacked (stencil_pairs).
typedef const __attribute__((address_space(1))) uint4 guint4;
typedef uint32_t u32x3v __attribute__((ext_vector_type(3)));
__device__ __forceinline__ uint32_t coarse_span(const Base& b, float inv_s, uint32_t nc, uint32_t& offset_pair)
{
    const uint32_t cmask = nc - 1u;
    const float ux = (float(b.ix) + b.fx) * inv_s;
    const float uz = (float(b.iz) + b.fz) * inv_s;
    const float rx = floorf(ux + 0.5f), rz = floorf(uz + 0.5f);
    const uint32_t cx = uint32_t(int32_t(rx)) & cmask, cz = uint32_t(int32_t(rz)) & cmask;
    offset_pair = pack_bf16(ux - rx, uz - rz);
    // record cx of row cz: < 512 * 514 * 16 bytes, fits 32 bits
    return (cz * (nc + 2u) + cx) << 4;
}
// The record's three node pairs (its fourth dword, the zero, is not read:
// a 16-byte load's unused register would be reused while the load is in
// flight, and the compiler then waits for it).
__device__ __forceinline__ u32x3v coarse_record(guint4* __restrict__ coarse, uint32_t byte_offset)
{
    typedef const __attribute__((address_space(1))) char gchar;
    typedef const __attribute__((address_space(1))) u32x3v gu32x3;
    return *(gu32x3*)((gchar*)coarse + byte_offset);
}
__device__ __forceinline__ void stencil_pairs(u32x3v c0, u32x3v c1, u32x3v c2, uint32_t offset_pair, uint32_t* pk)
{
#pragma unroll
    for (int bz = 0; bz < 3; ++bz) {
        pk[3 * bz] = c0[bz];
        pk[3 * bz + 1] = c1[bz];
        pk[3 * bz + 2] = c2[bz];
    }
    pk[9] = offset_pair;
}
__device__ __forceinline__ void coarse_features_bf16(const Base& b, float inv_s, uint32_t nc,
                                                     guint4* __restrict__ coarse, uint32_t* pk)
{
    uint32_t op;
    const uint32_t off = coarse_span(b, inv_s, nc, op);
    stencil_pairs(coarse_record(coarse, off), coarse_record(coarse, off + 16u), coarse_record(coarse, off + 32u),
                  op, pk);
}

// Level lv's ten feature pairs into row l of the wavefront's X tile.
__device__ __forceinline__ void store_level_pairs(uint32_t* X, int l, int lv, const uint32_t* pk)
{
    uint2* dst = reinterpret_cast<uint2*>(X + l * (kLevels * kFeatPerLevel / 2) + lv * (kFeatPerLevel / 2));
#pragma unroll
    for (int c = 0; c < kFeatPerLevel / 4; ++c) dst[c] = make_uint2(pk[2 * c], pk[2 * c + 1]);
}

constexpr int kPropWaves = 4;                 // wavefronts per block
constexpr int kXP = kLevels * kFeatPerLevel;  // X row pitch, bf16: the 60 features
static_assert(kXP == 60 && kIn == 64, "X holds features 0..59; inputs 60..63 are zero");
// A row pitch of 30 dwords puts the 16 rows an operand load touches on 16
// distinct bank pairs, with no padding: 64 x 60 bf16 = 7.5 KB a wavefront,
// so three 4-wave blocks (weights included) fit a CU's LDS.

__global__ __launch_bounds__(64 * kPropWaves) __attribute__((amdgpu_waves_per_eu(3))) void proposal_kernel(
    const ocean_hip_level* __restrict__ levels, uint32_t level_count, const float2* __restrict__ xz,
    uint32_t count, const uint16_t* __restrict__ pb, const float* __restrict__ pf,
    const int32_t* __restrict__ spacing, const uint4* __restrict__ coarse, float2* __restrict__ prop)
{
    __shared__ InferLds w;
    __shared__ __align__(16) uint32_t tiles[kPropWaves][64 * kXP / 2];
    const int l = threadIdx.x & 63;
    const int i = l & 15, g = l >> 4;
    uint32_t* X = tiles[threadIdx.x >> 6];
    load_infer_batched<64 * kPropWaves>(w, pb, pf);
    const uint32_t used = level_count < uint32_t(kLevels) ? level_count : uint32_t(kLevels);
    // feature columns of levels not in use stay zero
    for (uint32_t c = used * (kFeatPerLevel / 2); c < uint32_t(kXP / 2); ++c) X[l * (kXP / 2) + c] = 0u;
    // each level's constants, once per block
    double gl[kLevels];
    float inv_s[kLevels], sp[kLevels];
    uint32_t nl[kLevels], ncl[kLevels];
    guint4* cgl[kLevels];
#pragma unroll
    for (int lv = 0; lv < kLevels; ++lv) {
        gl[lv] = 0.0;
        inv_s[lv] = sp[lv] = 0.0f;
        nl[lv] = ncl[lv] = 0u;
        cgl[lv] = nullptr;
        if (uint32_t(lv) < used) {
            const ocean_hip_level L = levels[lv];
            const int32_t s = spacing[lv];
            gl[lv] = cells_per_metre(L);
            sp[lv] = float(s);
            inv_s[lv] = 1.0f / float(s);  // s is a power of two: exact
            nl[lv] = L.n;
            ncl[lv] = L.n / uint32_t(s);
            if (ncl[lv] <= OCEAN_HIP_MLP_COARSE_SIDE)
                cgl[lv] = (guint4*)(coarse + size_t(lv) * OCEAN_HIP_MLP_COARSE_RECORDS);
        }
    }
    // The common case, all three levels in use and on coarse grids, gathers
    // in straight-line code: the nine record loads are all issued before the
    // first wait.
    const bool straight = used == uint32_t(kLevels) && cgl[0] != nullptr && cgl[1] != nullptr && cgl[2] != nullptr;
    __syncthreads();
    const uint32_t waves = gridDim.x * kPropWaves;
    const uint32_t first = (blockIdx.x * kPropWaves + (threadIdx.x >> 6)) * 64u;
    // Each round's query point is loaded one round ahead, so that load's
    // latency overlaps the previous round's network.
    float2 p_next = first + uint32_t(l) < count ? xz[first + uint32_t(l)] : make_float2(0.0f, 0.0f);
    for (uint32_t base = first; base < count; base += waves * 64u) {
        const uint32_t q = base + uint32_t(l);
        const bool live = q < count;
        const float2 p = p_next;
        const uint32_t qn = q + waves * 64u;
        if (qn < count) p_next = xz[qn];
        wave_lds_sync();  // the previous round's reads of X are done
        if (straight) {
            // Lanes past the end gather at a stale or zero point: every
            // offset is wrapped into its grid, the rows are independent
            // columns of the products, and their outputs are not stored.
            //
            // The wavefront's 64 spans of a level are 192 records; load k
            // (k = 0, 1, 2) has lane l fetch record 64 k + l of that list,
            // record a = (64 k + l) % 3 of query row (64 k + l) / 3, so
            // consecutive lanes read one span's consecutive records and a
            // load touches about 64 / 3 spans' cache lines instead of 64.
            // The span's offset comes from its query's lane by a lane
            // shuffle, and the record's three node pairs (column a, rows
            // cz - 1, cz, cz + 1) go straight to that row's feature slots
            // 3 bz + a.
            uint32_t off[kLevels], op[kLevels];
#pragma unroll
            for (int lv = 0; lv < kLevels; ++lv)
                off[lv] = coarse_span(base_cell_g(p.x, p.y, gl[lv], nl[lv]), inv_s[lv], ncl[lv], op[lv]);
            // this lane's share of the loads, derived each round (an empty
            // asm keeps the compiler from holding it across the network)
            uint32_t lr = uint32_t(l);
            asm volatile("" : "+v"(lr));
            uint32_t span_row[3], span_rec[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                span_row[k] = (64u * uint32_t(k) + lr) / 3u;
                span_rec[k] = (64u * uint32_t(k) + lr) - 3u * span_row[k];
            }
#pragma unroll
            for (int lv = 0; lv < kLevels; ++lv) {
                u32x3v c[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const uint32_t o = uint32_t(__shfl(int(off[lv]), int(span_row[k]), 64));
                    c[k] = coarse_record(cgl[lv], o + 16u * span_rec[k]);
                }
                X[l * (kXP / 2) + lv * (kFeatPerLevel / 2) + 9] = op[lv];
#pragma unroll
                for (int k = 0; k < 3; ++k)
#pragma unroll
                    for (int bz = 0; bz < 3; ++bz)
                        X[span_row[k] * (kXP / 2) + lv * (kFeatPerLevel / 2) + 3 * bz + span_rec[k]] = c[k][bz];
            }
        } else {
#pragma unroll
            for (int lv = 0; lv < kLevels; ++lv) {
                if (uint32_t(lv) < used) {
                    uint32_t pk[kFeatPerLevel / 2];
                    if (live) {
                        const Base b = base_cell_g(p.x, p.y, gl[lv], nl[lv]);
                        if (cgl[lv] != nullptr) {
                            coarse_features_bf16(b, inv_s[lv], ncl[lv], cgl[lv], pk);
                        } else {
                            float f[kFeatPerLevel];
                            level_features(levels[lv], b, int(sp[lv]), f);
#pragma unroll
                            for (int k = 0; k < kFeatPerLevel / 2; ++k) pk[k] = pack_bf16(f[2 * k], f[2 * k + 1]);
                        }
                    } else {
#pragma unroll
                        for (int k = 0; k < kFeatPerLevel / 2; ++k) pk[k] = 0u;
                    }
                    store_level_pairs(X, l, lv, pk);
                }
            }
        }
        wave_lds_sync();
        // The weights' LDS addresses, re-derived every round through an empty
        // asm: otherwise the compiler hoists all 36 loop-invariant A fragments
        // out of the loop into 72 registers, which halves the wavefronts the
        // feature gathers can overlap.
        uint32_t wofs = uint32_t(i * kPW + 4 * g);
        asm volatile("" : "+v"(wofs));
        const uint16_t* W1 = w.wt + wofs;
        const uint16_t* W2 = W1 + kHid * kPW;
        const uint16_t* W3 = W1 + 2 * kHid * kPW;
        const uint16_t* Xh = reinterpret_cast<const uint16_t*>(X);
#pragma unroll 1
        for (int qt = 0; qt < 4; ++qt) {
            // B fragments of X^T: inputs 16 kb + 4 g .. + 3 of query 16 qt + i;
            // inputs 60..63 (kb = 3, g = 3) are the zero features
            const uint16_t* xr = Xh + (16 * qt + i) * kXP + 4 * g;
            bf16x4 x[4];
            x[0] = ld4(xr);
            x[1] = ld4(xr + 16);
            x[2] = ld4(xr + 32);
            const bf16x4 x3 = ld4(g < 3 ? xr + 48 : xr);
            x[3] = g < 3 ? x3 : bf16x4{0, 0, 0, 0};
            bf16x4 h1[4], h2[4];
#pragma unroll
            for (int it = 0; it < 4; ++it) {
                f32x4 acc = ld_bias(w.b1 + 16 * it + 4 * g);
#pragma unroll
                for (int kb = 0; kb < 4; kb += 2)
                    acc = mfma32(ld4(W1 + 16 * it * kPW + 16 * kb), ld4(W1 + 16 * it * kPW + 16 * kb + 16),
                                 x[kb], x[kb + 1], acc);
                h1[it] = relu_bf16(acc);
            }
#pragma unroll
            for (int it = 0; it < 4; ++it) {
                f32x4 acc = ld_bias(w.b2 + 16 * it + 4 * g);
#pragma unroll
                for (int kb = 0; kb < 4; kb += 2)
                    acc = mfma32(ld4(W2 + 16 * it * kPW + 16 * kb), ld4(W2 + 16 * it * kPW + 16 * kb + 16),
                                 h1[kb], h1[kb + 1], acc);
                h2[it] = relu_bf16(acc);
            }
            f32x4 acc = ld_bias(w.b3 + 4 * g);
#pragma unroll
            for (int kb = 0; kb < 4; kb += 2)
                acc = mfma32(ld4(W3 + 16 * kb), ld4(W3 + 16 * kb + 16), h2[kb], h2[kb + 1], acc);
            // lane (i, g) holds outputs 4 g .. 4 g + 3 of query base + 16 qt + i:
            // g = 0 levels 0 and 1, g = 1 level 2 (outputs 4, 5)
            const uint32_t qo = base + uint32_t(16 * qt + i);
            if (qo < count && g < 2) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int lv = 2 * g + h;
                    if (lv < kLevels && uint32_t(lv) < used) {
                        const float s_lv = lv == 0 ? sp[0] : (lv == 1 ? sp[1] : sp[2]);
                        const float du = acc[2 * h] * s_lv;
                        const float dv = acc[2 * h + 1] * s_lv;
                        prop[size_t(lv) * count + qo] = make_float2(du, dv);
                    }
                }
            }
        }
    }
}

// Converged queries, one lane per query; LEARNED starts each of the first
// kLevels levels from prop (cells) and falls back to the query point.
template <bool LEARNED>
__global__ __launch_bounds__(256) void converged_kernel(
    const ocean_hip_level* __restrict__ levels, uint32_t level_count, uint32_t lv_begin,
    uint32_t lv_end, const float2* __restrict__ xz, uint32_t count, float tol,
    const float2* __restrict__ prop, float* __restrict__ heights, uint8_t* __restrict__ iters,
    uint8_t* __restrict__ flags)
{
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    if (q >= count) return;
    const float2 p = xz[q];
    // Levels [lv_begin, lv_end) of level_count; a later range continues the
    // earlier ranges' sum and flags (the same float additions in level order).
    float height = lv_begin == 0 ? 0.0f : heights[q];
    uint8_t fl = (lv_begin == 0 || flags == nullptr) ? uint8_t(0) : flags[q];
    for (uint32_t lv = lv_begin; lv < lv_end; ++lv) {
        const ocean_hip_level L = levels[lv];
        const Base b = base_cell(p.x, p.y, L);
        const float inv_cell = float(L.n) / L.patch_length;
        const bool use = LEARNED && lv < uint32_t(kLevels);
        float du0 = 0.0f, dv0 = 0.0f;
        if (use) {
            const float2 st = prop[size_t(lv) * count + q];
            du0 = st.x;
            dv0 = st.y;
            if (!(fabsf(du0) < 1.0e4f) || !(fabsf(dv0) < 1.0e4f)) {
                du0 = dv0 = 0.0f;  // not finite or far off the patch: abandoned at once
                fl |= 2u;
            }
        }
        TapCache cache;
        const Solve s = solve_level(L, b, du0, dv0, tol * inv_cell, kMaxIter, use, cache);
        float dx, dz, h;
        tap<true>(global_texels(L.disp), L.n, b, b.fx + s.du, b.fz + s.dv, cache, dx, dz, h);
        height += h;
        if (iters) iters[size_t(q) * level_count + lv] = uint8_t(s.iters > 255 ? 255 : s.iters);
        if (!s.converged) fl |= 1u;
        if (s.fell_back) fl |= 2u;
    }
    heights[q] = height;
    if (flags) flags[q] = fl;
}

// Training pairs: features and the solved start offsets (tight tolerance).
__global__ __launch_bounds__(256) void pairs_kernel(const ocean_hip_level* __restrict__ levels,
                                                    const float2* __restrict__ xz, uint32_t count,
                                                    float tol, const int32_t* __restrict__ spacing,
                                                    float* __restrict__ feats, float* __restrict__ targets,
                                                    uint8_t* __restrict__ ok)
{
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    if (q >= count) return;
    const float2 p = xz[q];
    float f[kIn];
#pragma unroll
    for (int k = 0; k < kIn; ++k) f[k] = 0.0f;
    float t[kOut];
#pragma unroll
    for (int k = 0; k < kOut; ++k) t[k] = 0.0f;
    bool good = true;
    for (int lv = 0; lv < kLevels; ++lv) {
        const ocean_hip_level L = levels[lv];
        const Base b = base_cell(p.x, p.y, L);
        const int s = spacing[lv];
        level_features(L, b, s, f + lv * kFeatPerLevel);
        const float inv_cell = float(L.n) / L.patch_length;
        TapCache cache;
        const Solve r = solve_level(L, b, 0.0f, 0.0f, tol * inv_cell, 255, false, cache);
        good = good && r.converged;
        t[2 * lv] = r.du / float(s);
        t[2 * lv + 1] = r.dv / float(s);
    }
    for (int k = 0; k < kIn; ++k) feats[size_t(q) * kIn + k] = f[k];
    for (int k = 0; k < kOut; ++k) targets[size_t(q) * kOut + k] = t[k];
    ok[q] = good ? 1u : 0u;
}

// Forward only (rows of fp32 features -> fp32 outputs), the proposal
// kernel's network code, 64 rows per wave.
__global__ __launch_bounds__(256) void forward_kernel(const float* __restrict__ feats, uint32_t rows,
                                                      const uint16_t* __restrict__ pb,
                                                      const float* __restrict__ pf,
                                                      float* __restrict__ out)
{
    extern __shared__ __align__(16) unsigned char smem[];
    InferLds& w = *reinterpret_cast<InferLds*>(smem);
    uint16_t* tile = reinterpret_cast<uint16_t*>(smem + sizeof(InferLds)) + (threadIdx.x >> 6) * kWaveTile;
    load_infer(w, pb, pf);
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    float f[kIn];
    for (int k = 0; k < kIn; ++k) f[k] = r < rows ? feats[size_t(r) * kIn + k] : 0.0f;
    store_row(tile, f);
    __syncthreads();
    float y[kOut];
    forward_wave(w, tile, y);
    if (r < rows)
        for (int k = 0; k < kOut; ++k) out[size_t(r) * kOut + k] = y[k];
}

// ---- training: one block = 2 waves = 128 rows of the minibatch -----------------

constexpr int kTrainRows = 128;

struct TrainTiles {
    uint16_t X[kTrainRows * kIn];
    uint16_t H1[kTrainRows * kHid];
    uint16_t H2[kTrainRows * kHid];
    uint16_t dY[kTrainRows * kOut];
    uint16_t dZ2[kTrainRows * kHid];
    uint16_t dZ1[kTrainRows * kHid];
    float sq[kTrainRows];
};

// dst[a x b tile grid] = sum over the block's rows of A^T[.][r] * B[r][.]:
// dW tile (kt, jt) of A (pitch NA) and B (pitch NB); ONES: A is all ones (bias
// gradient, row 0 of the result is used).
template <int NA, int NB, bool ONES>
__device__ __forceinline__ f32x4 grad_tile(const uint16_t* __restrict__ A, const uint16_t* __restrict__ B,
                                           int kt, int jt)
{
    const int l = threadIdx.x & 63;
    const int li = l & 15, lk = (l >> 4) * 4;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const short one = short(0x3f80);
#pragma unroll
    for (int rb = 0; rb < kTrainRows / 16; ++rb) {
        const int r = rb * 16 + lk;
        bf16x4 a;
        if (ONES) {
            a[0] = a[1] = a[2] = a[3] = one;
        } else {
            a = ld4s(A + r * NA + kt * 16 + li, NA);
        }
        const bf16x4 b = ld4s(B + r * NB + jt * 16 + li, NB);
        acc = mfma16(a, b, acc);
    }
    return acc;
}

__device__ __forceinline__ void store_grad_tile(float* __restrict__ dst, int pitch, int kt, int jt,
                                                f32x4 acc, bool row0_only)
{
    const int l = threadIdx.x & 63;
    const int j = jt * 16 + (l & 15);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int i = (l >> 4) * 4 + e;
        if (row0_only) {
            if (i == 0) dst[j] = acc[e];
        } else {
            dst[(kt * 16 + i) * pitch + j] = acc[e];
        }
    }
}

__global__ __launch_bounds__(128) void train_kernel(const float* __restrict__ feats,
                                                    const float* __restrict__ targets,
                                                    const uint32_t* __restrict__ rows, uint32_t batch,
                                                    const uint16_t* __restrict__ pb,
                                                    const float* __restrict__ pf,
                                                    float* __restrict__ partial,
                                                    float* __restrict__ loss_partial)
{
    extern __shared__ __align__(16) unsigned char smem[];
    TrainWeights& w = *reinterpret_cast<TrainWeights*>(smem);
    TrainTiles& t = *reinterpret_cast<TrainTiles*>(smem + sizeof(TrainWeights));
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int lr = wave * 64 + l;  // row inside the block
    const uint32_t row = rows[blockIdx.x * kTrainRows + lr];
    load_train(w, pb, pf);
    for (int k = 0; k < kIn; ++k) t.X[lr * kIn + k] = to_bf16(feats[size_t(row) * kIn + k]);
    __syncthreads();
    // forward, each wave on its 64 rows
    float Yf[kOut];
    {
        float* Y = reinterpret_cast<float*>(t.dZ2) + wave * 64 * kOut;  // scratch until dZ2 is written
        dense<kIn, kHid, true>(t.X + wave * 64 * kIn, w.w1t, w.b1, t.H1 + wave * 64 * kHid, nullptr);
        __syncthreads();
        dense<kHid, kHid, true>(t.H1 + wave * 64 * kHid, w.w2t, w.b2, t.H2 + wave * 64 * kHid, nullptr);
        __syncthreads();
        dense<kHid, kOut, false>(t.H2 + wave * 64 * kHid, w.w3t, w.b3, nullptr, Y);
        __syncthreads();
        for (int k = 0; k < kOut; ++k) Yf[k] = Y[l * kOut + k];
        __syncthreads();
    }
    // loss 1/(B*6) sum (y - t)^2 and dY = 2 (y - t) / (B*6)
    const float scale = 2.0f / (float(batch) * float(kOutUsed));
    float sq = 0.0f;
    for (int k = 0; k < kOut; ++k) {
        float g = 0.0f;
        if (k < kOutUsed) {
            const float d = Yf[k] - targets[size_t(row) * kOut + k];
            sq += d * d;
            g = d * scale;
        }
        t.dY[lr * kOut + k] = to_bf16(g);
    }
    t.sq[lr] = sq;
    __syncthreads();
    // dH2 = dY W3^T, dZ2 = dH2 * (H2 > 0)
    {
        const int li = l & 15, lk = (l >> 4) * 4;
        const uint16_t* dY = t.dY + wave * 64 * kOut;
        const uint16_t* H2 = t.H2 + wave * 64 * kHid;
        uint16_t* dZ2 = t.dZ2 + wave * 64 * kHid;
        for (int rt = 0; rt < 4; ++rt)
            for (int ct = 0; ct < kHid / 16; ++ct) {
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                const bf16x4 a = ld4(dY + (rt * 16 + li) * kOut + lk);
                const bf16x4 b = ld4(w.w3 + (ct * 16 + li) * kOut + lk);
                acc = mfma16(a, b, acc);
                const int j = ct * 16 + li;
                for (int e = 0; e < 4; ++e) {
                    const int i = rt * 16 + lk + e;
                    dZ2[i * kHid + j] = (H2[i * kHid + j] & 0x7fffu) != 0u && (H2[i * kHid + j] & 0x8000u) == 0u
                                            ? to_bf16(acc[e])
                                            : uint16_t(0);
                }
            }
    }
    __syncthreads();
    // dH1 = dZ2 W2^T, dZ1 = dH1 * (H1 > 0)
    {
        const int li = l & 15, lk = (l >> 4) * 4;
        const uint16_t* dZ2 = t.dZ2 + wave * 64 * kHid;
        const uint16_t* H1 = t.H1 + wave * 64 * kHid;
        uint16_t* dZ1 = t.dZ1 + wave * 64 * kHid;
        for (int rt = 0; rt < 4; ++rt)
            for (int ct = 0; ct < kHid / 16; ++ct) {
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                for (int kb = 0; kb < kHid / 16; ++kb) {
                    const bf16x4 a = ld4(dZ2 + (rt * 16 + li) * kHid + kb * 16 + lk);
                    const bf16x4 b = ld4(w.w2 + (ct * 16 + li) * kHid + kb * 16 + lk);
                    acc = mfma16(a, b, acc);
                }
                const int j = ct * 16 + li;
                for (int e = 0; e < 4; ++e) {
                    const int i = rt * 16 + lk + e;
                    dZ1[i * kHid + j] = (H1[i * kHid + j] & 0x7fffu) != 0u && (H1[i * kHid + j] & 0x8000u) == 0u
                                            ? to_bf16(acc[e])
                                            : uint16_t(0);
                }
            }
    }
    __syncthreads();
    // weight and bias gradients over the block's 128 rows: 45 tiles over 2 waves
    float* P = partial + size_t(blockIdx.x) * kParams;
    for (int tile = wave; tile < 45; tile += 2) {
        if (tile < 16) {  // dW1 = X^T dZ1
            const int kt = tile / 4, jt = tile % 4;
            store_grad_tile(P + oW1, kHid, kt, jt, grad_tile<kIn, kHid, false>(t.X, t.dZ1, kt, jt), false);
        } else if (tile < 32) {  // dW2 = H1^T dZ2
            const int kt = (tile - 16) / 4, jt = (tile - 16) % 4;
            store_grad_tile(P + oW2, kHid, kt, jt, grad_tile<kHid, kHid, false>(t.H1, t.dZ2, kt, jt), false);
        } else if (tile < 36) {  // dW3 = H2^T dY
            const int kt = tile - 32;
            store_grad_tile(P + oW3, kOut, kt, 0, grad_tile<kHid, kOut, false>(t.H2, t.dY, kt, 0), false);
        } else if (tile < 40) {  // db1 = 1^T dZ1
            const int jt = tile - 36;
            store_grad_tile(P + ob1, 0, 0, jt, grad_tile<kIn, kHid, true>(nullptr, t.dZ1, 0, jt), true);
        } else if (tile < 44) {  // db2
            const int jt = tile - 40;
            store_grad_tile(P + ob2, 0, 0, jt, grad_tile<kHid, kHid, true>(nullptr, t.dZ2, 0, jt), true);
        } else {  // db3
            store_grad_tile(P + ob3, 0, 0, 0, grad_tile<kHid, kOut, true>(nullptr, t.dY, 0, 0), true);
        }
    }
    if (threadIdx.x == 0) {
        float s = 0.0f;
        for (int r = 0; r < kTrainRows; ++r) s += t.sq[r];
        loss_partial[blockIdx.x] = s;
    }
}

// Sum of the per-block partial gradients (block order) and, if lr > 0, one
// Adam step on the fp32 master parameters, refreshing their bf16 copies.
__global__ __launch_bounds__(256) void adam_kernel(const float* __restrict__ partial, uint32_t blocks,
                                                   float* __restrict__ grad, float* __restrict__ master,
                                                   float* __restrict__ m1, float* __restrict__ m2,
                                                   uint16_t* __restrict__ pb, float lr, float c1,
                                                   float c2)
{
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= uint32_t(kParams)) return;
    float g = 0.0f;
    for (uint32_t b = 0; b < blocks; ++b) g += partial[size_t(b) * kParams + p];
    grad[p] = g;
    if (!(lr > 0.0f)) return;
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1.0e-8f;
    const float a = beta1 * m1[p] + (1.0f - beta1) * g;
    const float v = beta2 * m2[p] + (1.0f - beta2) * g * g;
    m1[p] = a;
    m2[p] = v;
    const float mhat = a / c1, vhat = v / c2;
    const float w = master[p] - lr * mhat / (sqrtf(vhat) + eps);
    master[p] = w;
    int t, r;
    bf16_slots(int(p), t, r);
    if (t >= 0) {
        pb[t] = to_bf16(w);
        pb[r] = to_bf16(w);
    }
}

__global__ void to_bf16_kernel(const float* __restrict__ master, uint16_t* __restrict__ pb)
{
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= kParams) return;
    int t, r;
    bf16_slots(p, t, r);
    if (t >= 0) {
        pb[t] = to_bf16(master[p]);
        pb[r] = to_bf16(master[p]);
    }
}

__global__ void loss_sum_kernel(const float* __restrict__ part, uint32_t n, float* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float s = 0.0f;
    for (uint32_t i = 0; i < n; ++i) s += part[i];
    *out = s;
}

constexpr size_t kInferSmem = sizeof(InferLds) + 4 * kWaveTile * sizeof(uint16_t);
constexpr size_t kTrainSmem = sizeof(TrainWeights) + sizeof(TrainTiles);
static_assert(kInferSmem <= 160 * 1024 && kTrainSmem <= 160 * 1024, "LDS");

}  // namespace
}  // namespace oceanamd

using namespace oceanamd;

extern "C" uint32_t ocean_hip_mlp_train_blocks(uint32_t batch) { return batch / uint32_t(kTrainRows); }

extern "C" hipError_t ocean_hip_mlp_spacing_partial(const ocean_hip_level* levels, uint32_t level_count,
                                                    double* partial, hipStream_t s)
{
    if (level_count == 0) return hipSuccess;
    if (partial == nullptr || level_count > 256u) return hipErrorInvalidValue;
    hipLaunchKernelGGL(spacing_partial_kernel, dim3(kSpacingBlocks, level_count), dim3(256), 0, s, levels,
                       partial);
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_mlp_spacing(const ocean_hip_level* levels, uint32_t level_count,
                                            int32_t* spacing, double* partial, hipStream_t s)
{
    if (level_count == 0) return hipSuccess;
    if (partial == nullptr) {
        hipLaunchKernelGGL(spacing_kernel, dim3(level_count), dim3(256), 0, s, levels, spacing);
    } else {
        if (level_count > 256u) return hipErrorInvalidValue;
        hipLaunchKernelGGL(spacing_partial_kernel, dim3(kSpacingBlocks, level_count), dim3(256), 0, s,
                           levels, partial);
        hipLaunchKernelGGL(spacing_final_kernel, dim3(1), dim3(256), 0, s, levels, partial, level_count,
                           spacing);
    }
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_mlp_propose(const ocean_hip_level* levels, uint32_t level_count,
                                            const float* xz, uint32_t count, const ocean_hip_mlp* mlp,
                                            const double* spacing_partial, int32_t* spacing,
                                            float* coarse, float* prop, hipStream_t s)
{
    if (count == 0) return hipSuccess;
    if (mlp == nullptr || spacing_partial == nullptr || spacing == nullptr || prop == nullptr ||
        coarse == nullptr || level_count == 0)
        return hipErrorInvalidValue;
    uint4* coarse2 = reinterpret_cast<uint4*>(coarse);
    hipLaunchKernelGGL(coarse_kernel,
                       dim3(kCoarseBlocks, level_count < uint32_t(kLevels) ? level_count : uint32_t(kLevels)),
                       dim3(256), 0, s, levels, spacing_partial, spacing, coarse2);
    // resident proposal blocks, looked up once per device (a race between two
    // first calls only repeats the lookup)
    static std::atomic<int> resident[kMaxDevices];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    int prop_grid = resident[dev].load(std::memory_order_relaxed);
    if (prop_grid == 0) {
        int cus = 0, per = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, proposal_kernel, 64 * kPropWaves, 0) !=
                hipSuccess)
            return hipErrorInvalidValue;
        prop_grid = cus * (per > 0 ? per : 1);
        resident[dev].store(prop_grid, std::memory_order_relaxed);
    }
    const uint32_t need = (count + 64u * kPropWaves - 1u) / (64u * kPropWaves);
    const uint32_t pg = need < uint32_t(prop_grid) ? need : uint32_t(prop_grid);
    hipLaunchKernelGGL(proposal_kernel, dim3(pg), dim3(64 * kPropWaves), 0, s, levels, level_count,
                       reinterpret_cast<const float2*>(xz), count, mlp->bf16, mlp->master, spacing,
                       coarse2, reinterpret_cast<float2*>(prop));
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_height_converged(const ocean_hip_level* levels, uint32_t level_count,
                                                 const float* xz, uint32_t count, float tol,
                                                 const ocean_hip_mlp* mlp, const double* spacing_partial,
                                                 int32_t* spacing, float* coarse, float* prop,
                                                 float* heights, uint8_t* iters, uint8_t* flags,
                                                 hipStream_t s)
{
    if (count == 0) return hipSuccess;
    const dim3 grid((count + 255u) / 256u);
    const float2* p = reinterpret_cast<const float2*>(xz);
    float2* prop2 = reinterpret_cast<float2*>(prop);
    if (mlp != nullptr) {
        const hipError_t e = ocean_hip_mlp_propose(levels, level_count, xz, count, mlp, spacing_partial,
                                                   spacing, coarse, prop, s);
        if (e != hipSuccess) return e;
    }
    // One launch per level: all queries tap one level's 4 MB (at 512^2) at a
    // time, which an XCD's L2 holds, instead of every level's at once.
    for (uint32_t lv = 0; lv < level_count; ++lv) {
        if (mlp != nullptr)
            hipLaunchKernelGGL(converged_kernel<true>, grid, dim3(256), 0, s, levels, level_count, lv,
                               lv + 1, p, count, tol, prop2, heights, iters, flags);
        else
            hipLaunchKernelGGL(converged_kernel<false>, grid, dim3(256), 0, s, levels, level_count,
                               lv, lv + 1, p, count, tol, nullptr, heights, iters, flags);
    }
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_mlp_pairs(const ocean_hip_level* levels, const float* xz,
                                          uint32_t count, float tol, const int32_t* spacing,
                                          float* feats, float* targets, uint8_t* ok, hipStream_t s)
{
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(pairs_kernel, dim3((count + 255u) / 256u), dim3(256), 0, s, levels,
                       reinterpret_cast<const float2*>(xz), count, tol, spacing, feats, targets, ok);
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_mlp_forward(const ocean_hip_mlp* mlp, const float* feats,
                                            uint32_t rows, float* out, hipStream_t s)
{
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(forward_kernel, dim3((rows + 255u) / 256u), dim3(256), kInferSmem, s, feats,
                       rows, mlp->bf16, mlp->master, out);
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_mlp_sync_bf16(ocean_hip_mlp* mlp, hipStream_t s)
{
    hipLaunchKernelGGL(to_bf16_kernel, dim3((kParams + 255) / 256), dim3(256), 0, s, mlp->master,
                       mlp->bf16);
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_mlp_step(ocean_hip_mlp* mlp, const float* feats, const float* targets,
                                         const uint32_t* rows, uint32_t batch, float lr, float c1,
                                         float c2, float* partial, float* loss_partial, float* loss,
                                         hipStream_t s)
{
    const uint32_t blocks = batch / uint32_t(kTrainRows);
    if (blocks == 0 || blocks * uint32_t(kTrainRows) != batch) return hipErrorInvalidValue;
    hipLaunchKernelGGL(train_kernel, dim3(blocks), dim3(128), kTrainSmem, s, feats, targets, rows, batch,
                       mlp->bf16, mlp->master, partial, loss_partial);
    hipLaunchKernelGGL(adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, s, partial, blocks,
                       mlp->grad, mlp->master, mlp->m1, mlp->m2, mlp->bf16, lr, c1, c2);
    if (loss != nullptr) hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(64), 0, s, loss_partial, blocks, loss);
    return hipGetLastError();
}

extern "C" int ocean_hip_mlp_configure()
{
    // set once per device (the attributes are idempotent: concurrent first
    // calls both set them)
    static std::atomic<bool> done[kMaxDevices];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -1;
    if (done[dev].load(std::memory_order_acquire)) return 0;
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&forward_kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, int(kInferSmem)) != hipSuccess)
        return -1;
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&train_kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, int(kTrainSmem)) != hipSuccess)
        return -1;
    done[dev].store(true, std::memory_order_release);
    return 0;
}

// ---- the allocation-free converged query's scratch ------------------------
// A file-scope device array holds what ocean_hip_height_converged needs
// besides its inputs and outputs: the levels' stencil spacings, their partial
// sums, their coarse grids and the per-query start offsets, for up to
// OCEAN_HIP_STATIC_MAX_QUERIES
// queries. Each GPU holds its own copy until the process ends; every call on
// one GPU shares it, so calls that overlap on one GPU race. Host code never
// passes the array's name to a kernel (in host code it is a stand-in, not a
// device address): it uses the address hipGetSymbolAddress returns, looked
// up once per device with that device current.
namespace {
constexpr size_t kStaticSpacingBytes = 256;
constexpr size_t kStaticPartialBytes =
    (size_t(OCEAN_HIP_MLP_LEVELS) * OCEAN_HIP_MLP_SPACING_PARTIALS * sizeof(double) + 255) / 256 * 256;
constexpr size_t kStaticCoarseBytes = OCEAN_HIP_MLP_COARSE_BYTES;
constexpr size_t kStaticPropBytes = size_t(OCEAN_HIP_STATIC_MAX_QUERIES) * OCEAN_HIP_MLP_PROP * sizeof(float);
constexpr size_t kStaticBytes = kStaticSpacingBytes + kStaticPartialBytes + kStaticCoarseBytes + kStaticPropBytes;
static_assert(kStaticBytes == OCEAN_HIP_STATIC_SCRATCH_BYTES, "static scratch layout");
constexpr int kStaticMaxDevices = kMaxDevices;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kStaticBytes];

extern "C" hipError_t ocean_hip_static_scratch(ocean_hip_static_scratch_view* out)
{
    static std::mutex mu;
    static void* base[kStaticMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kStaticMaxDevices) return hipErrorInvalidDevice;
    void* p = nullptr;
    {
        std::lock_guard<std::mutex> lock(mu);
        if (base[dev] == nullptr) {
            e = hipGetSymbolAddress(&base[dev], HIP_SYMBOL(scratch));
            if (e != hipSuccess) {
                base[dev] = nullptr;
                return e;
            }
        }
        p = base[dev];
    }
    char* b = static_cast<char*>(p);
    out->spacing = reinterpret_cast<int32_t*>(b);
    out->spacing_partial = reinterpret_cast<double*>(b + kStaticSpacingBytes);
    out->coarse = reinterpret_cast<float*>(b + kStaticSpacingBytes + kStaticPartialBytes);
    out->prop = reinterpret_cast<float*>(b + kStaticSpacingBytes + kStaticPartialBytes + kStaticCoarseBytes);
    out->max_queries = OCEAN_HIP_STATIC_MAX_QUERIES;
    return hipSuccess;
}
