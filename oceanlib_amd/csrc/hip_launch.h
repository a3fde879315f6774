/* Thin C-ABI layer between the C++20 host code and the HIP kernels.
 * Plain pointers and sizes only; every pointer below is device memory. */
#ifndef OCEANAMD_HIP_LAUNCH_H
#define OCEANAMD_HIP_LAUNCH_H

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#if defined(__cplusplus)
extern "C" {
#endif

/* One ocean level's device state. */
typedef struct ocean_hip_level {
    uint32_t n;                 /* grid size N                                 */
    float patch_length;
    float choppiness;
    float foam_threshold;
    float inv_foam_threshold;   /* host-computed as in evolve.cpp:247-248      */
    const float* h0;            /* N*N float2 (re, im)                         */
    const float* omega;         /* N*N float                                   */
    const float* twiddles;      /* N/2 float2 (cos, +sin)                      */
    float* spec;                /* F*N*N float2 scratch between the passes     */
    float* disp;                /* N*N float4                                  */
    float* normal;              /* N*N float4                                  */
    float* vel;                 /* N*N float4, or NULL when velocity is off    */
} ocean_hip_level;

/* Same layout as ocean_surface_ex. */
typedef struct ocean_hip_surface {
    float height, offset_x, offset_z, normal_x, normal_y, normal_z, foam;
    float velocity_x, velocity_y, velocity_z;
} ocean_hip_surface;

/* Per-step constants of the foam advection, all computed on the host exactly
 * as foam.cpp:94-100 computes them. */
typedef struct ocean_hip_foam_params {
    float dt, inv_cell, decay, gain, scale, drift_x, drift_z;
} ocean_hip_foam_params;

/* Enqueue one frame (row pass + column pass) of one level on stream s. */
hipError_t ocean_hip_frame(const ocean_hip_level* level, double time_over_two_pi,
                           hipStream_t s);
/* The same with a pass mask (1 = row pass, 2 = column pass), for timing the
 * two kernels separately. */
hipError_t ocean_hip_frame_passes(const ocean_hip_level* level, double time_over_two_pi,
                                  hipStream_t s, unsigned passes);

/* Batched sample_at over `count` (x, z) pairs (device float2). `out` and/or
 * `heights` may be NULL. */
hipError_t ocean_hip_sample(const ocean_hip_level* level, const float* xz, uint32_t count,
                            ocean_hip_surface* out, float* heights, hipStream_t s);

/* One semi-Lagrangian foam step: cur -> next (N*N floats each). */
hipError_t ocean_hip_foam_step(const ocean_hip_level* level, const float* cur, float* next,
                               const ocean_hip_foam_params* p, hipStream_t s);

/* Cascade query: per point, each level's own inversion and sample, combined
 * in level order. `levels` is a device array of `level_count` descriptors. */
hipError_t ocean_hip_cascade_sample(const ocean_hip_level* levels, uint32_t level_count,
                                    const float* xz, uint32_t count, ocean_hip_surface* out,
                                    float* heights, hipStream_t s);

/* Foam world query: the ocean's inversion, then a periodic bilinear tap of
 * the foam field at the solved parameter position (foam.cpp:291-305). */
hipError_t ocean_hip_foam_query(const ocean_hip_level* level, const float* foam, float cell,
                                const float* xz, uint32_t count, float* out, hipStream_t s);

/* Mean of n floats in double: `partial` holds ocean_hip_coverage_blocks(n)
 * doubles of scratch; the mean lands in *out (device). */
uint32_t ocean_hip_coverage_blocks(uint32_t n);
hipError_t ocean_hip_coverage(const float* v, uint32_t n, double* partial, double* out,
                              hipStream_t s);

#if defined(__cplusplus)
}
#endif

#endif
