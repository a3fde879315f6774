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
    uint32_t shape;             /* launch shape of the per-level passes: row
                                   variant | column variant << 8 (0 = default),
                                   | precision mode << 16 (0 = bit-exact,
                                   1 = fast, 2 = fast with fp16 fields)      */
    const float* h0;            /* N*N float2 (re, im)                         */
    const float* omega;         /* N*N float                                   */
    const float* twiddles;      /* N/2 float2 (cos, +sin)                      */
    float* spec;                /* F*N*N float2 scratch between the passes     */
    float* disp;                /* N*N float4                                  */
    float* normal;              /* N*N float4                                  */
    float* vel;                 /* N*N float4, or NULL when velocity is off    */
    const float* kinv;          /* N*N float: 1/|k| (0 at DC)                  */
    const float* kx;            /* N float: signed wavenumber of each column   */
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

/* ---- the constant tables on the device -----------------------------------
 * h0 (N*N float2), omega (N*N) and the twiddles (N/2 float2) by the device
 * libm; flags[i] = 1 marks a cell (twiddle) whose float results could differ
 * from the host libm's and must be recomputed on the host. */
typedef struct ocean_hip_spectrum_params {
    uint32_t n;
    float patch_length;
    uint64_t seed;
    double dk;
    double gravity, depth, omega_p, alpha, gamma, wind, swell, wind_direction, cutoff_l, mu, q_den;
} ocean_hip_spectrum_params;
hipError_t ocean_hip_spectrum(const ocean_hip_spectrum_params* p, float* h0, float* omega,
                              uint8_t* flags, hipStream_t s);
hipError_t ocean_hip_twiddles(uint32_t n, float* tw, uint8_t* flags, hipStream_t s);
/* Writes host-computed cells: `fixes` is a device array of count records
 * {uint32 index; float h0_re, h0_im, omega}. */
hipError_t ocean_hip_spectrum_fix(const void* fixes, uint32_t count, float* h0, float* omega,
                                  hipStream_t s);

/* The per-cell 1/|k| and per-column kx planes of an N grid (time-independent;
 * the row pass reads them instead of recomputing a double sqrt and divide per
 * cell), by the reference's expressions (src/core/spectrum.hpp:85-92,
 * src/core/spectrum.cpp:300-306). */
hipError_t ocean_hip_wavenumbers(uint32_t n, float patch_length, float* kinv, float* kx,
                                 hipStream_t s);

/* Enqueue one frame (row pass + column pass) of one level on stream s. */
hipError_t ocean_hip_frame(const ocean_hip_level* level, double time_over_two_pi,
                           hipStream_t s);
/* The same with a pass mask (1 = row pass, 2 = column pass), for timing the
 * two kernels separately. The level's `shape` picks the launch shapes. */
hipError_t ocean_hip_frame_passes(const ocean_hip_level* level, double time_over_two_pi,
                                  hipStream_t s, unsigned passes);

/* The same frame with each pass's dispatch recording its own start and end:
 * events[0..1] bracket the row pass, events[2..3] the column pass. */
hipError_t ocean_hip_frame_timed(const ocean_hip_level* level, double time_over_two_pi,
                                 hipStream_t s, hipEvent_t* events);

/* Launch-shape variants of the per-level passes for grid size n and F fields
 * (4, or 6 with velocity; variant 0 is the default): how many row and column
 * variants exist; -1 if n is not a supported size. */
int ocean_hip_frame_shapes(uint32_t n, uint32_t fields, int* row_variants, int* col_variants);

/* RGBA16F handoff (reference examples/viewer/half.hpp:58-130): cells RGBA32F
 * texels of src to 4*cells halves in dst, the alpha from alpha[cell] when
 * alpha is non-NULL; src 16-byte and dst 8-byte aligned. _half: count floats
 * to count halves (src 16-byte, dst 8-byte aligned). */
hipError_t ocean_hip_pack_rgba16(const float* src, const float* alpha, uint16_t* dst,
                                 uint32_t cells, hipStream_t s);
hipError_t ocean_hip_pack_half(const float* src, uint16_t* dst, uint64_t count, hipStream_t s);

/* Batched sample_at over `count` (x, z) pairs (device float2). `out` and/or
 * `heights` may be NULL. */
hipError_t ocean_hip_sample(const ocean_hip_level* level, const float* xz, uint32_t count,
                            ocean_hip_surface* out, float* heights, hipStream_t s);

/* One semi-Lagrangian foam step: cur -> next (N*N floats each). */
hipError_t ocean_hip_foam_step(const ocean_hip_level* level, const float* cur, float* next,
                               const ocean_hip_foam_params* p, hipStream_t s);
/* One foam step and `count` queries of the same frame (ocean_hip_sample's
 * arguments) in one launch; the same arithmetic as the two separate calls. */
hipError_t ocean_hip_foam_step_sample(const ocean_hip_level* level, const float* cur, float* next,
                                      const ocean_hip_foam_params* p, const float* xz,
                                      uint32_t count, ocean_hip_surface* out, float* heights,
                                      hipStream_t s);

/* Cascade query: per point, each level's own inversion and sample, combined
 * in level order. `levels` is a device array of `level_count` descriptors. */
hipError_t ocean_hip_cascade_sample(const ocean_hip_level* levels, uint32_t level_count,
                                    const float* xz, uint32_t count, ocean_hip_surface* out,
                                    float* heights, hipStream_t s);

/* Foam world query: the ocean's inversion, then a periodic bilinear tap of
 * the foam field at the solved parameter position (foam.cpp:291-305). */
hipError_t ocean_hip_foam_query(const ocean_hip_level* level, const float* foam, float cell,
                                const float* xz, uint32_t count, float* out, hipStream_t s);

/* Mean of n > 0 floats as the sequential double sum in index order over n
 * (foam.cpp:307-313), bit for bit; lands in *out (device). */
hipError_t ocean_hip_coverage(const float* v, uint32_t n, double* out, hipStream_t s);

/* ---- batched frames ------------------------------------------------------
 * Levels with N in [64, 2048] can be stepped together: every level is cut
 * into N*N/2048 work items (rows for the row pass, columns for the column
 * pass) of 256 lanes, and one persistent launch per pass strides a block
 * through the item list. Items are packed as (level << 12) | block. */
int ocean_hip_batchable(uint32_t n);
uint32_t ocean_hip_batch_items(uint32_t n);
uint32_t ocean_hip_batch_item(uint32_t level, uint32_t block);
/* One (N, F) group of a batch: its work items over the batch's level array. */
typedef struct ocean_hip_batch_group {
    uint32_t n;                    /* N of every level in the group            */
    uint32_t fields;               /* 4, or 6 with velocity                    */
    const ocean_hip_level* levels; /* device array of the whole batch's levels */
    const uint32_t* items;         /* device array of this group's items       */
    uint32_t n_items;
    uint32_t grid_row;             /* persistent grid sizes (autotuned)        */
    uint32_t grid_col;
    uint32_t mode;                 /* precision mode of the group's kernels    */
    uint32_t radix_row;            /* FFT points per lane of each pass: 8 (or  */
    uint32_t radix_col;            /* 0) or 4 (autotuned; each item in halves) */
    uint32_t layout;               /* fields between the passes: 0 row-major, */
                                   /* 1 tiles of an item's 2048/N columns     */
} ocean_hip_batch_group;
/* The opt-in fast modes' kernels (mode 1: frame_fast_kernels.hip, mode 2:
 * frame_half_kernels.hip): the per-level passes of a level whose shape
 * carries that mode, their variant counts, and the batched kernels of a
 * group of that mode (NULL if none). */
hipError_t ocean_hip_frame_fast_1(const ocean_hip_level* level, double time_over_two_pi,
                                  hipStream_t s, unsigned passes, hipEvent_t* events);
hipError_t ocean_hip_frame_fast_2(const ocean_hip_level* level, double time_over_two_pi,
                                  hipStream_t s, unsigned passes, hipEvent_t* events);
int ocean_hip_frame_shapes_fast_1(uint32_t n, uint32_t fields, int* row_variants, int* col_variants);
int ocean_hip_frame_shapes_fast_2(uint32_t n, uint32_t fields, int* row_variants, int* col_variants);
void* ocean_hip_batch_kernel_fast_1(uint32_t n, int fields, int row, int per_item, int radix);
void* ocean_hip_batch_kernel_fast_2(uint32_t n, int fields, int row, int per_item, int radix);
/* Variant counts of the per-level passes of one precision mode. */
int ocean_hip_frame_shapes_mode(uint32_t n, uint32_t fields, uint32_t mode, int* row_variants,
                                int* col_variants);
/* The same for a layout of the fields between the passes (a level's shape
 * bits 24-31): 0 row-major, 1 / 2 / 3 tiles of 2 / 4 / 8 columns x N rows
 * (bit-exact mode, N >= 256); -1 where the layout is not available. */
int ocean_hip_frame_shapes_layout(uint32_t n, uint32_t fields, uint32_t mode, uint32_t layout,
                                  int* row_variants, int* col_variants);
/* Direct launches of one group (passes: 1 row, 2 column, 3 both), for tuning. */
hipError_t ocean_hip_batch_group_frame(const ocean_hip_batch_group* group,
                                       double time_over_two_pi, unsigned passes, hipStream_t s);
/* Every group's frame as one instantiated hipGraph; launch sets the time. */
typedef struct ocean_hip_batch_graph ocean_hip_batch_graph;
hipError_t ocean_hip_batch_graph_create(const ocean_hip_batch_group* groups, uint32_t count,
                                        ocean_hip_batch_graph** out);
hipError_t ocean_hip_batch_graph_launch(ocean_hip_batch_graph* g, double time_over_two_pi,
                                        hipStream_t s);
void ocean_hip_batch_graph_destroy(ocean_hip_batch_graph* g);
/* Resident blocks per CU of a group's batched kernel (occupancy). */
int ocean_hip_batch_occupancy(uint32_t n, uint32_t fields, int row);
/* Whether the batched kernels of (n, fields, mode) exist with `radix` FFT
 * points per lane (8 always; 4 for N <= 1024) and `layout` (0 always; 1 in
 * bit-exact mode for N = 256 .. 1024). */
int ocean_hip_batch_has_launch(uint32_t n, uint32_t fields, uint32_t mode, uint32_t radix,
                               uint32_t layout);

/* Per level of a device array of `count` levels: the checksum of its output
 * buffers (displacement, normal, velocity words w_i; sum of w_i * K_i mod
 * 2^64, K_i = (i * 0x9E3779B97F4A7C15 + 0x632BE59BD9B4E019) | 1) into out[]. */
hipError_t ocean_hip_level_checksums(const ocean_hip_level* levels, uint32_t count,
                                     unsigned long long* out, hipStream_t s);

/* ---- converged queries and the learned start-point proposal --------------
 * Parameters of the proposal MLP, fp32, in this order: W1 [IN][HIDDEN]
 * (row = input), b1 [HIDDEN], W2 [HIDDEN][HIDDEN], b2, W3 [HIDDEN][OUT], b3
 * [OUT]. Inputs 60..63 and outputs 6..15 are padding (zero). */
#define OCEAN_HIP_MLP_LEVELS 3
#define OCEAN_HIP_MLP_IN 64
#define OCEAN_HIP_MLP_HIDDEN 64
#define OCEAN_HIP_MLP_OUT 16
#define OCEAN_HIP_MLP_PARAMS 9360
#define OCEAN_HIP_MLP_BF16 18432   /* W1^T, W2^T, W3^T, W1, W2, W3 in bf16 */
#define OCEAN_HIP_MLP_PROP 8       /* floats of proposal scratch per query: the
                                      start offsets, level-major float2 per
                                      (level, query); 6 used at 3 levels */
#define OCEAN_HIP_MLP_SPACING_PARTIALS 64  /* doubles of scratch per level */
#define OCEAN_HIP_MLP_COARSE_SIDE 512      /* largest coarse grid side kept in scratch */
/* Coarse-grid scratch: per level, side rows of side + 2 16-byte records (a
 * node column's three rows as bf16 pairs, the row wrapped one record either
 * side), at a fixed level stride. */
#define OCEAN_HIP_MLP_COARSE_RECORDS (OCEAN_HIP_MLP_COARSE_SIDE * (OCEAN_HIP_MLP_COARSE_SIDE + 2))
#define OCEAN_HIP_MLP_COARSE_BYTES (OCEAN_HIP_MLP_LEVELS * OCEAN_HIP_MLP_COARSE_RECORDS * 16u)
typedef struct ocean_hip_mlp {
    uint16_t* bf16;   /* the weights rounded to bf16, in the kernels' layouts */
    float* master;    /* fp32 parameters (biases are read from here)        */
    float* m1;        /* Adam first / second moments                        */
    float* m2;
    float* grad;      /* the last summed gradient                           */
} ocean_hip_mlp;

/* Stencil spacing per level from its displacement RMS (device int32[count]);
 * with `partial` (count * SPACING_PARTIALS doubles, count <= 256) the sum is
 * spread over 64 blocks per level, else one block per level. */
hipError_t ocean_hip_mlp_spacing(const ocean_hip_level* levels, uint32_t level_count,
                                 int32_t* spacing, double* partial, hipStream_t s);
/* The per-block partial sums of the levels' displacement RMS alone
 * (count * SPACING_PARTIALS doubles, count <= 256); ocean_hip_mlp_propose
 * adds them up. */
hipError_t ocean_hip_mlp_spacing_partial(const ocean_hip_level* levels, uint32_t level_count,
                                         double* partial, hipStream_t s);
/* Heights at `count` points (device float2), each level inverted until the
 * fixed-point step is <= tol metres (at most 64 taps), summed over levels.
 * mlp != NULL: start from the proposal (needs `spacing_partial` from
 * ocean_hip_mlp_spacing_partial; writes `spacing`, MLP_LEVELS int32; uses
 * `coarse` scratch of MLP_COARSE_BYTES and `prop` scratch of
 * count * MLP_PROP floats).
 * iters (count * level_count bytes) and flags (count bytes: 1 = a level did
 * not converge, 2 = a proposal was abandoned) may be NULL. */
hipError_t ocean_hip_height_converged(const ocean_hip_level* levels, uint32_t level_count,
                                      const float* xz, uint32_t count, float tol,
                                      const ocean_hip_mlp* mlp, const double* spacing_partial,
                                      int32_t* spacing, float* coarse, float* prop, float* heights,
                                      uint8_t* iters, uint8_t* flags, hipStream_t s);
/* The proposal alone: each of the first min(level_count, MLP_LEVELS) levels'
 * proposed start offset in cells, level-major float2 (prop[lv * count + q]),
 * with `spacing_partial`, `spacing` and `coarse` as for
 * ocean_hip_height_converged. */
hipError_t ocean_hip_mlp_propose(const ocean_hip_level* levels, uint32_t level_count, const float* xz,
                                 uint32_t count, const ocean_hip_mlp* mlp, const double* spacing_partial,
                                 int32_t* spacing, float* coarse, float* prop, hipStream_t s);
/* The scratch of the allocation-free converged query: views into a file-scope
 * device array of OCEAN_HIP_STATIC_SCRATCH_BYTES on the current device (its
 * address looked up once per device), shared by every call on that device. */
#define OCEAN_HIP_STATIC_MAX_QUERIES 262144u
#define OCEAN_HIP_STATIC_SCRATCH_BYTES \
    (256u + ((OCEAN_HIP_MLP_LEVELS * OCEAN_HIP_MLP_SPACING_PARTIALS * 8u + 255u) / 256u) * 256u + \
     OCEAN_HIP_MLP_COARSE_BYTES + \
     OCEAN_HIP_STATIC_MAX_QUERIES * OCEAN_HIP_MLP_PROP * 4u)
typedef struct ocean_hip_static_scratch_view {
    int32_t* spacing;          /* OCEAN_HIP_MLP_LEVELS                          */
    double* spacing_partial;   /* OCEAN_HIP_MLP_LEVELS * SPACING_PARTIALS      */
    float* coarse;             /* OCEAN_HIP_MLP_COARSE_BYTES                   */
    float* prop;               /* max_queries * OCEAN_HIP_MLP_PROP             */
    uint32_t max_queries;
} ocean_hip_static_scratch_view;
hipError_t ocean_hip_static_scratch(ocean_hip_static_scratch_view* out);
/* Training pairs for the first MLP_LEVELS levels: features (count * IN),
 * targets (count * OUT) from a plain solve to `tol`, ok[q] = 1 if every
 * level converged. */
hipError_t ocean_hip_mlp_pairs(const ocean_hip_level* levels, const float* xz, uint32_t count,
                               float tol, const int32_t* spacing, float* feats, float* targets,
                               uint8_t* ok, hipStream_t s);
hipError_t ocean_hip_mlp_forward(const ocean_hip_mlp* mlp, const float* feats, uint32_t rows,
                                 float* out, hipStream_t s);
/* bf16 <- master (after the master parameters were written). */
hipError_t ocean_hip_mlp_sync_bf16(ocean_hip_mlp* mlp, hipStream_t s);
/* One minibatch (rows[batch] index feats/targets; batch a multiple of 128):
 * gradient of the mean squared error over the 6 outputs into mlp->grad and,
 * if lr > 0, one Adam step (c1 = 1 - 0.9^t, c2 = 1 - 0.999^t). partial holds
 * (batch / 128) * PARAMS floats, loss_partial batch / 128 floats; *loss (may
 * be NULL) gets the sum of squared errors. */
uint32_t ocean_hip_mlp_train_blocks(uint32_t batch);
hipError_t ocean_hip_mlp_step(ocean_hip_mlp* mlp, const float* feats, const float* targets,
                              const uint32_t* rows, uint32_t batch, float lr, float c1, float c2,
                              float* partial, float* loss_partial, float* loss, hipStream_t s);
/* Raises the kernels' dynamic LDS limits once per process; 0 on success. */
int ocean_hip_mlp_configure(void);

/* ---- iWave interaction field -------------------------------------------- */

/* Device view of one interaction field. Planes have a P-cell zero halo;
 * `cur`/`old` point at interior cell (0,0) and rows are `stride` floats. */
typedef struct ocean_hip_iwave {
    uint32_t n;
    uint32_t p;
    uint32_t stride;
    uint32_t pad_;
    float* cur;            /* h^n                                         */
    float* old;            /* h^{n-1}; overwritten with h^{n+1} by a step */
    const float* taps;     /* (2P+1)^2, row-major j-outer                 */
    const float* c1;       /* N*N: 1/(1 + alpha dt)                       */
    const float* c2;       /* N*N: 1 - alpha dt                           */
    float* out;            /* N*N float4: eta, dEta/dx, dEta/dz, dEta/dt  */
    float gdt2;            /* g dt^2                                      */
    float dt;
    float inv_two_cell;    /* 0.5 / cell                                  */
    float inv_dt;          /* 1 / dt                                      */
    float cell;
    int32_t origin_x;      /* grid origin in whole cells                  */
    int32_t origin_z;
    int32_t pad2_;
} ocean_hip_iwave;

/* One source stamp of one substep, with the host-side scalars of the
 * reference's inject() (src/interaction.cpp:322-348) already evaluated. */
typedef struct ocean_hip_stamp {
    float gx, gz;          /* source position, cells                      */
    float inv2s2;          /* 1 / (2 sigma^2)                             */
    float disp;            /* -strength * scale                           */
    float vel;             /* velocity_y * scale                          */
    int32_t x0, x1, z0, z1;/* clipped stamp box (inclusive)               */
    int32_t pad_;
} ocean_hip_stamp;

/* Stamps `count` sources, in order, into cur/old (mean removal included).
 * `means` is device scratch of `count` floats. */
hipError_t ocean_hip_iwave_inject(const ocean_hip_iwave* f, const ocean_hip_stamp* stamps,
                                  uint32_t count, float* means, hipStream_t s);
/* The same with the stamps read from host memory at enqueue and passed inside
 * the kernel arguments: count <= ocean_hip_iwave_arg_stamps(), else
 * hipErrorInvalidValue and nothing is enqueued. */
uint32_t ocean_hip_iwave_arg_stamps(void);
hipError_t ocean_hip_iwave_inject_args(const ocean_hip_iwave* f, const ocean_hip_stamp* host_stamps,
                                       uint32_t count, float* means, hipStream_t s);
/* Solid cells: Dirichlet zero (dirichlet != 0) or copy from the nearest
 * fluid cell. `solid[k]` is the k-th solid cell's index and `source[k]` its
 * nearest fluid cell (0xFFFFFFFF: none, the cell is left as it is), for k <
 * *count (device memory, as written by ocean_hip_iwave_mask_tables); at most
 * `max_count` entries. */
hipError_t ocean_hip_iwave_obstruct(const ocean_hip_iwave* f, const uint32_t* solid,
                                    const uint32_t* source, const uint32_t* count,
                                    uint32_t max_count, int dirichlet, hipStream_t s);
/* The obstruction tables of an n x n mask (n a power of two in 32..1024;
 * mask != 0 is solid), reference src/interaction.cpp:255-292: near[i] = the
 * nearest fluid cell of cell i as the reference's BFS assigns it (the fluid
 * cell at the smallest L1 distance, the smallest index among equals;
 * 0xFFFFFFFF when there is no fluid cell), solid[] = the solid cells in index
 * order, source[k] = near[solid[k]], *count = their number. The mask is read
 * once from `src` (device memory or mapped host memory, 16-byte aligned) and
 * copied to `keep` (device memory) unless `keep` is NULL, in which case `src`
 * must be device memory. `work` is device scratch of n*n + n uint32; near,
 * solid, source hold n*n each. */
hipError_t ocean_hip_iwave_mask_tables(const uint8_t* src, uint8_t* keep, uint32_t n,
                                       uint32_t* work, uint32_t* near, uint32_t* solid,
                                       uint32_t* source, uint32_t* count, hipStream_t s);
/* dst <- src shifted by whole cells (dst(x, z) = src(x + dx, z + dz), 0 outside). */
hipError_t ocean_hip_iwave_mask_shift(const uint8_t* src, uint8_t* dst, uint32_t n, int32_t dx,
                                      int32_t dz, hipStream_t s);
/* The least-squares stencil fit's normal equations (reference
 * src/core/iwave_kernel.cpp:286-336): ata_lower[r (r+1)/2 + c] (c <= r) and
 * atb[r], each summed over the `samples` rows of psi (sample-major, mm values
 * each) in order, with per-sample weight w and symbol sym. mm <= MAX_BASIS. */
#define OCEAN_HIP_IWAVE_MAX_BASIS 90
hipError_t ocean_hip_iwave_fit_normal(const double* psi, const double* w, const double* sym,
                                      uint32_t samples, uint32_t mm, double* ata_lower,
                                      double* atb, hipStream_t s);
/* The realised symbol of (2 radius + 1)^2 `taps` with cell `dx` at every
 * point (zi, xi) of the (scan+1)^2 grid over [0, k_max]^2, zi-major, with
 * the device's double cosine: out[2 pt] = realised symbol, out[2 pt + 1] =
 * the true symbol k tanh(k depth) (k when depth <= 0) at |k|. */
hipError_t ocean_hip_iwave_symbol_scan(const float* taps, uint32_t radius, double dx, double k_max,
                                       double depth, uint32_t scan, double* out, hipStream_t s);
/* One leapfrog step: old <- c1 (2 cur - c2 old - g dt^2 (G * cur)). */
hipError_t ocean_hip_iwave_step(const ocean_hip_iwave* f, hipStream_t s);
/* energy() of the reference: `scratch` holds 2*N*N doubles; result in *out. */
hipError_t ocean_hip_iwave_energy(const ocean_hip_iwave* f, double gravity, double* scratch,
                                  double* out, hipStream_t s);
/* RGBA output from cur/old. */
hipError_t ocean_hip_iwave_finalize(const ocean_hip_iwave* f, hipStream_t s);
/* dst interior <- src interior shifted by whole cells (0 where uncovered). */
hipError_t ocean_hip_iwave_shift(const float* src, float* dst, uint32_t n, uint32_t stride,
                                 int32_t dx, int32_t dz, hipStream_t s);
/* Interaction-only queries (src/interaction.cpp:729-773). out: 4 floats each. */
hipError_t ocean_hip_iwave_sample(const ocean_hip_iwave* f, const float* xz, uint32_t count,
                                  float* out, hipStream_t s);
/* WaterSurface::sample_at (src/interaction.cpp:863-901): the FFT surface of
 * `level_count` levels (one ocean = 1 level; a cascade = its levels, combined
 * as a cascade) plus the field (may be NULL). */
hipError_t ocean_hip_combined_sample(const ocean_hip_level* levels, uint32_t level_count,
                                     int as_cascade, const ocean_hip_iwave* field,
                                     const float* xz, uint32_t count, ocean_hip_surface* out,
                                     hipStream_t s);

#if defined(__cplusplus)
}
#endif

#endif
