/* Additive C entry points of the MI355X build.
 *
 * Nothing here changes or replaces a reference symbol (include/ocean/ocean.h
 * is the drop-in ABI). These are the entry points the GPU design adds, as
 * additive symbols in the sense of the reference's own versioning rules
 * (reference include/ocean/ocean.h:68-74,207-215): batched queries, device
 * pointers, and C access to the reference's C++-only FoamField and
 * CascadeStack (reference include/ocean/foam.hpp:84-142,
 * include/ocean/cascade.hpp:45-92). Each declaration names the reference
 * interface it extends.
 *
 * Status returns: 0 on success, otherwise an ocean_status (ocean.h) or, for
 * a device failure, OCEAN_ERROR_UNKNOWN with the HIP error available from
 * ocean_amd_last_error().
 *
 * Threads: the entry points that take host arrays stage them through
 * buffers the ocean (or batch, or field) owns, so on one handle they must not
 * run concurrently with each other; device-array entry points only enqueue
 * on the handle's stream. As with the reference, nothing runs concurrently
 * with an update of the same handle.
 */
#ifndef OCEANLIB_OCEAN_AMD_H
#define OCEANLIB_OCEAN_AMD_H

#include "ocean/ocean.h"

#if defined(__cplusplus)
extern "C" {
#endif

/* ---- device-resident outputs -------------------------------------------
 * Extends ocean_get_buffers (reference include/ocean/ocean.h:140-171,
 * src/c_api.cpp:180-192): the same N*N RGBA32F layout, but device pointers,
 * valid until ocean_destroy or ocean_amd_bind_outputs. `stream` is the
 * hipStream_t the ocean enqueues its frames on; work queued on it after
 * ocean_update sees that frame. */
typedef struct ocean_amd_device_buffers {
    const float* displacement;
    const float* normal;
    const float* velocity;     /* NULL unless compute_velocity */
    uint32_t     size;
    float        patch_length;
    void*        stream;
} ocean_amd_device_buffers;

int  ocean_amd_get_device_buffers(const ocean_sim* sim, ocean_amd_device_buffers* out);
/* Zero-copy outputs for a renderer (no reference counterpart; the consumer
 * the reference's viewer feeds by copying, examples/viewer/ocean_view.cpp:
 * 818-870): the following frames' column pass writes displacement, normal
 * (and velocity) straight into caller-owned device memory, e.g. a Vulkan
 * buffer exported with VK_KHR_external_memory_fd and imported with
 * hipImportExternalMemory + hipExternalMemoryGetMappedBuffer; the queries,
 * the foam step, the packs, ocean_get_buffers and ocean_amd_get_device_buffers
 * then read them there. Each pointer: N*N float4, 16-byte aligned, inside one
 * allocation HIP knows; NULL keeps or restores the ocean's own buffer for that
 * output. Re-evaluates the current frame into the bound buffers (enqueued).
 * The caller keeps the memory alive until it rebinds or destroys the ocean.
 * INVALID_ARG (binding unchanged): a pointer that fails those checks,
 * velocity without compute_velocity, or a level of a batch or cascade. */
int  ocean_amd_bind_outputs(ocean_sim* sim, float* displacement_dev, float* normal_dev,
                            float* velocity_dev);
/* Blocks until every frame and query enqueued on the ocean has finished. */
void ocean_amd_synchronize(const ocean_sim* sim);
/* First HIP error latched since creation (hipError_t value), 0 if none.
 * update() is noexcept in the reference (include/ocean/ocean.hpp:198-202), so
 * device failures are latched here instead of thrown. */
int  ocean_amd_last_error(const ocean_sim* sim);

/* Device time of one frame's two kernels, measured with hipEvents on the
 * ocean's own stream over `iters` frames at times t0 + i*dt: mean ms of the
 * row pass (evolve + row FFT) and of the column pass (column FFT + pack).
 * Leaves the ocean at the last frame timed. Diagnostic; no reference
 * counterpart (the reference's stage timing is bench/bench_update.cpp:211-252). */
int ocean_amd_profile_frame(ocean_sim* sim, double t0, double dt, uint32_t iters,
                            float* row_ms, float* col_ms);

/* Timing of the frame kernels inside the caller's own update loop:
 * _frame_timing(sim, k): from now on the row and column dispatches of every
 * k-th ocean_update record HIP start/end events themselves
 * (hipExtLaunchKernel, no marker packets between kernels; k = 0 stops); _frame_timing_read
 * waits for the ocean's stream and returns the mean ms of each pass over the
 * recorded frames (the last 64 at most), then clears them. Diagnostic; no
 * reference counterpart. */
int ocean_amd_frame_timing(ocean_sim* sim, uint32_t every);
int ocean_amd_frame_timing_read(ocean_sim* sim, float* row_ms, float* col_ms, uint32_t* samples);

/* ---- launch shapes of the frame kernels ---------------------------------
 * No reference counterpart (the reference's work split is its task count,
 * src/ocean.cpp:117-161). The two frame kernels of one ocean can run with
 * several launch shapes (rows / columns per workgroup); every shape computes
 * the same bits. _frame_shapes: how many row and column variants exist for
 * this ocean's size and field count (variant 0 = default; the last ones
 * spread the fields over several teams per row / column). _set_frame_shape
 * selects one;
 * _autotune_frame times every variant of each pass on the ocean's own stream
 * inside reps whole frames (the pass's dispatches record their own events;
 * reps <= 256), keeps the fastest, leaves the ocean at its current time and
 * returns the choice and the times. */
int ocean_amd_frame_shapes(const ocean_sim* sim, uint32_t* row_variants, uint32_t* col_variants);
int ocean_amd_set_frame_shape(ocean_sim* sim, uint32_t row_variant, uint32_t col_variant);
int ocean_amd_autotune_frame(ocean_sim* sim, uint32_t reps, uint32_t* row_variant,
                             uint32_t* col_variant, float* row_ms, float* col_ms);
/* The layout of the complex fields the row pass hands to the column pass:
 * 0 = row-major (the default), 1 / 2 / 3 = tiles of 2 / 4 / 8 columns x N
 * rows, so a column pass's loads of adjacent columns over consecutive rows
 * are contiguous runs (bit-exact mode, N >= 256). Setting a layout selects its default shapes; the shape counts and
 * indices above refer to the current layout, and the autotuner tunes every
 * available layout and keeps the fastest. Results do not depend on it.
 * OCEAN_ERROR_INVALID_ARG for a layout the ocean cannot use. */
int ocean_amd_set_frame_layout(ocean_sim* sim, uint32_t layout);
uint32_t ocean_amd_get_frame_layout(const ocean_sim* sim);

/* ---- opt-in fast mode ----------------------------------------------------
 * No reference counterpart: the reference's frame is bit-exact only, and it
 * names a real-to-complex FFT as its speed-up path (ARCHITECTURE.md ADR-008;
 * the four complex fields here already carry eight real spectra). Never the
 * default. Modes:
 *   EXACT     : the reference's arithmetic bit for bit.
 *   FAST      : the phase's sine and cosine from the hardware's sin/cos of the
 *               phase in revolutions, fused (FMA) butterfly rotations, and the
 *               normal's 1/|n| from a hardware reciprocal square root.
 *   FAST_HALF : FAST with the complex fields between the row and the column
 *               pass stored as fp16 pairs (half the bytes between the passes).
 * _set_precision applies from the next update and returns the launch shape to
 * the mode's default (the shapes of a fast mode are its own list; see
 * _frame_shapes); OCEAN_ERROR_INVALID_ARG for another value or a batch's
 * level. Queries, foam and packs read whatever the frame wrote. */
#define OCEAN_AMD_PRECISION_EXACT 0u
#define OCEAN_AMD_PRECISION_FAST 1u
#define OCEAN_AMD_PRECISION_FAST_HALF 2u
int      ocean_amd_set_precision(ocean_sim* sim, uint32_t mode);
uint32_t ocean_amd_get_precision(const ocean_sim* sim);

/* ---- batched queries ----------------------------------------------------
 * Batched form of ocean_sample_ex / ocean_height_at (reference
 * src/c_api.cpp:199-202,330-350; src/ocean.cpp:375-452): result q is exactly
 * what ocean_sample_ex(sim, xz[2q], xz[2q+1]) returns.
 *   _batch        : host arrays; blocks until the results are back.
 *   _batch_device : device arrays; enqueued on the ocean's stream, returns
 *                   at once. `out` and/or `heights` may be NULL.          */
int ocean_amd_sample_batch(const ocean_sim* sim, const float* xz, uint32_t count,
                           ocean_surface_ex* out);
int ocean_amd_sample_batch_device(const ocean_sim* sim, const float* xz_device,
                                  uint32_t count, ocean_surface_ex* out_device,
                                  float* heights_device);

/* The ocean's constant tables as the frame kernels read them (copied from
 * the device): h0 (2*N*N floats), omega (N*N), and how many entries the
 * device build flagged and the host recomputed at creation. The device build
 * (reference src/core/spectrum.cpp:264-339, src/core/rng.cpp:7-17 on the
 * GPU) equals ocean_amd_build_spectrum bit for bit. Any pointer may be NULL. */
int ocean_amd_get_tables(const ocean_sim* sim, float* h0, float* omega, uint32_t* repaired);

/* ---- host-side tables (no device needed) -------------------------------
 * The constant tables the frame kernels read, as the library builds them:
 * h0 (2*N*N floats, re/im interleaved) and omega (N*N), reference
 * src/core/spectrum.cpp:264-339; twiddles (N/2 pairs cos, +sin), reference
 * src/core/fft.cpp:24-33. */
int ocean_amd_build_spectrum(const ocean_desc* desc, float* h0, float* omega);
int ocean_amd_build_twiddles(uint32_t n, float* tw);
/* The interaction field's (2P+1)^2 convolution taps as the library fits them
 * (reference src/core/iwave_kernel.cpp:177-437, least-squares method: the
 * normal equations and the realised-symbol scan on the current device) and
 * the largest realised symbol; returns the tap count or -1. */
int ocean_amd_build_iwave_taps(uint32_t radius, double cell, double depth, float* taps,
                               double* max_symbol);
/* The same fit with the scan's summary: sym[0..3] = max_symbol, min_symbol,
 * band_low, peak_relative_error (the reference's IWaveKernel fields). */
int ocean_amd_build_iwave_stencil(uint32_t radius, double cell, double depth, float* taps,
                                  double* sym);

/* ---- persistent advected foam ------------------------------------------
 * C access to the reference's C++ FoamField (include/ocean/foam.hpp:38-142,
 * src/foam.cpp). Same descriptor fields and defaults (FoamDesc), same
 * fixed-step accumulator; the field lives in device memory. */
typedef struct ocean_amd_foam ocean_amd_foam;

typedef struct ocean_amd_foam_desc {
    size_t   struct_size;
    float    decay;          /* 1/s                                        */
    float    source_gain;
    float    advect_scale;
    float    wind_drift_x;   /* m/s                                        */
    float    wind_drift_z;
    float    fixed_dt;       /* s                                          */
    uint32_t max_substeps;
} ocean_amd_foam_desc;

void            ocean_amd_foam_desc_init(ocean_amd_foam_desc* desc);
ocean_amd_foam* ocean_amd_foam_create(const ocean_sim* sim, const ocean_amd_foam_desc* desc,
                                      ocean_status* out_status);
void            ocean_amd_foam_destroy(ocean_amd_foam* foam);
void            ocean_amd_foam_update(ocean_amd_foam* foam, float dt);
/* ocean_amd_foam_update(foam, dt) then ocean_amd_sample_batch_device on the
 * foam's ocean, with the last foam substep and the queries in one launch: the
 * same results as the two calls, one kernel boundary fewer per frame. */
int             ocean_amd_foam_update_sample_device(ocean_amd_foam* foam, float dt,
                                                    const float* xz_device, uint32_t count,
                                                    ocean_surface_ex* out_device,
                                                    float* heights_device);
void            ocean_amd_foam_clear(ocean_amd_foam* foam);
/* N*N floats: host mirror (copied on request) / device pointer. */
const float*    ocean_amd_foam_data(const ocean_amd_foam* foam);
const float*    ocean_amd_foam_device_data(const ocean_amd_foam* foam);
uint32_t        ocean_amd_foam_last_substeps(const ocean_amd_foam* foam);
float           ocean_amd_foam_at(const ocean_amd_foam* foam, float world_x, float world_z);
double          ocean_amd_foam_coverage(const ocean_amd_foam* foam);

/* ---- cascades ------------------------------------------------------------
 * C access to the reference's C++ CascadeStack (include/ocean/cascade.hpp:
 * 45-92, src/cascade.cpp). update() enqueues every level on one stream;
 * queries sum levels in the reference's order (src/cascade.cpp:82-153). */
typedef struct ocean_amd_cascade ocean_amd_cascade;

ocean_amd_cascade* ocean_amd_cascade_create(const ocean_desc* levels, uint32_t count,
                                            ocean_status* out_status);
void          ocean_amd_cascade_destroy(ocean_amd_cascade* stack);
void          ocean_amd_cascade_update(ocean_amd_cascade* stack, double time);
uint32_t      ocean_amd_cascade_level_count(const ocean_amd_cascade* stack);
ocean_buffers ocean_amd_cascade_get_buffers(const ocean_amd_cascade* stack, uint32_t level);
void          ocean_amd_cascade_sample(const ocean_amd_cascade* stack, float world_x,
                                       float world_z, ocean_surface_ex* out);
int           ocean_amd_cascade_sample_batch(const ocean_amd_cascade* stack, const float* xz,
                                             uint32_t count, ocean_surface_ex* out);

/* ---- batches of seas ------------------------------------------------------
 * No reference counterpart: the reference steps each ocean (or cascade level)
 * on its own, one after another (src/cascade.cpp:42-55; src/c_api.cpp:148-152
 * per ocean_update call). A batch steps many independent levels of uneven N
 * with one persistent launch per (N, velocity) group and pass, replayed as a
 * hipGraph. Every level's buffers and queries equal those of an ocean_sim
 * created from the same desc and updated to the same time. */
typedef struct ocean_amd_batch ocean_amd_batch;

typedef struct ocean_amd_batch_group {
    uint32_t size;        /* N of the group's levels                      */
    uint32_t velocity;    /* 1 when the group carries velocity            */
    uint32_t levels;
    uint32_t items;       /* work items per pass                          */
    uint32_t grid_row;    /* persistent grid sizes in use                 */
    uint32_t grid_col;
    float    row_ms;      /* device time of each pass at the last autotune */
    float    col_ms;
    uint32_t radix_row;   /* FFT points per lane of each pass, as tuned:  */
    uint32_t radix_col;   /* 8, or 4 (stages in pairs, twice the lanes)   */
    uint32_t layout;      /* fields between the passes: 0 row-major, 1    */
                          /* tiles as wide as a work item's columns       */
} ocean_amd_batch_group;

ocean_amd_batch* ocean_amd_batch_create(const ocean_desc* levels, uint32_t count,
                                        ocean_status* out_status);
void          ocean_amd_batch_destroy(ocean_amd_batch* batch);
/* Every level to the same absolute time; enqueues and returns. */
void          ocean_amd_batch_update(ocean_amd_batch* batch, double time);
uint32_t      ocean_amd_batch_level_count(const ocean_amd_batch* batch);
/* One level as a read-only ocean_sim owned by the batch: every function of
 * ocean.h and of this header that takes a const ocean_sim* works on it
 * (buffers, velocity, queries, device buffers). Never pass it to
 * ocean_update or ocean_destroy. */
const ocean_sim* ocean_amd_batch_level(const ocean_amd_batch* batch, uint32_t level);
/* Queries over levels [first, first + level_count): one level answers as
 * ocean_sample_ex, several as a cascade of those levels. */
int           ocean_amd_batch_sample(const ocean_amd_batch* batch, uint32_t first,
                                     uint32_t level_count, const float* xz, uint32_t count,
                                     ocean_surface_ex* out);
int           ocean_amd_batch_sample_device(const ocean_amd_batch* batch, uint32_t first,
                                            uint32_t level_count, const float* xz_dev,
                                            uint32_t count, ocean_surface_ex* out_dev,
                                            float* heights_dev);
/* Times candidate grid sizes (reps launches each) and keeps the fastest. */
void          ocean_amd_batch_autotune(ocean_amd_batch* batch, uint32_t reps);
/* Precision mode of every level (as ocean_amd_set_precision), from the next
 * update; the grid sizes are kept. */
int           ocean_amd_batch_set_precision(ocean_amd_batch* batch, uint32_t mode);
uint32_t      ocean_amd_batch_group_count(const ocean_amd_batch* batch);
int           ocean_amd_batch_group_info(const ocean_amd_batch* batch, uint32_t group,
                                         ocean_amd_batch_group* out);
/* Group `group`'s launch choices by hand (what ocean_amd_batch_autotune
 * picks): the layout of the fields between the passes (0 row-major; 1 tiles
 * as wide as a work item's 2048/N columns, bit-exact mode, N = 256 .. 1024),
 * FFT points per lane of each pass (8, or 4 for N <= 1024) and each pass's
 * grid, 1 .. the group's items (twice that with 4 points per lane).
 * OCEAN_ERROR_INVALID_ARG, nothing changed, for a choice that does not exist.
 * Results do not depend on these choices. */
int           ocean_amd_batch_set_group_launch(ocean_amd_batch* batch, uint32_t group,
                                               uint32_t layout, uint32_t radix_row,
                                               uint32_t grid_row, uint32_t radix_col,
                                               uint32_t grid_col);
/* Checksums of levels [first, first + count) after the work queued so far
 * (blocks until done): per level its buffers (displacement, normal, then
 * velocity when on) as uint32 words w_i, i counted across the buffers, summed
 * as sum(w_i * K_i) mod 2^64 with K_i = (i * 0x9E3779B97F4A7C15 +
 * 0x632BE59BD9B4E019) | 1. For replay and lockstep checks between machines:
 * equal buffers give equal sums, and any one changed word changes the sum. */
int           ocean_amd_batch_checksums(const ocean_amd_batch* batch, uint32_t first,
                                        uint32_t count, uint64_t* out);
int           ocean_amd_batch_synchronize(const ocean_amd_batch* batch);
int           ocean_amd_batch_last_error(const ocean_amd_batch* batch);

/* ---- converged height queries and the learned start point ----------------
 * No reference counterpart. ocean_height_at takes the reference's four
 * fixed-point steps (src/ocean.cpp:375-400); these iterate the same map on
 * every level until the step is <= tol metres (at most 64 taps a level) and
 * return the sum of the levels' heights at their solved points, as a cascade
 * sums them (src/cascade.cpp:118-130). All arrays are device memory; the
 * work is enqueued on the batch's / cascade's stream. With a proposal, each
 * of the first three levels starts from the network's proposed point; the
 * plain iteration refines it to the same tolerance and restarts from the
 * query point when its step does not fall. iters (count * level_count bytes:
 * taps per level) and flags (count bytes: 1 = a level hit 64 taps, 2 = a
 * proposal was abandoned) may be NULL. Calls sharing a proposal are
 * serialised on its scratch. */
#define OCEAN_AMD_PROPOSAL_PARAMS 9360u
#define OCEAN_AMD_PROPOSAL_INPUTS 64u   /* 60 used */
#define OCEAN_AMD_PROPOSAL_OUTPUTS 16u  /* 6 used  */
#define OCEAN_AMD_PROPOSAL_MAX_BATCH 8192u

typedef struct ocean_amd_proposal ocean_amd_proposal;

/* params: OCEAN_AMD_PROPOSAL_PARAMS floats (W1 [64][64], b1 [64], W2 [64][64],
 * b2 [64], W3 [64][16], b3 [16]; rows are inputs), or NULL for zeros. */
ocean_amd_proposal* ocean_amd_proposal_create(const float* params, uint32_t max_queries,
                                              ocean_status* out_status);
void ocean_amd_proposal_destroy(ocean_amd_proposal* model);
int  ocean_amd_proposal_set_params(ocean_amd_proposal* model, const float* params);
int  ocean_amd_proposal_get_params(const ocean_amd_proposal* model, float* out);
/* rows x 64 features -> rows x 16 outputs (device). */
int  ocean_amd_proposal_forward(const ocean_amd_proposal* model, const float* feats_dev,
                                uint32_t rows, float* out_dev);
/* One Adam step on rows_dev[batch] (batch a multiple of 128, <= MAX_BATCH);
 * *loss (may be NULL) receives the minibatch's mean squared error. */
int  ocean_amd_proposal_train_step(ocean_amd_proposal* model, const float* feats_dev,
                                   const float* targets_dev, const uint32_t* rows_dev,
                                   uint32_t batch, float lr, float* loss);
/* The minibatch gradient (PARAMS floats, host) and loss, without a step. */
int  ocean_amd_proposal_gradient(ocean_amd_proposal* model, const float* feats_dev,
                                 const float* targets_dev, const uint32_t* rows_dev,
                                 uint32_t batch, float* grad, float* loss);
int  ocean_amd_proposal_synchronize(const ocean_amd_proposal* model);

int  ocean_amd_batch_height_converged(const ocean_amd_batch* batch, uint32_t first,
                                      uint32_t level_count, const float* xz_dev, uint32_t count,
                                      float tol, const ocean_amd_proposal* model,
                                      float* heights_dev, uint8_t* iters_dev, uint8_t* flags_dev);
/* No counterpart in the reference. For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): ocean_amd_batch_height_converged with the same arguments and
 * results, but the proposal's per-call scratch (the levels' stencil spacings,
 * their coarse grids and every query's start offsets) comes from a file-scope device array of
 * the library instead of the model's own, so the call allocates nothing and
 * the model's max_queries does not bound it. Limit: count <=
 * OCEAN_AMD_STATIC_MAX_QUERIES (larger counts return OCEAN_ERROR_INVALID_ARG;
 * split them into calls of at most that many). Each GPU holds one copy of the
 * array (about 21.0 MB) until the process ends, shared by every call on that
 * GPU: calls that overlap on one GPU (other streams, other threads) race on
 * it. Without a model no scratch is used. */
#define OCEAN_AMD_STATIC_MAX_QUERIES 262144u
int  ocean_amd_batch_height_converged_static(const ocean_amd_batch* batch, uint32_t first,
                                             uint32_t level_count, const float* xz_dev,
                                             uint32_t count, float tol,
                                             const ocean_amd_proposal* model, float* heights_dev,
                                             uint8_t* iters_dev, uint8_t* flags_dev);
int  ocean_amd_cascade_height_converged(const ocean_amd_cascade* stack, const float* xz_dev,
                                        uint32_t count, float tol, const ocean_amd_proposal* model,
                                        float* heights_dev, uint8_t* iters_dev,
                                        uint8_t* flags_dev);
/* The network's proposal alone, for batch levels [first, first + 3): each
 * level's proposed start offset from the query point in cells, level-major
 * (starts_dev: 3 * count float2, starts_dev[2 * (lv * count + q)] = du,
 * [+1] = dv), and the stencil spacings used (spacing_dev: 3 int32, may be
 * NULL), all device. count <= the model's max_queries. */
int  ocean_amd_proposal_starts(const ocean_amd_batch* batch, uint32_t first, const float* xz_dev,
                               uint32_t count, const ocean_amd_proposal* model, float* starts_dev,
                               int32_t* spacing_dev);
/* Training pairs from batch levels [first, first + 3): features (count x 64),
 * targets (count x 16: each level's solved start offset in stencil units,
 * plain iteration to tol), ok (count bytes), spacing (3 int32), all device. */
int  ocean_amd_proposal_pairs(const ocean_amd_batch* batch, uint32_t first, const float* xz_dev,
                              uint32_t count, float tol, float* feats_dev, float* targets_dev,
                              uint8_t* ok_dev, int32_t* spacing_dev);

/* ---- renderer handoff: RGBA16F textures -----------------------------------
 * The viewer's upload format (reference examples/viewer/half.hpp:58-130,
 * examples/viewer/ocean_view.cpp:818-870, ARCHITECTURE.md ADR-026), made on
 * the device: 4 halves per cell, the viewer's float_to_half (round to nearest
 * even, half subnormals, overflow to infinity, NaN to the quiet NaN
 * sign | 0x7E00). The displacement texture's alpha is the persistent foam of
 * `foam` (which must follow this ocean) when foam is non-NULL, else the
 * displacement's own w; the normal is converted as is. Either output may be
 * NULL; outputs are 8-byte aligned arrays of 4*N*N uint16.
 *   _device : device arrays (e.g. memory a renderer exported and HIP
 *             imported), enqueued on the ocean's stream; returns at once.
 *   host    : host arrays (e.g. a mapped staging buffer); blocks until
 *             written. Half the bytes of ocean_get_buffers.               */
int ocean_amd_pack_rgba16(const ocean_sim* sim, const ocean_amd_foam* foam,
                          uint16_t* displacement, uint16_t* normal);
int ocean_amd_pack_rgba16_device(const ocean_sim* sim, const ocean_amd_foam* foam,
                                 uint16_t* displacement_dev, uint16_t* normal_dev);
/* Any device array of `count` floats (16-byte aligned) to halves (8-byte
 * aligned), the same conversion, enqueued on the ocean's stream: for other
 * RGBA32F data a renderer uploads, e.g. ocean_get_velocity's. */
int ocean_amd_pack_half_device(const ocean_sim* sim, const float* src_dev, uint16_t* dst_dev,
                               uint64_t count);
/* The interaction field's RGBA buffer (ocean_interaction_get_buffers) the same
 * way, on the field's own stream (_device) or into host memory. */
int ocean_amd_interaction_pack_rgba16(const ocean_interaction* field, uint16_t* out);
int ocean_amd_interaction_pack_rgba16_device(const ocean_interaction* field, uint16_t* out_dev);
/* InteractionField::energy() of the C++ API (reference
 * src/interaction.cpp:803-846, which the reference's C API does not export):
 * the scheme's discrete energy, computed on the device in the reference's
 * summation order. NaN for a NULL field. */
double ocean_amd_interaction_energy(const ocean_interaction* field);
/* The obstruction tables the field applies (reference
 * src/interaction.cpp:255-292, which exports no accessor for them): the solid
 * cells in index order and each one's nearest fluid cell (0xFFFFFFFF: none),
 * built on the device. Copies at most `capacity` of each into the host arrays
 * (either may be NULL) and returns the number of solid cells; 0 for a NULL
 * field or one without a mask. */
uint32_t ocean_amd_interaction_obstruction_tables(const ocean_interaction* field, uint32_t* solid,
                                                  uint32_t* source, uint32_t capacity);

#if defined(__cplusplus)
}
#endif

#endif /* OCEANLIB_OCEAN_AMD_H */
