/* oceanlib C ABI, MI355X implementation.
 *
 * Drop-in replacement for the reference's C header (include/ocean/ocean.h:35-415
 * in the reference tree): every type below has the reference's field order,
 * sizes and alignment, and every function has its name, signature and
 * semantics, so a binary or binding built against the reference links and runs
 * against this library unchanged.
 *
 * What differs is underneath: the surface is synthesised by HIP kernels on one
 * GPU and stays resident in device memory. Pointers returned by
 * ocean_get_buffers() / ocean_get_velocity() point into a host mirror that is
 * refreshed from the device when they are requested, so their contents are the
 * surface at the time of the last ocean_update() as in the reference. Entry
 * points that the GPU design adds (batched queries, device pointers, batched
 * oceans) are additive and live in ocean/ocean_amd.h.
 *
 * Conventions carried over from the reference:
 *   - opaque handles; no exception crosses this boundary;
 *   - struct_size versioning: a smaller struct_size is an older caller's
 *     prefix and the rest takes defaults, 0 or a larger one is INVALID_ARG
 *     (reference src/c_api.cpp:46-60);
 *   - every entry point tolerates NULL handles (reference src/c_api.cpp:167-202).
 */
#ifndef OCEANLIB_OCEAN_H
#define OCEANLIB_OCEAN_H

#include <stddef.h>
#include <stdint.h>

#if defined(__cplusplus)
extern "C" {
#endif

/* ---- handle and status (reference include/ocean/ocean.h:35-48) ---------- */

typedef struct ocean_sim ocean_sim;

typedef enum ocean_status {
    OCEAN_OK                  = 0,
    OCEAN_ERROR_INVALID_ARG   = 1,
    OCEAN_ERROR_OUT_OF_MEMORY = 2,
    OCEAN_ERROR_UNKNOWN       = 3
} ocean_status;

/* Static string for a status; never NULL. */
const char* ocean_status_string(ocean_status status);

/* ---- host scheduler hook (reference include/ocean/ocean.h:54-66) --------
 * Accepted for ABI compatibility. The GPU path has no host-side stages to
 * distribute, so a supplied hook is stored and never invoked. */
typedef void (*ocean_task_fn)(void* ctx, uint32_t index);
typedef void (*ocean_parallel_for_fn)(void* user, ocean_task_fn task, void* ctx,
                                      uint32_t count);

/* ---- configuration (reference include/ocean/ocean.h:68-134) -------------
 * ocean_spectrum_desc is embedded by value before later ocean_desc fields, so
 * it is frozen; new ocean_desc fields are only ever appended at its end. */
typedef struct ocean_spectrum_desc {
    float wind_speed;        /* U10, m/s                                   */
    float fetch;             /* m of open water under the wind             */
    float wind_direction;    /* radians, 0 = +X                            */
    float peak_enhancement;  /* JONSWAP gamma (1 = Pierson-Moskowitz)      */
    float swell;             /* 0 wind sea .. 1 narrow swell               */
    float small_wave_cutoff; /* m; shorter waves are rolled off            */
    float gravity;           /* m/s^2                                      */
} ocean_spectrum_desc;

typedef struct ocean_desc {
    size_t struct_size;      /* = sizeof(ocean_desc); the version tag       */

    uint32_t size;           /* N, power of two in [16, 4096]               */
    float    patch_length;   /* tile edge, m                                */

    ocean_spectrum_desc spectrum;

    float    choppiness;     /* lambda                                      */
    float    foam_threshold; /* Jacobian below which foam appears           */
    uint64_t seed;

    ocean_parallel_for_fn parallel_for;
    void*                 parallel_for_user;
    uint32_t              thread_count;

    float   water_depth;      /* <= 0: deep water                           */
    int32_t compute_velocity; /* non-zero: also synthesise orbital velocity */
} ocean_desc;

/* Defaults identical to the reference's C++ OceanDesc; sets struct_size. */
void ocean_desc_init(ocean_desc* desc);

/* ---- outputs (reference include/ocean/ocean.h:140-171) -------------------
 * N*N RGBA32F texels, row-major i = z*N + x:
 *   displacement = (dx, height, dz, foam), normal = (nx, ny, nz, jacobian). */
typedef struct ocean_buffers {
    const float* displacement;
    const float* normal;
    uint32_t     size;
    float        patch_length;
} ocean_buffers;

typedef struct ocean_surface {
    float height;
    float offset_x;
    float offset_z;
    float normal_x;
    float normal_y;
    float normal_z;
    float foam;
} ocean_surface;

/* ---- lifetime (reference include/ocean/ocean.h:173-191) ------------------ */

/* Validates, builds the spectrum, allocates all device memory and evaluates
 * t = 0. NULL plus *out_status on failure. */
ocean_sim* ocean_create(const ocean_desc* desc, ocean_status* out_status);
void       ocean_destroy(ocean_sim* sim);

/* ---- simulation (reference include/ocean/ocean.h:193-203) ----------------
 * Absolute time; the surface is a pure function of (seed, desc, time). The
 * frame is enqueued on the ocean's stream; allocation-free. */
void          ocean_update(ocean_sim* sim, double time);
ocean_buffers ocean_get_buffers(const ocean_sim* sim);
double        ocean_get_time(const ocean_sim* sim);

/* ---- queries (reference include/ocean/ocean.h:205-253) ------------------- */
float         ocean_height_at(const ocean_sim* sim, float world_x, float world_z);
ocean_surface ocean_sample_at(const ocean_sim* sim, float world_x, float world_z);

typedef struct ocean_surface_ex {
    float height;
    float offset_x;
    float offset_z;
    float normal_x;
    float normal_y;
    float normal_z;
    float foam;
    float velocity_x;
    float velocity_y;
    float velocity_z;
} ocean_surface_ex;

void         ocean_sample_ex(const ocean_sim* sim, float world_x, float world_z,
                             ocean_surface_ex* out);
const float* ocean_get_velocity(const ocean_sim* sim);

/* ---- introspection (reference include/ocean/ocean.h:255-280) -------------
 * There are no CPU SIMD levels on this path: ocean_simd_level() names the GPU
 * kernel set ("gfx950"), and ocean_force_simd_level() accepts the reference's
 * names, changes nothing and returns the same string. */
const char* ocean_simd_level(void);
const char* ocean_force_simd_level(const char* name);
void        ocean_version(uint32_t* major, uint32_t* minor, uint32_t* patch);

/* ---- local interaction (reference include/ocean/ocean.h:282-415) -------- */

typedef struct ocean_interaction ocean_interaction;

typedef enum ocean_obstruction {
    OCEAN_OBSTRUCTION_NEUMANN   = 0,
    OCEAN_OBSTRUCTION_DIRICHLET = 1
} ocean_obstruction;

typedef enum ocean_source_kind {
    OCEAN_SOURCE_IMPULSE    = 0,
    OCEAN_SOURCE_CONTINUOUS = 1
} ocean_source_kind;

typedef struct ocean_interaction_desc {
    size_t struct_size;

    uint32_t size;          /* N, power of two in [32, 1024]               */
    float    extent;        /* world edge length, m                        */
    uint32_t kernel_radius; /* P; (2P+1)^2 taps                            */
    float    damping;       /* 1/s                                         */
    uint32_t absorb_cells;
    float    fixed_dt;      /* 0 = derived default                         */
    uint32_t max_substeps;
    uint32_t max_sources;
    float    gravity;
    float    water_depth;
    int32_t  obstruction;   /* ocean_obstruction                           */

    ocean_parallel_for_fn parallel_for;
    void*                 parallel_for_user;
    uint32_t              thread_count;
} ocean_interaction_desc;

void ocean_interaction_desc_init(ocean_interaction_desc* desc);

ocean_interaction* ocean_interaction_create(const ocean_interaction_desc* desc,
                                            ocean_status* out_status);
void ocean_interaction_destroy(ocean_interaction* field);

typedef struct ocean_disturbance {
    float world_x;
    float world_z;
    float radius;
    float strength;
    float velocity_y;
    float velocity_x;
    float velocity_z;
    int32_t kind;           /* ocean_source_kind */
} ocean_disturbance;

void     ocean_interaction_add(ocean_interaction* field, const ocean_disturbance* d);
uint64_t ocean_interaction_dropped(const ocean_interaction* field);
void     ocean_interaction_update(ocean_interaction* field, float dt);
int32_t  ocean_interaction_recenter(ocean_interaction* field, float world_x,
                                    float world_z);
void     ocean_interaction_set_obstruction(ocean_interaction* field,
                                           const uint8_t* mask);

typedef struct ocean_interaction_buffers {
    const float* field;     /* N*N*4: eta, dEta/dx, dEta/dz, dEta/dt; not periodic */
    uint32_t     size;
    float        extent;
    float        origin_x;
    float        origin_z;
} ocean_interaction_buffers;

ocean_interaction_buffers ocean_interaction_get_buffers(const ocean_interaction* field);

typedef struct ocean_interaction_sample {
    float height;
    float slope_x;
    float slope_z;
    float velocity_y;
} ocean_interaction_sample;

void  ocean_interaction_sample_at(const ocean_interaction* field, float world_x,
                                  float world_z, ocean_interaction_sample* out);
float ocean_interaction_height_at(const ocean_interaction* field, float world_x,
                                  float world_z);
float ocean_interaction_fixed_dt(const ocean_interaction* field);
float ocean_interaction_dt_limit(const ocean_interaction* field);
float ocean_interaction_dispersion_error(const ocean_interaction* field);

void ocean_sample_combined(const ocean_sim* sim, const ocean_interaction* field,
                           float world_x, float world_z, ocean_surface_ex* out);

#if defined(__cplusplus)
}
#endif

#endif /* OCEANLIB_OCEAN_H */
