/* CPU oracle for the iWave interaction field: plain-C restatement of the
 * reference's InteractionField (src/interaction.cpp), its kernel builder
 * (src/core/iwave_kernel.cpp) and step/finalize (src/core/iwave_step.cpp).
 * TEST INFRASTRUCTURE ONLY - see ocean_oracle.h. */
#include "ocean_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

static const double kPi = 3.14159265358979323846;

/* ---------------- kernel builder: iwave_kernel.cpp:72-437 ---------------- */

/* iwave_kernel.cpp:72-77 */
static int orbit_size(int a, int b)
{
    if (a == 0) return 1;
    if (b == 0) return 4;
    if (a == b) return 4;
    return 8;
}

/* iwave_kernel.cpp:79-108 */
static int orbit_of(int a, int b, int (*out)[2])
{
    int n = 0;
    if (a == 0) {
        out[n][0] = 0; out[n][1] = 0; ++n;
        return n;
    }
    if (b == 0) {
        out[n][0] = a; out[n][1] = 0; ++n;
        out[n][0] = -a; out[n][1] = 0; ++n;
        out[n][0] = 0; out[n][1] = a; ++n;
        out[n][0] = 0; out[n][1] = -a; ++n;
        return n;
    }
    const int sx[2] = {a, -a};
    const int sz[2] = {b, -b};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) { out[n][0] = sx[i]; out[n][1] = sz[j]; ++n; }
    if (a != b)
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) { out[n][0] = sz[j]; out[n][1] = sx[i]; ++n; }
    return n;
}

/* iwave_kernel.cpp:110-150 */
static int cholesky_solve(double* a, double* b, int n)
{
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j <= i; ++j) {
            double s = a[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) s -= a[(size_t)i * n + k] * a[(size_t)j * n + k];
            if (i == j) {
                if (!(s > 0.0)) return 0;
                a[(size_t)i * n + i] = sqrt(s);
            } else {
                a[(size_t)i * n + j] = s / a[(size_t)j * n + j];
            }
        }
    }
    for (int i = 0; i < n; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= a[(size_t)i * n + k] * b[k];
        b[i] = s / a[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int k = i + 1; k < n; ++k) s -= a[(size_t)k * n + i] * b[k];
        b[i] = s / a[(size_t)i * n + i];
    }
    return 1;
}

/* iwave_kernel.cpp:154-159 */
static double iwave_symbol(double k, double depth)
{
    if (k <= 0.0) return 0.0;
    if (depth <= 0.0) return k;
    return k * tanh(k * depth);
}

typedef struct {
    int radius, width;
    float cell;
    float* taps;
    double max_symbol, min_symbol, band_low, peak_relative_error;
} kernel_t;

/* iwave_kernel.cpp:161-173 (realised symbol, with the float-stored cell) */
static double realised_symbol(const kernel_t* k, double kx, double kz)
{
    const int p = k->radius;
    const double dx = k->cell;
    double sum = 0.0;
    for (int j = -p; j <= p; ++j)
        for (int i = -p; i <= p; ++i)
            sum += (double)k->taps[(size_t)(j + p) * k->width + (i + p)] *
                   cos(kx * i * dx + kz * j * dx);
    return sum;
}

/* iwave_kernel.cpp:177-437, LeastSquares method, fit_samples 96, band 1/8. */
static int build_kernel(int radius, double cell, double depth, kernel_t* out)
{
    const int fit_samples = 96;
    const double band_low_frac = 0.125;
    const int p = radius;
    const int width = 2 * p + 1;
    const double k_max = kPi / cell;
    out->radius = radius;
    out->width = width;
    out->cell = (float)cell;
    out->taps = calloc((size_t)width * width, sizeof(float));

    int ai[200], bi[200], m = 0;
    for (int a = 0; a <= p; ++a)
        for (int b = 0; b <= a; ++b) { ai[m] = a; bi[m] = b; ++m; }
    double g[200];
    memset(g, 0, sizeof(g));

    const int mm = m - 1;
    double* ata = calloc((size_t)mm * mm, sizeof(double));
    double* atb = calloc((size_t)mm, sizeof(double));
    double* psi = calloc((size_t)mm, sizeof(double));
    const double k_lo = band_low_frac * k_max;
    const double kOutOfBandWeight = 0.02;
    const int ns = fit_samples;
    int orb[8][2];
    for (int zi = 0; zi <= ns; ++zi) {
        for (int xi = 0; xi <= ns; ++xi) {
            const double kx = k_max * xi / ns;
            const double kz = k_max * zi / ns;
            const double kk = sqrt(kx * kx + kz * kz);
            const double s = iwave_symbol(kk, depth);
            const int in_band = (kk >= k_lo && kk <= k_max);
            const double ref = in_band ? s : iwave_symbol((kk < k_lo) ? k_lo : k_max, depth);
            double w = 1.0 / (ref * ref);
            if (!in_band) w *= kOutOfBandWeight;
            for (int t = 1; t < m; ++t) {
                const int n = orbit_of(ai[t], bi[t], orb);
                double phi = 0.0;
                for (int e = 0; e < n; ++e) phi += cos(kx * orb[e][0] * cell + kz * orb[e][1] * cell);
                psi[t - 1] = phi - n;
            }
            for (int r = 0; r < mm; ++r) {
                const double wr = w * psi[r];
                atb[r] += wr * s;
                for (int c = 0; c <= r; ++c) ata[(size_t)r * mm + c] += wr * psi[c];
            }
        }
    }
    double trace = 0.0;
    for (int r = 0; r < mm; ++r) trace += ata[(size_t)r * mm + r];
    const double ridge = 1e-12 * trace / mm;
    for (int r = 0; r < mm; ++r) {
        ata[(size_t)r * mm + r] += ridge;
        for (int c = r + 1; c < mm; ++c) ata[(size_t)r * mm + c] = ata[(size_t)c * mm + r];
    }
    const int ok = cholesky_solve(ata, atb, mm);
    if (ok) {
        double sum_rest = 0.0;
        for (int t = 1; t < m; ++t) {
            g[t] = atb[t - 1];
            sum_rest += orbit_size(ai[t], bi[t]) * g[t];
        }
        g[0] = -sum_rest;
    }
    free(ata); free(atb); free(psi);
    if (!ok) return 0;

    for (int t = 0; t < m; ++t) {
        const int n = orbit_of(ai[t], bi[t], orb);
        const float v = (float)g[t];
        for (int e = 0; e < n; ++e)
            out->taps[(size_t)(orb[e][1] + p) * width + (orb[e][0] + p)] = v;
    }

    const int kScan = 192;
    double worst = 0.0, peak = 0.0, trough = 0.0;
    out->band_low = band_low_frac * k_max;
    for (int zi = 0; zi <= kScan; ++zi) {
        for (int xi = 0; xi <= kScan; ++xi) {
            const double kx = k_max * xi / kScan;
            const double kz = k_max * zi / kScan;
            const double kk = sqrt(kx * kx + kz * kz);
            const double rs = realised_symbol(out, kx, kz);
            if (rs > peak) peak = rs;
            if (rs < trough) trough = rs;
            if (kk >= out->band_low && kk <= k_max) {
                const double tr = iwave_symbol(kk, depth);
                if (tr > 0.0) {
                    const double e = fabs(rs / tr - 1.0);
                    if (e > worst) worst = e;
                }
            }
        }
    }
    out->max_symbol = peak;
    out->min_symbol = trough;
    out->peak_relative_error = worst;
    return 1;
}

EXPORT int orc_iwave_taps(uint32_t radius, double cell, double depth, float* taps,
                          double* max_symbol)
{
    kernel_t k;
    if (!build_kernel((int)radius, cell, depth, &k)) return -1;
    const int w = 2 * (int)radius + 1;
    memcpy(taps, k.taps, sizeof(float) * (size_t)w * w);
    if (max_symbol) *max_symbol = k.max_symbol;
    free(k.taps);
    return w * w;
}

/* The fit's taps and its scan summary: sym[0..3] = max_symbol, min_symbol,
 * band_low, peak_relative_error (iwave_kernel.cpp:400-436). */
EXPORT int orc_iwave_stencil(uint32_t radius, double cell, double depth, float* taps, double* sym)
{
    kernel_t k;
    if (!build_kernel((int)radius, cell, depth, &k)) return -1;
    const int w = 2 * (int)radius + 1;
    memcpy(taps, k.taps, sizeof(float) * (size_t)w * w);
    sym[0] = k.max_symbol;
    sym[1] = k.min_symbol;
    sym[2] = k.band_low;
    sym[3] = k.peak_relative_error;
    free(k.taps);
    return w * w;
}

/* The realised symbol of `taps` (radius, float-stored cell) at every point
 * (zi, xi) of the (scan+1)^2 grid over [0, pi/cell]^2, zi-major, as the
 * scan of iwave_kernel.cpp:408-432 evaluates it. */
EXPORT void orc_iwave_symbol_scan(const float* taps, uint32_t radius, float cell, double k_max,
                                  int scan, double* out)
{
    kernel_t k;
    k.radius = (int)radius;
    k.width = 2 * (int)radius + 1;
    k.cell = cell;
    k.taps = (float*)taps;
    for (int zi = 0; zi <= scan; ++zi)
        for (int xi = 0; xi <= scan; ++xi) {
            const double kx = k_max * xi / scan;
            const double kz = k_max * zi / scan;
            out[(size_t)zi * (scan + 1) + xi] = realised_symbol(&k, kx, kz);
        }
}

/* ---------------- the field: interaction.cpp:71-908 ---------------------- */

typedef struct {
    ocean_interaction_desc desc;
    kernel_t kernel;
    uint32_t n;
    size_t stride, cells;
    float cell, dt, dt_limit;
    float *buf_a, *buf_b, *cur, *old;
    float *c1, *c2, *out;
    uint8_t* mask;
    uint32_t *nearest, *solid, *bfs;
    uint32_t solid_count;
    int has_mask;
    ocean_disturbance* queue;
    uint32_t queue_count;
    uint64_t dropped;
    double accum;
    uint32_t substeps_last;
    int clamped_last;
    int32_t origin_x, origin_z, shift_x, shift_z;
    int moved;
} orc_iw;

/* interaction.cpp:71-75 */
static float ricker(float r2, float inv_two_sigma2)
{
    const float u = r2 * inv_two_sigma2;
    return (1.0f - u) * expf(-u);
}

static float* base_of(orc_iw* m, float* b)
{
    return b + (size_t)m->desc.kernel_radius * m->stride + m->desc.kernel_radius;
}

/* interaction.cpp:212-241 */
static void build_damping(orc_iw* m)
{
    const float a0 = m->desc.damping;
    const int w = (int)m->desc.absorb_cells;
    const float denom = m->dt * (float)(w > 1 ? w : 1);
    const float a_edge = 4.0f / (denom > 1e-6f ? denom : 1e-6f);
    const uint32_t n = m->n;
    for (uint32_t z = 0; z < n; ++z) {
        for (uint32_t x = 0; x < n; ++x) {
            const int dx = (int)x < (int)(n - 1 - x) ? (int)x : (int)(n - 1 - x);
            const int dz = (int)z < (int)(n - 1 - z) ? (int)z : (int)(n - 1 - z);
            const int d = dx < dz ? dx : dz;
            float alpha = a0;
            if (w > 0 && d < w) {
                const float t = (float)(w - d) / (float)w;
                alpha += a_edge * t * t;
            }
            const float ad = alpha * m->dt;
            const size_t i = (size_t)z * n + x;
            m->c1[i] = 1.0f / (1.0f + ad);
            m->c2[i] = 1.0f - ad;
        }
    }
}

/* interaction.cpp:255-292: a multi-source BFS from every fluid cell
 * (mask == 0) in index order; nearest[i] = the fluid cell that claims cell i
 * (0xFFFFFFFF when the mask has no fluid cell). `queue` holds n*n entries. */
static void bfs_nearest(const uint8_t* mask, uint32_t n, uint32_t* nearest, uint32_t* queue)
{
    const size_t cells = (size_t)n * n;
    uint32_t head = 0, tail = 0;
    for (size_t i = 0; i < cells; ++i) {
        if (mask[i] == 0) {
            nearest[i] = (uint32_t)i;
            queue[tail++] = (uint32_t)i;
        } else {
            nearest[i] = 0xFFFFFFFFu;
        }
    }
    while (head < tail) {
        const uint32_t i = queue[head++];
        const uint32_t x = i % n, z = i / n;
        const int ddx[4] = {1, -1, 0, 0};
        const int ddz[4] = {0, 0, 1, -1};
        for (int e = 0; e < 4; ++e) {
            const int nx = (int)x + ddx[e], nz = (int)z + ddz[e];
            if (nx < 0 || nz < 0 || nx >= (int)n || nz >= (int)n) continue;
            const uint32_t j = (uint32_t)nz * n + (uint32_t)nx;
            if (nearest[j] == 0xFFFFFFFFu) {
                nearest[j] = nearest[i];
                queue[tail++] = j;
            }
        }
    }
}

static void rebuild_mask_tables(orc_iw* m)
{
    m->solid_count = 0;
    if (!m->has_mask) return;
    for (size_t i = 0; i < m->cells; ++i)
        if (m->mask[i] != 0) m->solid[m->solid_count++] = (uint32_t)i;
    bfs_nearest(m->mask, m->n, m->nearest, m->bfs);
}

/* The reference's nearest-fluid table of a bare n x n mask. */
EXPORT void orc_nearest_fluid(const uint8_t* mask, uint32_t n, uint32_t* nearest)
{
    uint32_t* queue = malloc(sizeof(uint32_t) * (size_t)n * n);
    bfs_nearest(mask, n, nearest, queue);
    free(queue);
}

/* interaction.cpp:295-318 */
static void apply_obstruction(orc_iw* m)
{
    if (!m->has_mask || m->solid_count == 0) return;
    const uint32_t n = m->n;
    for (uint32_t s = 0; s < m->solid_count; ++s) {
        const uint32_t i = m->solid[s];
        if (m->desc.obstruction == OCEAN_OBSTRUCTION_DIRICHLET) {
            m->cur[(i / n) * m->stride + (i % n)] = 0.0f;
        } else {
            const uint32_t f = m->nearest[i];
            if (f == 0xFFFFFFFFu) continue;
            m->cur[(i / n) * m->stride + (i % n)] = m->cur[(f / n) * m->stride + (f % n)];
        }
    }
}

/* interaction.cpp:322-391 */
static void inject(orc_iw* m, const ocean_disturbance* d, float sub_time, float scale)
{
    const float cell = m->cell;
    const float sigma = d->radius > 0.25f * cell ? d->radius : 0.25f * cell;
    const float inv2s2 = 1.0f / (2.0f * sigma * sigma);
    const float wx = d->world_x + d->velocity_x * sub_time;
    const float wz = d->world_z + d->velocity_z * sub_time;
    const float gx = (wx - (float)m->origin_x * cell) / cell;
    const float gz = (wz - (float)m->origin_z * cell) / cell;
    const int rad = (int)ceilf(4.0f * sigma / cell);
    const int n = (int)m->n;
    int x0 = (int)floorf(gx) - rad; if (x0 < 0) x0 = 0;
    int x1 = (int)floorf(gx) + rad; if (x1 > n - 1) x1 = n - 1;
    int z0 = (int)floorf(gz) - rad; if (z0 < 0) z0 = 0;
    int z1 = (int)floorf(gz) + rad; if (z1 > n - 1) z1 = n - 1;
    if (x1 < x0 || z1 < z0) return;
    double shape_sum = 0.0;
    for (int z = z0; z <= z1; ++z) {
        for (int x = x0; x <= x1; ++x) {
            const float px = ((float)x + 0.5f) - gx;
            const float pz = ((float)z + 0.5f) - gz;
            shape_sum += ricker((px * px + pz * pz) * cell * cell, inv2s2);
        }
    }
    const float shape_mean =
        (float)(shape_sum / ((double)(x1 - x0 + 1) * (double)(z1 - z0 + 1)));
    const float disp = -d->strength * scale;
    const float vel = d->velocity_y * scale;
    for (int z = z0; z <= z1; ++z) {
        for (int x = x0; x <= x1; ++x) {
            const float px = ((float)x + 0.5f) - gx;
            const float pz = ((float)z + 0.5f) - gz;
            const float r2 = (px * px + pz * pz) * cell * cell;
            const float s = ricker(r2, inv2s2) - shape_mean;
            const float dd = disp * s;
            const float dv = vel * s;
            const size_t ci = (size_t)z * m->stride + x;
            m->cur[ci] += dd;
            m->old[ci] += dd - dv * m->dt;
        }
    }
}

/* iwave_step.cpp:7-54 */
static void step(orc_iw* m)
{
    const int p = (int)m->desc.kernel_radius;
    const int w = 2 * p + 1;
    const size_t stride = m->stride;
    const uint32_t n = m->n;
    const float gdt2 = m->desc.gravity * m->dt * m->dt;
    for (uint32_t z = 0; z < n; ++z) {
        const size_t row = (size_t)z * stride;
        const size_t orow = (size_t)z * n;
        for (uint32_t x = 0; x < n; ++x) {
            const float* centre = m->cur + row + x;
            float acc = 0.0f;
            const float* tap = m->kernel.taps;
            for (int j = -p; j <= p; ++j) {
                const float* srow = centre + (ptrdiff_t)j * (ptrdiff_t)stride - p;
                for (int i = 0; i < w; ++i) acc += tap[i] * srow[i];
                tap += w;
            }
            const float hc = *centre;
            const float hp = m->old[row + x];
            const float v = m->c1[orow + x] * (2.0f * hc - m->c2[orow + x] * hp - gdt2 * acc);
            m->old[row + x] = (v > -1.0e-30f && v < 1.0e-30f) ? 0.0f : v;
        }
    }
}

/* iwave_step.cpp:93-124 */
static void finalize(orc_iw* m)
{
    const float inv_two_cell = 0.5f / m->cell;
    const float inv_dt = 1.0f / m->dt;
    const size_t stride = m->stride;
    for (uint32_t z = 0; z < m->n; ++z) {
        const size_t row = (size_t)z * stride;
        float* dst = m->out + (size_t)z * m->n * 4;
        for (uint32_t x = 0; x < m->n; ++x) {
            const float* c = m->cur + row + x;
            const float gx = (c[1] - c[-1]) * inv_two_cell;
            const float gz = (c[stride] - c[-(ptrdiff_t)stride]) * inv_two_cell;
            const float vy = (*c - m->old[row + x]) * inv_dt;
            dst[4 * x + 0] = *c;
            dst[4 * x + 1] = gx;
            dst[4 * x + 2] = gz;
            dst[4 * x + 3] = vy;
        }
    }
}

static int is_pow2(uint32_t v) { return v != 0 && (v & (v - 1)) == 0; }

/* interaction.cpp:438-475 validate + 477-507 constructor + Impl ctor 136-205 */
EXPORT orc_iw* orc_iw_create(const ocean_interaction_desc* d)
{
    if (!is_pow2(d->size) || d->size < 32 || d->size > 1024) return NULL;
    if (!(d->extent > 0.0f)) return NULL;
    if (d->kernel_radius < 1 || d->kernel_radius > 12) return NULL;
    if (2 * d->kernel_radius >= d->size) return NULL;
    if (!(d->damping >= 0.0f)) return NULL;
    if (d->absorb_cells * 2 >= d->size) return NULL;
    if (d->max_substeps < 1 || d->max_sources < 1) return NULL;
    if (!(d->gravity > 0.0f)) return NULL;
    if (d->fixed_dt < 0.0f) return NULL;

    orc_iw* m = calloc(1, sizeof(*m));
    m->desc = *d;
    m->n = d->size;
    m->cell = d->extent / (float)d->size;
    m->cells = (size_t)m->n * m->n;
    const size_t p = d->kernel_radius;
    m->stride = ((size_t)m->n + 2 * p + 15u) & ~(size_t)15u;
    const size_t total = ((size_t)m->n + 2 * p) * m->stride;
    if (!build_kernel((int)d->kernel_radius, m->cell, d->water_depth, &m->kernel)) {
        free(m);
        return NULL;
    }
    if (!(m->kernel.min_symbol >= -1e-3 * m->kernel.max_symbol && m->kernel.max_symbol > 0.0)) {
        free(m->kernel.taps);
        free(m);
        return NULL;
    }
    m->dt_limit = (float)(2.0 / sqrt((double)d->gravity * m->kernel.max_symbol));
    if (d->fixed_dt > 0.0f) {
        m->dt = d->fixed_dt;
    } else {
        const float a = 1.0f / 60.0f, b = 0.35f * m->dt_limit;
        m->dt = b < a ? b : a;
    }
    if (d->fixed_dt > 0.0f && d->fixed_dt > m->dt_limit) {
        free(m->kernel.taps);
        free(m);
        return NULL;
    }
    m->buf_a = calloc(total, sizeof(float));
    m->buf_b = calloc(total, sizeof(float));
    m->cur = base_of(m, m->buf_a);
    m->old = base_of(m, m->buf_b);
    m->c1 = calloc(m->cells, sizeof(float));
    m->c2 = calloc(m->cells, sizeof(float));
    m->out = calloc(m->cells * 4, sizeof(float));
    m->mask = calloc(m->cells, 1);
    m->nearest = calloc(m->cells, sizeof(uint32_t));
    m->solid = calloc(m->cells, sizeof(uint32_t));
    m->bfs = calloc(m->cells, sizeof(uint32_t));
    m->queue = calloc(d->max_sources, sizeof(ocean_disturbance));
    build_damping(m);
    finalize(m);
    return m;
}

EXPORT void orc_iw_destroy(orc_iw* m)
{
    if (!m) return;
    free(m->kernel.taps); free(m->buf_a); free(m->buf_b); free(m->c1); free(m->c2);
    free(m->out); free(m->mask); free(m->nearest); free(m->solid); free(m->bfs);
    free(m->queue); free(m);
}

/* interaction.cpp:511-524 */
EXPORT void orc_iw_add(orc_iw* m, const ocean_disturbance* d)
{
    if (m->queue_count >= m->desc.max_sources) {
        ++m->dropped;
        return;
    }
    m->queue[m->queue_count++] = *d;
}

EXPORT uint64_t orc_iw_dropped(const orc_iw* m) { return m->dropped; }

/* interaction.cpp:531-587 */
EXPORT void orc_iw_update(orc_iw* m, float dt)
{
    m->moved = 0;
    if (!(dt > 0.0f)) dt = 0.0f;
    m->accum += (double)dt;
    uint32_t steps = 0;
    if (m->accum >= m->dt) steps = (uint32_t)(m->accum / (double)m->dt);
    m->clamped_last = steps > m->desc.max_substeps;
    if (m->clamped_last) {
        steps = m->desc.max_substeps;
        m->accum = 0.0;
    } else {
        m->accum -= (double)steps * (double)m->dt;
    }
    m->substeps_last = steps;
    if (steps == 0) {
        uint32_t keep = 0;
        for (uint32_t i = 0; i < m->queue_count; ++i)
            if (m->queue[i].kind != OCEAN_SOURCE_CONTINUOUS) m->queue[keep++] = m->queue[i];
        m->queue_count = keep;
        return;
    }
    for (uint32_t s = 0; s < steps; ++s) {
        const float sub_time = (float)s * m->dt;
        for (uint32_t i = 0; i < m->queue_count; ++i) {
            const ocean_disturbance* d = &m->queue[i];
            if (d->kind != OCEAN_SOURCE_CONTINUOUS) {
                if (s == 0) inject(m, d, 0.0f, 1.0f);
            } else {
                inject(m, d, sub_time, m->dt);
            }
        }
        apply_obstruction(m);
        step(m);
        float* t = m->cur; m->cur = m->old; m->old = t;
    }
    m->queue_count = 0;
    finalize(m);
}

EXPORT uint32_t orc_iw_last_substeps(const orc_iw* m) { return m->substeps_last; }

static void shift_plane(orc_iw* m, float* pl, int32_t dx, int32_t dz)
{
    const int32_t n = (int32_t)m->n;
    const int32_t zlo = (dz > 0) ? 0 : n - 1, zhi = (dz > 0) ? n : -1, zst = (dz > 0) ? 1 : -1;
    for (int32_t z = zlo; z != zhi; z += zst) {
        const int32_t sz = z + dz;
        float* dst = pl + (ptrdiff_t)z * m->stride;
        if (sz < 0 || sz >= n) {
            memset(dst, 0, (size_t)n * sizeof(float));
            continue;
        }
        const float* src = pl + (ptrdiff_t)sz * m->stride;
        const int32_t lo = dx < 0 ? -dx : 0;
        const int32_t hi = (n < n - dx) ? n : n - dx;
        if (hi > lo) memmove(dst + lo, src + lo + dx, (size_t)(hi - lo) * sizeof(float));
        if (lo > 0) memset(dst, 0, (size_t)lo * sizeof(float));
        if (hi < n) memset(dst + hi, 0, (size_t)(n - hi) * sizeof(float));
    }
}

static void shift_mask(orc_iw* m, int32_t dx, int32_t dz)
{
    const int32_t n = (int32_t)m->n;
    const int32_t zlo = (dz > 0) ? 0 : n - 1, zhi = (dz > 0) ? n : -1, zst = (dz > 0) ? 1 : -1;
    for (int32_t z = zlo; z != zhi; z += zst) {
        const int32_t sz = z + dz;
        uint8_t* dst = m->mask + (ptrdiff_t)z * n;
        if (sz < 0 || sz >= n) {
            memset(dst, 0, (size_t)n);
            continue;
        }
        const uint8_t* src = m->mask + (ptrdiff_t)sz * n;
        const int32_t lo = dx < 0 ? -dx : 0;
        const int32_t hi = (n < n - dx) ? n : n - dx;
        if (hi > lo) memmove(dst + lo, src + lo + dx, (size_t)(hi - lo));
        if (lo > 0) memset(dst, 0, (size_t)lo);
        if (hi < n) memset(dst + hi, 0, (size_t)(n - hi));
    }
}

/* interaction.cpp:601-705; returns recentered() */
EXPORT int32_t orc_iw_recenter(orc_iw* m, float wx, float wz)
{
    const float half = 0.5f * m->desc.extent;
    const int32_t tx = (int32_t)lroundf((wx - half) / m->cell);
    const int32_t tz = (int32_t)lroundf((wz - half) / m->cell);
    const int32_t dx = tx - m->origin_x, dz = tz - m->origin_z;
    m->shift_x = dx;
    m->shift_z = dz;
    if (dx == 0 && dz == 0) {
        m->moved = 0;
        return 0;
    }
    m->moved = 1;
    const int32_t n = (int32_t)m->n;
    m->origin_x = tx;
    m->origin_z = tz;
    if (abs(dx) >= n || abs(dz) >= n) {
        const size_t p = m->desc.kernel_radius;
        const size_t total = ((size_t)m->n + 2 * p) * m->stride;
        memset(m->buf_a, 0, total * sizeof(float));
        memset(m->buf_b, 0, total * sizeof(float));
        memset(m->mask, 0, m->cells);
        rebuild_mask_tables(m);
        finalize(m);
        return 1;
    }
    shift_plane(m, m->cur, dx, dz);
    shift_plane(m, m->old, dx, dz);
    if (m->has_mask) {
        shift_mask(m, dx, dz);
        rebuild_mask_tables(m);
    }
    finalize(m);
    return 1;
}

/* interaction.cpp:716-726 */
EXPORT void orc_iw_set_obstruction(orc_iw* m, const uint8_t* mask)
{
    if (mask == NULL) {
        memset(m->mask, 0, m->cells);
        m->has_mask = 0;
        m->solid_count = 0;
        return;
    }
    memcpy(m->mask, mask, m->cells);
    m->has_mask = 1;
    rebuild_mask_tables(m);
}

/* interaction.cpp:729-773 */
EXPORT void orc_iw_sample_at(const orc_iw* m, float wx, float wz, ocean_interaction_sample* s)
{
    memset(s, 0, sizeof(*s));
    const float gx = (wx - (float)m->origin_x * m->cell) / m->cell - 0.5f;
    const float gz = (wz - (float)m->origin_z * m->cell) / m->cell - 0.5f;
    if (gx < 0.0f || gz < 0.0f || gx >= (float)(m->n - 1) || gz >= (float)(m->n - 1)) return;
    const int x0 = (int)gx, z0 = (int)gz;
    const float tx = gx - (float)x0, tz = gz - (float)z0;
    const float* f = m->out;
    const size_t i00 = ((size_t)z0 * m->n + x0) * 4, i10 = i00 + 4;
    const size_t i01 = i00 + (size_t)m->n * 4, i11 = i01 + 4;
    const float w00 = (1.0f - tx) * (1.0f - tz), w10 = tx * (1.0f - tz);
    const float w01 = (1.0f - tx) * tz, w11 = tx * tz;
    float v[4];
    for (int c = 0; c < 4; ++c)
        v[c] = f[i00 + c] * w00 + f[i10 + c] * w10 + f[i01 + c] * w01 + f[i11 + c] * w11;
    s->height = v[0];
    s->slope_x = v[1];
    s->slope_z = v[2];
    s->velocity_y = v[3];
}

/* interaction.cpp:803-846: E = 1/2 (sum v^2 + g sum h^n (G h^{n-1})), one
 * sequential double sum each in z-major cell order; the tap products are
 * float products added to a double accumulator. */
EXPORT double orc_iw_energy(const orc_iw* m)
{
    const int p = (int)m->desc.kernel_radius;
    const int w = 2 * p + 1;
    const size_t stride = m->stride;
    const double inv_dt = 1.0 / (double)m->dt;
    double kinetic = 0.0, potential = 0.0;
    for (uint32_t z = 0; z < m->n; ++z) {
        const size_t row = (size_t)z * stride;
        for (uint32_t x = 0; x < m->n; ++x) {
            const double hc = m->cur[row + x];
            const double hp = m->old[row + x];
            const double v = (hc - hp) * inv_dt;
            kinetic += v * v;
            const float* centre = m->old + row + x;
            const float* tap = m->kernel.taps;
            double acc = 0.0;
            for (int j = -p; j <= p; ++j) {
                const float* srow = centre + (ptrdiff_t)j * (ptrdiff_t)stride - p;
                for (int i = 0; i < w; ++i) {
                    const float prod = tap[i] * srow[i];
                    acc += (double)prod;
                }
                tap += w;
            }
            potential += hc * acc;
        }
    }
    return 0.5 * (kinetic + (double)m->desc.gravity * potential);
}

EXPORT const float* orc_iw_field(const orc_iw* m) { return m->out; }
/* The solid cells in index order and their nearest fluid cells; returns the count. */
EXPORT uint32_t orc_iw_mask_tables(const orc_iw* m, uint32_t* solid, uint32_t* source)
{
    for (uint32_t k = 0; k < m->solid_count; ++k) {
        solid[k] = m->solid[k];
        source[k] = m->nearest[m->solid[k]];
    }
    return m->solid_count;
}
EXPORT float orc_iw_fixed_dt(const orc_iw* m) { return m->dt; }
EXPORT float orc_iw_dt_limit(const orc_iw* m) { return m->dt_limit; }
EXPORT float orc_iw_origin_x(const orc_iw* m) { return (float)m->origin_x * m->cell; }
EXPORT float orc_iw_origin_z(const orc_iw* m) { return (float)m->origin_z * m->cell; }

/* WaterSurface::sample_at composition, interaction.cpp:863-901, given the
 * FFT surface sample `s` (ocean or cascade) at the same point. */
EXPORT void orc_compose(ocean_surface_ex* s, const ocean_interaction_sample* e)
{
    s->height += e->height;
    s->velocity_y += e->velocity_y;
    if (s->normal_y > 1.0e-6f) {
        const float inv = 1.0f / s->normal_y;
        const float sx = -s->normal_x * inv + e->slope_x;
        const float sz = -s->normal_z * inv + e->slope_z;
        const float len = sqrtf(sx * sx + sz * sz + 1.0f);
        const float il = 1.0f / len;
        s->normal_x = -sx * il;
        s->normal_y = il;
        s->normal_z = -sz * il;
    }
}
