/* CPU oracle for the wave-synthesis path: a plain-C restatement of the
 * reference library's arithmetic, in the reference's operation order.
 *
 * TEST INFRASTRUCTURE ONLY. It is the checker the GPU path is compared with;
 * the shipped library never links, loads or calls it, and a missing GPU build
 * is an error in the library, never a reason to run this code instead.
 *
 * Build: make -C oracle oracle (gcc -O2 -ffp-contract=off -fno-fast-math),
 * mirroring the reference's FP policy (reference CMakeLists.txt:76-82).
 *
 * Parity pins (each function's check lives in tests/test_oracle_*.py):
 *   - PCG32 known answers, reference tests/test_rng.cpp:17-25;
 *   - every function here is compared bit for bit against the reference's own
 *     sources compiled into oracle/_ref (oracle/Makefile, ref_shim.cpp) on
 *     seeded inputs, and against committed fixtures in tests/golden/;
 *   - closed forms: sincos vs libm double to 2e-7, FFT vs naive DFT, chop-0
 *     exact special cases, Parseval.
 */
#ifndef OCEAN_ORACLE_H
#define OCEAN_ORACLE_H

#include <stdint.h>
#include "../include/ocean/ocean.h"

typedef struct orc_ocean orc_ocean;
typedef struct orc_foam orc_foam;

typedef struct orc_foam_params {
    float decay, source_gain, advect_scale, wind_drift_x, wind_drift_z, fixed_dt;
    uint32_t max_substeps;
} orc_foam_params;

#endif
