/* CPU oracle of the renderer handoff format: IEEE binary32 -> binary16 and
 * the RGBA16F pack with the alpha substituted. TEST INFRASTRUCTURE ONLY: the
 * shipped library never links it.
 *
 * Restates reference examples/viewer/half.hpp:58-100 (float_to_half) and
 * half.hpp:108-125 (pack_rgba16_scalar). Pinned by tests/test_pack_oracle.py:
 * against numpy's float16 conversion (round to nearest even, overflow to inf)
 * on every non-NaN float32 bit pattern, NaNs against the closed form
 * sign | 0x7E00, and against the reference's own header compiled into
 * oracle/_ref (ref_shim.cpp) on the critical exponent range and random data. */
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* half.hpp:58-100 */
EXPORT uint16_t orc_float_to_half(float value)
{
    uint32_t x;
    memcpy(&x, &value, sizeof x);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t e32 = (x >> 23) & 0xFFu;
    const uint32_t m32 = x & 0x007FFFFFu;
    if (e32 == 0xFFu) return (uint16_t)(sign | 0x7C00u | (m32 ? 0x0200u : 0u));
    const int32_t e = (int32_t)e32 - 127 + 15;
    if (e >= 0x1F) return (uint16_t)(sign | 0x7C00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        const uint32_t m = m32 | 0x00800000u;
        const int s = 14 - e;
        uint32_t h = m >> s;
        const uint32_t rem = m & ((1u << s) - 1u);
        const uint32_t tie = 1u << (s - 1);
        if (rem > tie || (rem == tie && (h & 1u))) ++h;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((uint32_t)e << 10) | (m32 >> 13);
    const uint32_t rem = m32 & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
    return (uint16_t)(sign | h);
}

/* Every float whose bit pattern is in [first, first + count). */
EXPORT void orc_half_range(uint32_t first, uint32_t count, uint16_t* out)
{
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t bits = first + i;
        float f;
        memcpy(&f, &bits, sizeof f);
        out[i] = orc_float_to_half(f);
    }
}

/* half.hpp:108-125: RGB from src, A from alpha (or src's own w when NULL). */
EXPORT void orc_pack_rgba16(uint16_t* dst, const float* src, const float* alpha, size_t cells)
{
    for (size_t c = 0; c < cells; ++c) {
        dst[4 * c + 0] = orc_float_to_half(src[4 * c + 0]);
        dst[4 * c + 1] = orc_float_to_half(src[4 * c + 1]);
        dst[4 * c + 2] = orc_float_to_half(src[4 * c + 2]);
        dst[4 * c + 3] = orc_float_to_half(alpha ? alpha[c] : src[4 * c + 3]);
    }
}
