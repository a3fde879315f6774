// The opt-in fast modes of the frame passes (no reference counterpart: the
// reference's only frame is the bit-exact one, and it names a real-to-complex
// FFT as its own speed-up path, ARCHITECTURE.md ADR-008). Opted into per
// ocean or per batch; never the default.
//
//   kModeFast     : the phase's sine and cosine from the hardware's v_sin /
//                   v_cos on the phase in revolutions (instead of the double
//                   fold to radians and the float polynomial), the butterfly
//                   rotations fused into FMAs, and the normal's 1/|n| from
//                   v_rsq (instead of a correctly rounded sqrt and divide).
//   kModeFastHalf : the same, with the row pass's complex fields stored as
//                   fp16 pairs: 4 instead of 8 bytes per field and cell
//                   between the passes.
// The real-to-complex saving the reference names is already in every mode:
// eight real spectra travel as four complex fields (A + iB).
//
// The same block code as the bit-exact kernels (frame_device.hpp), so the
// launch shapes, LDS layout and XCD mapping are unchanged; only the default
// shape and the field-split shapes are instantiated. This file holds
// kModeFast; frame_half_kernels.hip the same entry points for kModeFastHalf
// (two translation units build in parallel).
#include "frame_launch.hpp"

using namespace oceanamd;

extern "C" hipError_t ocean_hip_frame_fast_1(const ocean_hip_level* L, double t2, hipStream_t s,
                                             unsigned passes, hipEvent_t* events)
{
    return launch_frame_mode<kModeFast>(L, t2, s, passes, events);
}

extern "C" int ocean_hip_frame_shapes_fast_1(uint32_t n, uint32_t fields, int* row_variants,
                                             int* col_variants)
{
    return frame_shapes_mode<kModeFast>(n, int(fields), *row_variants, *col_variants);
}

extern "C" void* ocean_hip_batch_kernel_fast_1(uint32_t n, int fields, int row, int per_item,
                                               int radix)
{
    return batch_kernel_mode<kModeFast>(n, fields, row != 0, per_item != 0, radix);
}
