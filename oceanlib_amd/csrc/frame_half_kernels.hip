// The kModeFastHalf instantiations of the opt-in fast mode's frame passes:
// kModeFast (frame_fast_kernels.hip) with the row pass's complex fields
// stored as fp16 pairs between the passes.
#include "frame_launch.hpp"

using namespace oceanamd;

extern "C" hipError_t ocean_hip_frame_fast_2(const ocean_hip_level* L, double t2, hipStream_t s,
                                             unsigned passes, hipEvent_t* events)
{
    return launch_frame_mode<kModeFastHalf>(L, t2, s, passes, events);
}

extern "C" int ocean_hip_frame_shapes_fast_2(uint32_t n, uint32_t fields, int* row_variants,
                                             int* col_variants)
{
    return frame_shapes_mode<kModeFastHalf>(n, int(fields), *row_variants, *col_variants);
}

extern "C" void* ocean_hip_batch_kernel_fast_2(uint32_t n, int fields, int row, int per_item,
                                               int radix)
{
    return batch_kernel_mode<kModeFastHalf>(n, fields, row != 0, per_item != 0, radix);
}
