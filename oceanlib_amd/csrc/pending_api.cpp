// Entry points whose GPU implementation has not landed yet. Each one fails
// loudly (NULL handle + OCEAN_ERROR_UNKNOWN, or a no-op on a NULL handle);
// none of them computes anything on the CPU.
#include "ocean/ocean.h"
#include "ocean/ocean_amd.h"

#include <cstring>

extern "C" {

void ocean_interaction_desc_init(ocean_interaction_desc* d)
{
    if (d == nullptr) return;
    std::memset(d, 0, sizeof(*d));
    d->struct_size = sizeof(*d);
}
ocean_interaction* ocean_interaction_create(const ocean_interaction_desc*, ocean_status* st)
{
    if (st) *st = OCEAN_ERROR_UNKNOWN;
    return nullptr;
}
void ocean_interaction_destroy(ocean_interaction*) {}
void ocean_interaction_add(ocean_interaction*, const ocean_disturbance*) {}
uint64_t ocean_interaction_dropped(const ocean_interaction*) { return 0; }
void ocean_interaction_update(ocean_interaction*, float) {}
int32_t ocean_interaction_recenter(ocean_interaction*, float, float) { return 0; }
void ocean_interaction_set_obstruction(ocean_interaction*, const uint8_t*) {}
ocean_interaction_buffers ocean_interaction_get_buffers(const ocean_interaction*)
{
    ocean_interaction_buffers b;
    std::memset(&b, 0, sizeof(b));
    return b;
}
void ocean_interaction_sample_at(const ocean_interaction*, float, float,
                                 ocean_interaction_sample* o)
{
    if (o) std::memset(o, 0, sizeof(*o));
}
float ocean_interaction_height_at(const ocean_interaction*, float, float) { return 0.0f; }
float ocean_interaction_fixed_dt(const ocean_interaction*) { return 0.0f; }
float ocean_interaction_dt_limit(const ocean_interaction*) { return 0.0f; }
float ocean_interaction_dispersion_error(const ocean_interaction*) { return 0.0f; }
void ocean_sample_combined(const ocean_sim* sim, const ocean_interaction*, float x, float z,
                           ocean_surface_ex* out)
{
    ocean_sample_ex(sim, x, z, out);
}

}  // extern "C"
