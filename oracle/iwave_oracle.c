/* CPU oracle for the iWave interaction field (reference src/interaction.cpp,
 * src/core/iwave_step.cpp). TEST INFRASTRUCTURE ONLY - see ocean_oracle.h. */
#include "ocean_oracle.h"
