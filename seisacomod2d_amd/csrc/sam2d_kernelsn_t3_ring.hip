// sam2d_kernelsn_t3_ring.hip -- instantiations of the recomputation form of the three-step kernel
// (k_stepn, sam2d_kernelsn.h) for every workgroup shape, prefetch depth and
// KZ aliasing.
#include "sam2d_kernelsn.h"

namespace sam2d {

hipError_t launch_stepn_t3_ring(int waves, int pf, bool kza, dim3 g, hipStream_t s, const StepNArgs &a)
{
    return launchn_waves<3, true>(waves, pf, kza, g, s, a);
}

}  // namespace sam2d
