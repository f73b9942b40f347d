// sam2d_kernelsn_t3.hip -- instantiations of the 3-step kernel
// (k_stepn, sam2d_kernelsn.h) for every workgroup shape, prefetch depth and
// KZ aliasing.
#include "sam2d_kernelsn.h"

namespace sam2d {

hipError_t launch_stepn_t3(int waves, int pf, bool kza, dim3 g, hipStream_t s, const StepNArgs &a)
{
    return launchn_waves<3, false>(waves, pf, kza, g, s, a);
}

}  // namespace sam2d
