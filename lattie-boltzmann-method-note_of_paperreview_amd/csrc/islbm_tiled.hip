// islbm_tiled.hip -- column-marching, LDS-staged stream-collide kernel (FP32 and 16-bit
// storage).  Placeholder until the tiled kernel lands: reports no support.
#include "islbm_internal.h"

namespace islbm {

bool tiled_supported(const LaunchGeom&, Store, bool) { return false; }

hipError_t launch_tiled(hipStream_t, const StepLaunch&) { return hipErrorNotSupported; }

}  // namespace islbm
