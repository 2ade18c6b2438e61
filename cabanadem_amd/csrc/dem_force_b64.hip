// dem_force_b64.hip -- the force kernels of systems without bonds for
// tiles of 64 slots (dem_set_shape), in a translation unit of their own.
#include "dem_force_kernels.h"

namespace demhip {

hipError_t launch_force_b64(bool ref, bool ens, ForceKind kind, const StepArgs& a, bool sample,
                            int grid_blocks, hipStream_t s) {
  if (ens) return ref ? launch_force_mode<true, false, true, 64>(kind, a, sample, grid_blocks, s)
                      : launch_force_mode<false, false, true, 64>(kind, a, sample, grid_blocks, s);
  return ref ? launch_force_mode<true, false, false, 64>(kind, a, sample, grid_blocks, s)
             : launch_force_mode<false, false, false, 64>(kind, a, sample, grid_blocks, s);
}

} // namespace demhip
