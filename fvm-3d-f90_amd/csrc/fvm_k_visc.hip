// Viscous-residual kernels (exact numerics unless FVM_FAST is defined by the
// including file).
#include "fvm_launch.hpp"
#include "fvm_visc.hpp"

namespace fvm {
#if FVM_FAST
namespace fast {
#else
namespace exact {
#endif

hipError_t launch_viscous(int dim, const ViscArgs& a, int grid, hipStream_t s) {
  if (dim == 3)
    hipLaunchKernelGGL((viscous3d_kernel<kViscTile[0], kViscTile[1]>), dim3(grid),
                       dim3(kViscTile[0] * kViscTile[1]), 0, s, a);
  else
    hipLaunchKernelGGL((viscous2d_kernel<kViscTile[0], kViscTile[1]>), dim3(grid),
                       dim3(kViscTile[0] * kViscTile[1]), 0, s, a);
  return hipGetLastError();
}

}  // namespace exact / fast
}  // namespace fvm
