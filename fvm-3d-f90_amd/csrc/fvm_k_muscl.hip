// Stage kernels of the muscl reconstruction family (exact numerics unless
// FVM_FAST is defined by the including file).
#include "fvm_instantiate.hpp"

namespace fvm {
#if FVM_FAST
namespace fast {
#else
namespace exact {
#endif

hipError_t launch_stage2d_muscl(int flux, int tile, bool conn, const StageArgs& a, int grid, hipStream_t s) {
  return Launcher<ReconMuscl>::stage2d(flux, tile, conn, a, grid, s);
}

hipError_t launch_stage3d_muscl(int flux, int tile, bool conn, const StageArgs& a, int grid, hipStream_t s) {
  return Launcher<ReconMuscl>::stage3d(flux, tile, conn, a, grid, s);
}

}  // namespace exact / fast
}  // namespace fvm
