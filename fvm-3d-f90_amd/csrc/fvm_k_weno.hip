// Stage kernels of the weno reconstruction family (exact numerics unless
// FVM_FAST is defined by the including file).
#include "fvm_instantiate.hpp"

namespace fvm {
#if FVM_FAST
namespace fast {
#else
namespace exact {
#endif

hipError_t launch_stage2d_weno(int flux, int tile, bool conn, const StageArgs& a, int grid, hipStream_t s) {
  return Launcher<ReconWeno>::stage2d(flux, tile, conn, a, grid, s);
}

hipError_t launch_stage3d_weno(int flux, int tile, bool conn, const StageArgs& a, int grid, hipStream_t s) {
  return Launcher<ReconWeno>::stage3d(flux, tile, conn, a, grid, s);
}

}  // namespace exact / fast
}  // namespace fvm
