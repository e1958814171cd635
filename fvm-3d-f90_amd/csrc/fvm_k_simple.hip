// Stage-kernel instantiations for the simple reconstruction family.
#include "fvm_launch.hpp"
#include "fvm_stage.hpp"

namespace fvm {
#if FVM_FAST
namespace fast {
#else
namespace exact {
#endif

template <int FLUX>
static hipError_t go2d(const StageArgs& a, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL((stage2d_kernel<ReconSimple, FLUX>), dim3(nblocks), dim3(64), 0, s, a);
  return hipGetLastError();
}

template <int FLUX, int TX, int TY>
static hipError_t go3d(const StageArgs& a, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL((stage3d_kernel<ReconSimple, FLUX, TX, TY>), dim3(nblocks), dim3(TX * TY), 0, s, a);
  return hipGetLastError();
}

template <int FLUX>
static hipError_t go3d_ty(int tile, const StageArgs& a, int nblocks, hipStream_t s) {
  switch (tile) {
    case 0: return go3d<FLUX, 64, 4>(a, nblocks, s);
    case 1: return go3d<FLUX, 32, 8>(a, nblocks, s);
    case 2: return go3d<FLUX, 64, 2>(a, nblocks, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_stage2d_simple(int flux, const StageArgs& a, int nblocks, hipStream_t s) {
  switch (flux) {
    case FLUX_ROE: return go2d<FLUX_ROE>(a, nblocks, s);
    case FLUX_VANLEER: return go2d<FLUX_VANLEER>(a, nblocks, s);
    case FLUX_AUSM: return go2d<FLUX_AUSM>(a, nblocks, s);
    case FLUX_AUSMPLUS: return go2d<FLUX_AUSMPLUS>(a, nblocks, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_stage3d_simple(int flux, int tile, const StageArgs& a, int nblocks, hipStream_t s) {
  switch (flux) {
    case FLUX_ROE: return go3d_ty<FLUX_ROE>(tile, a, nblocks, s);
    case FLUX_VANLEER: return go3d_ty<FLUX_VANLEER>(tile, a, nblocks, s);
    case FLUX_AUSM: return go3d_ty<FLUX_AUSM>(tile, a, nblocks, s);
    case FLUX_AUSMPLUS: return go3d_ty<FLUX_AUSMPLUS>(tile, a, nblocks, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace exact / fast
}  // namespace fvm
