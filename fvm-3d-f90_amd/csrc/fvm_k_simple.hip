// Stage-kernel instantiations for the simple reconstruction family.
#include "fvm_launch.hpp"

namespace fvm {

template <int FLUX>
static hipError_t go2d(const StageArgs& a, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL((stage2d_kernel<ReconSimple, FLUX>), dim3(nblocks), dim3(64), 0, s, a);
  return hipGetLastError();
}

template <int FLUX, int TY>
static hipError_t go3d(const StageArgs& a, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL((stage3d_kernel<ReconSimple, FLUX, TY>), dim3(nblocks), dim3(64 * TY), 0, s, a);
  return hipGetLastError();
}

template <int FLUX>
static hipError_t go3d_ty(int ty, const StageArgs& a, int nblocks, hipStream_t s) {
  switch (ty) {
    case 2: return go3d<FLUX, 2>(a, nblocks, s);
    case 4: return go3d<FLUX, 4>(a, nblocks, s);
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

hipError_t launch_stage3d_simple(int flux, int ty, const StageArgs& a, int nblocks, hipStream_t s) {
  switch (flux) {
    case FLUX_ROE: return go3d_ty<FLUX_ROE>(ty, a, nblocks, s);
    case FLUX_VANLEER: return go3d_ty<FLUX_VANLEER>(ty, a, nblocks, s);
    case FLUX_AUSM: return go3d_ty<FLUX_AUSM>(ty, a, nblocks, s);
    case FLUX_AUSMPLUS: return go3d_ty<FLUX_AUSMPLUS>(ty, a, nblocks, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace fvm
