// fvm_launch.hpp -- host-side launchers of the stage kernels (one translation
// unit per reconstruction family so the instantiations build in parallel).
#pragma once
#include <hip/hip_runtime.h>

#include "fvm_stage.hpp"

namespace fvm {

// 3-D tile heights the library is built with (rows of 62 columns).
constexpr int kTileRows[] = {2, 4};

hipError_t launch_stage2d_simple(int flux, const StageArgs& a, int nblocks, hipStream_t s);
hipError_t launch_stage2d_muscl(int flux, const StageArgs& a, int nblocks, hipStream_t s);
hipError_t launch_stage2d_weno(int flux, const StageArgs& a, int nblocks, hipStream_t s);
hipError_t launch_stage3d_simple(int flux, int ty, const StageArgs& a, int nblocks, hipStream_t s);
hipError_t launch_stage3d_muscl(int flux, int ty, const StageArgs& a, int nblocks, hipStream_t s);
hipError_t launch_stage3d_weno(int flux, int ty, const StageArgs& a, int nblocks, hipStream_t s);

}  // namespace fvm
