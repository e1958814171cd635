// fvm_launch.hpp -- host-side launchers of the stage kernels (one translation
// unit per reconstruction family so the instantiations build in parallel).
// `conn`: some block of the launch has a face connected to another block
// (selects the kernel variant that also writes the neighbour's ghosts).
#pragma once
#include <hip/hip_runtime.h>

#include "fvm_types.hpp"

namespace fvm {

// 3-D tile shapes the library is built with: id -> (TX, TY); ids 4 and 5
// are the register-reconstruction kernel (stage3d_reg_kernel, simple and
// MUSCL reconstructions only)
constexpr int kNumTiles3d = 6;
constexpr int kTile3d[kNumTiles3d][2] = {{64, 4}, {32, 8}, {64, 2}, {16, 16}, {32, 8}, {16, 16}};
constexpr bool tile3d_reg(int tile) { return tile >= 4; }
// the register-reconstruction kernel takes stencils of at most 3 vertices
constexpr bool tile3d_ok(int recon, int tile) { return !tile3d_reg(tile) || recon != RECON_WENO; }
// 2-D stage shapes: id 0 the row kernel (62 columns marching over rows),
// then the tiled kernel's (TX, TY)
constexpr int kNumTiles2d = 3;
constexpr int kTile2d[kNumTiles2d][2] = {{62, 1}, {32, 4}, {16, 8}};
// viscous-residual tile (TX, TY); 3-D launches march over k-chunks
constexpr int kViscTile[2] = {32, 8};

namespace exact {
hipError_t launch_stage2d_simple(int flux, int tile, bool conn, const StageArgs& a, int grid, hipStream_t s);
hipError_t launch_stage2d_muscl(int flux, int tile, bool conn, const StageArgs& a, int grid, hipStream_t s);
hipError_t launch_stage2d_weno(int flux, int tile, bool conn, const StageArgs& a, int grid, hipStream_t s);
hipError_t launch_stage3d_simple(int flux, int tile, bool conn, const StageArgs& a, int grid, hipStream_t s);
hipError_t launch_stage3d_muscl(int flux, int tile, bool conn, const StageArgs& a, int grid, hipStream_t s);
hipError_t launch_stage3d_weno(int flux, int tile, bool conn, const StageArgs& a, int grid, hipStream_t s);
hipError_t launch_viscous(int dim, const ViscArgs& a, int grid, hipStream_t s);

}  // namespace exact
namespace fast {
hipError_t launch_stage2d_simple(int flux, int tile, bool conn, const StageArgs& a, int grid, hipStream_t s);
hipError_t launch_stage2d_muscl(int flux, int tile, bool conn, const StageArgs& a, int grid, hipStream_t s);
hipError_t launch_stage2d_weno(int flux, int tile, bool conn, const StageArgs& a, int grid, hipStream_t s);
hipError_t launch_stage3d_simple(int flux, int tile, bool conn, const StageArgs& a, int grid, hipStream_t s);
hipError_t launch_stage3d_muscl(int flux, int tile, bool conn, const StageArgs& a, int grid, hipStream_t s);
hipError_t launch_stage3d_weno(int flux, int tile, bool conn, const StageArgs& a, int grid, hipStream_t s);
hipError_t launch_viscous(int dim, const ViscArgs& a, int grid, hipStream_t s);

}  // namespace fast

}  // namespace fvm
