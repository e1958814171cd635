// fvm_instantiate.hpp -- kernel instantiation and flux / tile / connectivity
// dispatch of one reconstruction family, included once per family and
// numerics (exact or fast) so the instantiations build in parallel.
#pragma once
#include "fvm_launch.hpp"
#include "fvm_stage.hpp"

namespace fvm {
#if FVM_FAST
namespace fast {
#else
namespace exact {
#endif

template <class RC>
struct Launcher {
  template <int FLUX, bool CONN>
  static hipError_t go2d(const StageArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((stage2d_kernel<RC, FLUX, CONN>), dim3(grid), dim3(64), 0, s, a);
    return hipGetLastError();
  }

  template <int FLUX, int TX, int TY, bool CONN>
  static hipError_t go2d_tiled(const StageArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((stage2d_tile_kernel<RC, FLUX, TX, TY, CONN>), dim3(grid), dim3(TX * TY), 0, s, a);
    return hipGetLastError();
  }

  template <int FLUX, bool CONN>
  static hipError_t go2d_tile(int tile, const StageArgs& a, int grid, hipStream_t s) {
    switch (tile) {
      case 0: return go2d<FLUX, CONN>(a, grid, s);
      case 1: return go2d_tiled<FLUX, kTile2d[1][0], kTile2d[1][1], CONN>(a, grid, s);
      case 2: return go2d_tiled<FLUX, kTile2d[2][0], kTile2d[2][1], CONN>(a, grid, s);
      default: return hipErrorInvalidValue;
    }
  }

  template <int FLUX, int TX, int TY, bool CONN>
  static hipError_t go3d(const StageArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((stage3d_kernel<RC, FLUX, TX, TY, CONN>), dim3(grid), dim3(TX * TY), 0, s, a);
    return hipGetLastError();
  }

  template <int FLUX, int TX, int TY, bool CONN>
  static hipError_t go3d_reg(const StageArgs& a, int grid, hipStream_t s) {
    if constexpr (RC::RV <= 1) {
      hipLaunchKernelGGL((stage3d_reg_kernel<RC, FLUX, TX, TY, CONN>), dim3(grid), dim3(TX * TY), 0, s, a);
      return hipGetLastError();
    } else {
      return hipErrorInvalidValue;
    }
  }

  template <int FLUX, bool CONN>
  static hipError_t go3d_tile(int tile, const StageArgs& a, int grid, hipStream_t s) {
    switch (tile) {
      case 0: return go3d<FLUX, kTile3d[0][0], kTile3d[0][1], CONN>(a, grid, s);
      case 1: return go3d<FLUX, kTile3d[1][0], kTile3d[1][1], CONN>(a, grid, s);
      case 2: return go3d<FLUX, kTile3d[2][0], kTile3d[2][1], CONN>(a, grid, s);
      case 3: return go3d<FLUX, kTile3d[3][0], kTile3d[3][1], CONN>(a, grid, s);
      case 4: return go3d_reg<FLUX, kTile3d[4][0], kTile3d[4][1], CONN>(a, grid, s);
      case 5: return go3d_reg<FLUX, kTile3d[5][0], kTile3d[5][1], CONN>(a, grid, s);
      default: return hipErrorInvalidValue;
    }
  }

  template <bool CONN>
  static hipError_t stage2d_c(int flux, int tile, const StageArgs& a, int grid, hipStream_t s) {
    switch (flux) {
      case FLUX_ROE: return go2d_tile<FLUX_ROE, CONN>(tile, a, grid, s);
      case FLUX_VANLEER: return go2d_tile<FLUX_VANLEER, CONN>(tile, a, grid, s);
      case FLUX_AUSM: return go2d_tile<FLUX_AUSM, CONN>(tile, a, grid, s);
      case FLUX_AUSMPLUS: return go2d_tile<FLUX_AUSMPLUS, CONN>(tile, a, grid, s);
      default: return hipErrorInvalidValue;
    }
  }

  template <bool CONN>
  static hipError_t stage3d_c(int flux, int tile, const StageArgs& a, int grid, hipStream_t s) {
    switch (flux) {
      case FLUX_ROE: return go3d_tile<FLUX_ROE, CONN>(tile, a, grid, s);
      case FLUX_VANLEER: return go3d_tile<FLUX_VANLEER, CONN>(tile, a, grid, s);
      case FLUX_AUSM: return go3d_tile<FLUX_AUSM, CONN>(tile, a, grid, s);
      case FLUX_AUSMPLUS: return go3d_tile<FLUX_AUSMPLUS, CONN>(tile, a, grid, s);
      default: return hipErrorInvalidValue;
    }
  }

  static hipError_t stage2d(int flux, int tile, bool conn, const StageArgs& a, int grid, hipStream_t s) {
    return conn ? stage2d_c<true>(flux, tile, a, grid, s) : stage2d_c<false>(flux, tile, a, grid, s);
  }
  static hipError_t stage3d(int flux, int tile, bool conn, const StageArgs& a, int grid, hipStream_t s) {
    return conn ? stage3d_c<true>(flux, tile, a, grid, s) : stage3d_c<false>(flux, tile, a, grid, s);
  }
};

}  // namespace exact / fast
}  // namespace fvm
