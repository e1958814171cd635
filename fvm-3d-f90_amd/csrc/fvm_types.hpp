// fvm_types.hpp -- types shared by the host runtime and both kernel
// families (exact: bit-identical to the reference arithmetic; fast: FMA
// contraction and reciprocals, validated to the stated tolerance).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fvm {

// ---- scheme codes, src/defs/type_defs_scheme.f90:11-36 -------------------
enum : int {
  RK_SSP3 = 101, RK4 = 1001,
  RECON_SIMPLE = 200, RECON_MUSCL = 201, RECON_WENO = 202,
  LIM_NONE = 250, LIM_MINMOD = 251, LIM_VANLEER = 252, LIM_VANALBADA = 253,
  FLUX_ROE = 300, FLUX_VANLEER = 301, FLUX_AUSM = 303, FLUX_AUSMPLUS = 1303,
};

struct Params {
  double gamma, kappa, entfix;
  double g_gm1;   // gamma / (gamma - 1), evaluated once as the reference does per face
  int limiter;
  int weno_textbook;
};

struct BlockGeom {
  int nx, ny, nz;        // real vertices (nz = 1 in 2-D)
  int ox, oy, oz;        // array offset of real vertex 0 along i, j, k
  int64_t px, pxy, cs;   // row pitch, plane pitch, component stride (elements)
  double dx, dy, dz;
  double rdx, rdy, rdz;  // 1/dx, 1/dy, 1/dz (fast family only)
  int bc[6];             // i-, i+, j-, j+, k-, k+ : BC_EXTRAP / BC_PERIODIC
  int ng;                // ghost layers written per side
};

enum : int { BC_EXTRAP = 0, BC_PERIODIC = 1 };

// RK stage kinds: the combine each launch applies
enum : int {
  ST_SSP1 = 0,   // out = q0 + dt r
  ST_SSP2 = 1,   // out = (3/4 q0 + 1/4 qin) + (1/4 dt) r
  ST_SSP3 = 2,   // out = (1/3 q0 + 2/3 qin) + (2/3 dt) r
  ST_RK4_1 = 3,  // acc = q0 + dt/6 r ; out = q0 + dt/2 r
  ST_RK4_2 = 4,  // acc = acc + dt/3 r ; out = q0 + dt/2 r
  ST_RK4_3 = 5,  // acc = acc + dt/3 r ; out = q0 + dt r
  ST_RK4_4 = 6,  // out = acc + dt/6 r
};

struct StageArgs {
  const double* qin;     // stencil input, ghosts valid
  const double* q0;      // u^n, read at the cell only (may alias qout)
  double* qout;          // stage output (interior + ghost scatter)
  double* acc;           // RK4 accumulator (cell only)
  const double* dt_dev;  // if non-null, dt is read here (CFL mode)
  double dt;
  BlockGeom g;
  Params P;
  int stage;
  int tiles_x, tiles_y, tiles_z;  // tile grid; blockIdx.x is remapped over it
  int chunk;                      // planes (3-D) or rows (2-D) per workgroup
  double* norm_part;              // per-workgroup sums of rhs^2 (NC each), or null
  unsigned long long* smax;       // CFL wave-speed max of the output, or null
};

}  // namespace fvm
