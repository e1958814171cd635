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

// Face kinds.  BC_CONN: the face abuts another block; its ghost layers are
// written by that block's stage kernel.  `shared` faces hold the interface
// vertex in both blocks (a grid split through a vertex plane); abutting faces
// join vertex n-1 of one block to vertex 0 of the next one spacing apart.
// BC_PATCH: the face's ghost layers are filled after every stage from 1-to-1
// connection patches (donor ranges with a direction transform).
enum : int { BC_EXTRAP = 0, BC_PERIODIC = 1, BC_CONN = 2, BC_PATCH = 3 };

// One 1-to-1 connection patch of a receiving block face: the receiver's
// ghost points beyond the patch take the donor vertices the linear index map
// d = dmin + S (r - rmin) gives (S: a signed permutation), extended through
// the ghost layers.  Indices 0-based.
struct PatchDesc {
  int recv, donor;                 // block indices in the context's table
  int dn, side;                    // normal direction of the receiving face, 0: low, 1: high
  int ta, tb;                      // the two tangential directions
  int na, nb;                      // patch extent along ta, tb (vertices)
  int nc;
  int map_dir[3], map_sign[3];     // receiver direction e -> donor direction, sign
  int64_t rmin[3], dmin[3];        // corresponding corners
  int64_t start;                   // first point of this patch in the launch's point list
  int fixed_state;                 // 1: the ghosts take `fixed` (farfield / inflow faces)
  int neg_all;                     // 1: every momentum component negated (no-slip wall
                                   // mirror), else each along map_sign
  double fixed[5];                 // conserved state, device component order
};

// Wall faces: a whole-face mirror patch of the block onto itself.  Slip
// (wall, inviscid wall, symmetry): the momentum normal to the face is
// negated in the ghosts.  No-slip (adiabatic viscous wall): every momentum
// component is negated in the ghosts, and each stage sets the wall vertices'
// momentum to zero, taking its kinetic energy out of the total energy
// (pressure, density and temperature unchanged).
enum : int { WALL_NONE = 0, WALL_SLIP = 1, WALL_NOSLIP = 2 };

// CFL wave-speed maximum of one solve (bits of a non-negative double, or of
// a NaN: as unsigned integers those order like the values and NaN wins),
// spread over kSmaxWays cache lines.  Each workgroup of a launch adds its
// maximum to line (workgroup id % kSmaxWays) with one device-scope atomic:
// atomics on one line are serialised (~10 ns each on MI355X, so one line
// taking a 1250-workgroup launch's atomics adds ~13 us to the launch);
// readers take the maximum over the lines.
constexpr int kSmaxWays = 32;
struct alignas(128) SmaxLine {
  unsigned long long v;
  unsigned long long pad[15];
};
struct SmaxWays {
  SmaxLine line[kSmaxWays];
};

__device__ __forceinline__ void smax_add(SmaxWays* w, double s) {
  atomicMax(&w->line[blockIdx.x % kSmaxWays].v, (unsigned long long)__double_as_longlong(s));
}
__device__ __forceinline__ double smax_get(const SmaxWays* w) {
  unsigned long long m = 0ull;
#pragma unroll
  for (int i = 0; i < kSmaxWays; ++i) m = w->line[i].v > m ? w->line[i].v : m;
  return __longlong_as_double((long long)m);
}
// The same read by a whole wave (all 64 lanes active): lane l < kSmaxWays
// loads line l, so the lines are fetched by one vector load instead of
// kSmaxWays scalar ones, and a butterfly gives every lane the maximum.
__device__ __forceinline__ double smax_get_wave(const SmaxWays* w) {
  static_assert(kSmaxWays <= 64 && (kSmaxWays & (kSmaxWays - 1)) == 0, "one line per lane");
  const int l = threadIdx.x & 63;
  unsigned long long m = l < kSmaxWays ? w->line[l].v : 0ull;
#pragma unroll
  for (int off = kSmaxWays / 2; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  return __longlong_as_double((long long)__shfl(m, 0, 64));
}
__device__ __forceinline__ void smax_clear(SmaxWays* w) {
#pragma unroll
  for (int i = 0; i < kSmaxWays; ++i) w->line[i].v = 0ull;
}

// One block as the stage kernels see it.
struct BlockDesc {
  double* q[4];          // state arrays: 0 = u^n, 1, 2 = stage buffers, 3 = RK4 accumulator
  BlockGeom g;
  int nbr[6];            // neighbour block per face (BC_CONN), else -1
  int nbr_shared[6];     // 1: shared interface vertex, 0: abutting
  int tiles_x, tiles_y, tiles_z, chunk;
  int tile_base;         // first tile of this block in the launch's tile list
  const double* dt_dev;            // the block's dt for this step (device memory: the
                                   // context's time slot, or the solve's own in serving)
  SmaxWays* smax;                  // per-block CFL max accumulator, or null
  const int* halt;                 // per-block: if set, the block's tiles are skipped
  const double* rv;                // viscous residual of the stage input (NC-1 components:
                                   // momentum, energy; stride g.cs), or null: inviscid
  int wall[6];                     // WALL_* per face
  int noslip;                      // bit f: face f is a no-slip wall
};

// Central viscous terms: Sutherland viscosity, T = p / (rho R), k = mu cp/Pr.
struct ViscParams {
  double mu_ref, t_ref, s;         // Sutherland: mu_ref (T/t_ref)^1.5 (t_ref + s) / (T + s)
  double r_gas;
  double cpr;                      // cp / Pr = gamma R / ((gamma - 1) Pr)
  double gamma;
  double rr_gas, rt_ref, sfac;     // 1 / r_gas, 1 / t_ref, t_ref + s (fast family)
};

struct ViscArgs {
  BlockGeom g;
  const double* q;                 // stage input state
  double* rv;                      // out: viscous residual, NC-1 components
  int tiles_x, tiles_y, chunk;     // launch decomposition (3-D: k-chunks of `chunk` planes)
  ViscParams V;
  // connected blocks: the context's table, this block's index in it and the
  // stage input's state array index; stencil points beyond a connected face
  // are read from the neighbour (null table: a single block)
  const BlockDesc* tab;
  int self, in_buf;
  const PatchDesc* patches;        // connection patches (stencil points beyond a
  int npatch;                      // patched face are read from the donor)
};

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

// Relative tolerance on t_end: a run whose time is this close counts as done
// (t + (t_end - t) can round one ulp short of t_end).
constexpr double kEndTol = 1.0e-14;

// Time-stepping state of one solve, kept on the device so a run of steps
// needs no host round trip: dt = min(cfl / smax, t_end - t), t += dt, and
// halt once t reaches t_end.  The wave-speed maximum smax lives beside it in
// an SmaxWays.
struct SolveSlot {
  double t, t_end, cfl, dt;
  int steps;
  int nan;                        // dt was NaN, zero or infinite
  int halt;                       // t reached t_end: later steps are skipped
  int pad;
};

struct StageArgs {
  const BlockDesc* blocks;        // descriptor table (device memory)
  int nblocks;                    // blocks in the table
  int tile0, ntiles;              // tiles [tile0, tile0 + ntiles) of the table's tile list;
                                  // a workgroup loops over them
  int in_buf, q0_buf, out_buf;    // state array index (BlockDesc::q) per role
  int use_acc;                    // RK4: accumulator q[3]
  Params P;
  int stage;
  double* norm_part;              // per-tile sums of rhs^2 (NC each), or null
  SmaxWays* smax;                 // CFL wave-speed max of the output, or null
  int block_smax;                 // 1: accumulate into each block's own smax instead
  const int* halt;                // if non-null and set, the launch does nothing (run ended)
  // a plain context's time slot in CFL mode (null otherwise): the first stage
  // of a step computes dt from it (cfl_dt), the second advances t and the
  // step count and clears the CFL maximum (commit); norm_if_end: the last
  // stage leaves norm partials only if the step reached t_end
  SolveSlot* slot;
  SmaxWays* slot_smax;            // the slot's wave-speed max (with slot)
  int cfl_dt, commit, norm_if_end;
};

}  // namespace fvm
