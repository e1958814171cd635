// fvm_stage.hpp -- fused RK-stage kernels for gfx950.
//
// One launch computes, for every real vertex of a block, the inviscid
// residual of all swept directions and the Runge-Kutta combine of that stage:
//   rhs = ((0 - dFx/dx) - dFy/dy) - dFz/dz          (class_space_discret.f90:678-732)
//   out = RK combine(q0, qin, acc, rhs)             (scheme_Runge_Kutta.f90:30-63)
// and writes the new values of the owned vertices.  Readers map a point
// outside the block to the vertex its ghost would copy (bmap), so no ghost
// layer is written or read, except across a connected face: there the stage
// writes its boundary vertices into the neighbour's ghost layers.
//
// Data: structure-of-arrays FP64, i fastest, each component a separate
// ghost-inclusive array; loads along i are coalesced.
//
// 3-D kernel: a workgroup owns a TX x TY tile of (i, j) columns and marches
// along k.  Each owner thread keeps a register queue of the reconstruction
// variables of its column, so every vertex is loaded from HBM once per plane
// and z-face fluxes are computed once and carried.  The current k plane,
// with its i/j halo, is staged in LDS; per-vertex reconstructions (both face
// values of each vertex share the limited slopes) and face fluxes of the
// in-plane directions are computed once per tile into LDS.
// 2-D kernel: the same with a one-row tile marching along j.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fvm_numerics.hpp"

namespace fvm {
#if FVM_FAST
namespace fast {
#else
namespace exact {
#endif

__device__ __forceinline__ int64_t gidx(const BlockGeom& g, int i, int j, int k) {
  return (int64_t)(k + g.oz) * g.pxy + (int64_t)(j + g.oy) * g.px + (i + g.ox);
}

// 32-bit element offset of (i, j, k) within one component array (blocks are
// limited to 2^29 elements per component, so byte offsets fit 32 bits and the
// loads address a uniform component base plus a 32-bit lane offset).
__device__ __forceinline__ uint32_t gidx32(const BlockGeom& g, int i, int j, int k) {
  return (uint32_t)((k + g.oz) * (int)g.pxy + (j + g.oy) * (int)g.px + (i + g.ox));
}
__device__ __forceinline__ double ld_comp(const double* __restrict__ q, int64_t cs, int c, uint32_t idx) {
  return *(const double*)((const char*)(q + c * cs) + (idx << 3));
}
__device__ __forceinline__ void st_comp(double* __restrict__ q, int64_t cs, int c, uint32_t idx, double v) {
  *(double*)((char*)(q + c * cs) + (idx << 3)) = v;
}

// ---------------------------------------------------------------- RK combine
struct RKCoef { double a, b, c, d; };

__device__ __forceinline__ RKCoef rk_coef(int stage, double dt) {
  RKCoef k{0.0, 0.0, 0.0, 0.0};
  switch (stage) {
    case ST_SSP1: k.c = dt; break;
    case ST_SSP2: k.a = 3.0 / 4.0; k.b = 1.0 / 4.0; k.c = 1.0 / 4.0 * dt; break;
    case ST_SSP3: k.a = 1.0 / 3.0; k.b = 2.0 / 3.0; k.c = 2.0 / 3.0 * dt; break;
    case ST_RK4_1: k.c = dt / 6.0; k.d = dt / 2.0; break;
    case ST_RK4_2: k.c = dt / 3.0; k.d = dt / 2.0; break;
    case ST_RK4_3: k.c = dt / 3.0; k.d = dt; break;
    default: k.c = dt / 6.0; break;
  }
  return k;
}

// Returns the new stage value; updates the RK4 accumulator in place.
__device__ __forceinline__ double rk_apply(int stage, const RKCoef& k, double q0, double qc,
                                           double* accp, double r) {
  switch (stage) {
    case ST_SSP1: return q0 + k.c * r;
    case ST_SSP2:
    case ST_SSP3: return k.a * q0 + k.b * qc + k.c * r;
    case ST_RK4_1: *accp = q0 + k.c * r; return q0 + k.d * r;
    case ST_RK4_2:
    case ST_RK4_3: *accp = *accp + k.c * r; return q0 + k.d * r;
    default: return *accp + k.c * r;
  }
}

// The combine of rk_apply over the NC components of one vertex, on operands
// loaded ahead: x is the stage input (SSP stages 2, 3) or the accumulator
// (RK4 stages 2-4); RK4 stages 1-3 store the new accumulator at ci.  One
// branch on the (uniform) stage per vertex.
template <int NC>
__device__ __forceinline__ void rk_apply_all(int stage, const RKCoef& k, const double q0[NC],
                                             const double x[NC], const double r[NC], double out[NC],
                                             double* __restrict__ acc, int64_t cs, uint32_t ci) {
  switch (stage) {
    case ST_SSP1:
#pragma unroll
      for (int m = 0; m < NC; ++m) out[m] = q0[m] + k.c * r[m];
      break;
    case ST_SSP2:
    case ST_SSP3:
#pragma unroll
      for (int m = 0; m < NC; ++m) out[m] = k.a * q0[m] + k.b * x[m] + k.c * r[m];
      break;
    case ST_RK4_1:
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        st_comp(acc, cs, m, ci, q0[m] + k.c * r[m]);
        out[m] = q0[m] + k.d * r[m];
      }
      break;
    case ST_RK4_2:
    case ST_RK4_3:
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        st_comp(acc, cs, m, ci, x[m] + k.c * r[m]);
        out[m] = q0[m] + k.d * r[m];
      }
      break;
    default:
#pragma unroll
      for (int m = 0; m < NC; ++m) out[m] = x[m] + k.c * r[m];
      break;
  }
}

// Per-launch time-slot work of a plain context's stage (StageArgs::slot);
// returns false when the launch has nothing to do (the run reached its end
// time).  The first stage of a CFL step computes dt with advance_slot's
// operations in every workgroup (workgroup 0 stores it, and the halt and NaN
// flags); the second advances t and the step count and clears the CFL maximum
// the last stage accumulates (no other workgroup reads them meanwhile); the
// last stage of a step that may be skipped leaves norms only if it reached
// t_end, read from the committed t.
__device__ __forceinline__ bool stage_clock(const StageArgs& a, double& dt, bool& want_norm) {
  want_norm = a.norm_part != nullptr;
  if (a.halt && *a.halt) return false;
  SolveSlot* S = a.slot;
  if (!S) return true;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (a.cfl_dt) {
    const double t = S->t, t_end = S->t_end;
    if (isfinite(t_end) && !(t < t_end - kEndTol * fabs(t_end))) {
      if (lead) S->halt = 1;
      return false;
    }
    const double sm = smax_get_wave(a.slot_smax);   // every thread of the workgroup is here
    double d = S->cfl / sm;
    const bool bad = !(d > 0.0) || !(d < 1.0e300);
    const double rest = t_end - t;
    if (d > rest) d = rest;
    dt = d;
    if (lead) {
      S->dt = d;
      if (bad) S->nan = 1;
    }
  }
  if (a.commit && lead) {
    S->t = S->t + S->dt;
    S->steps = S->steps + 1;
    smax_clear(a.slot_smax);
  }
  if (a.norm_if_end && want_norm) {
    const double t = S->t, t_end = S->t_end;
    want_norm = !(t < t_end - kEndTol * fabs(t_end));
  }
  return true;
}

// Loads the operands rk_apply_all needs at one vertex (uniform stage): PARTS
// bit 0 the RK base state q0, bit 1 the stage input or accumulator x.
template <int NC, int PARTS>
__device__ __forceinline__ void rk_prefetch(int stage, bool owner, uint32_t ci, int64_t cs,
                                            const double* __restrict__ q0p,
                                            const double* __restrict__ qin,
                                            const double* __restrict__ acc, double q0[NC],
                                            double x[NC]) {
  if constexpr (PARTS & 1) {
#pragma unroll
    for (int m = 0; m < NC; ++m) q0[m] = 0.0;
    if (owner && stage != ST_RK4_4) {
#pragma unroll
      for (int m = 0; m < NC; ++m) q0[m] = ld_comp(q0p, cs, m, ci);
    }
  }
  if constexpr (PARTS & 2) {
#pragma unroll
    for (int m = 0; m < NC; ++m) x[m] = 0.0;
    if (owner && (stage == ST_SSP2 || stage == ST_SSP3)) {
#pragma unroll
      for (int m = 0; m < NC; ++m) x[m] = ld_comp(qin, cs, m, ci);
    } else if (owner && (stage == ST_RK4_2 || stage == ST_RK4_3 || stage == ST_RK4_4)) {
#pragma unroll
      for (int m = 0; m < NC; ++m) x[m] = ld_comp(acc, cs, m, ci);
    }
  }
}

// ---------------------------------------------------------------- recon policies
// RV: vertex reconstruction radius; H = RV + 1: stencil radius of a cell.
// CONS: the reconstructed variables are conserved (converted at the face).
// SAME: both face values of a vertex are the same number.
struct ReconSimple {   // scheme_reconstruct_simple.f90:42-85
  static constexpr int RV = 0, H = 1;
  static constexpr bool CONS = false, SAME = true;
  template <int NC>
  __device__ __forceinline__ static void vertex(const double (*V)[NC], const Params&, double L[NC], double R[NC]) {
#pragma unroll
    for (int c = 0; c < NC; ++c) { L[c] = V[0][c]; R[c] = V[0][c]; }
  }
};

struct ReconMuscl {    // scheme_reconstruct_MUSCL.f90:47-226
  static constexpr int RV = 1, H = 2;
  static constexpr bool CONS = false, SAME = false;
  template <int NC, int LIM>
  __device__ __forceinline__ static void vertex_lim(const double (*V)[NC], double kappa, double L[NC], double R[NC]) {
    muscl_vertex_all<LIM, NC>(V, kappa, L, R);
  }
  // one uniform branch on the limiter per vertex, none per component
  template <int NC>
  __device__ __forceinline__ static void vertex(const double (*V)[NC], const Params& P, double L[NC], double R[NC]) {
    switch (P.limiter) {
      case LIM_MINMOD: vertex_lim<NC, LIM_MINMOD>(V, P.kappa, L, R); break;
      case LIM_VANLEER: vertex_lim<NC, LIM_VANLEER>(V, P.kappa, L, R); break;
      case LIM_VANALBADA: vertex_lim<NC, LIM_VANALBADA>(V, P.kappa, L, R); break;
      default: vertex_lim<NC, LIM_NONE>(V, P.kappa, L, R); break;
    }
  }
};

struct ReconWeno {     // scheme_WENO.f90:28-121, then q -> p at the face
  static constexpr int RV = 2, H = 3;
  static constexpr bool CONS = true, SAME = false;
  template <int NC>
  __device__ __forceinline__ static void vertex(const double (*V)[NC], const Params& P, double L[NC], double R[NC]) {
    double qL[NC], qR[NC];
    // textbook WENO: both faces of the vertex share the smoothness
    // indicators (in the 3-D stage this spills a few registers, and is still
    // faster than two independent evaluations)
    if (kFast && P.weno_textbook) {
#pragma unroll
      for (int c = 0; c < NC; ++c) weno_js_pair(V[0][c], V[1][c], V[2][c], V[3][c], V[4][c], qL[c], qR[c]);
      cons_to_prim<NC>(qL, L, P.gamma);
      cons_to_prim<NC>(qR, R, P.gamma);
      return;
    }
    if (!kFast && P.weno_textbook) {
      // textbook: the right state is the mirror image of the stencil; both
      // faces share the smoothness indicators
#pragma unroll
      for (int c = 0; c < NC; ++c) weno_js_pair_exact(V[0][c], V[1][c], V[2][c], V[3][c], V[4][c], qL[c], qR[c]);
      cons_to_prim<NC>(qL, L, P.gamma);
      cons_to_prim<NC>(qR, R, P.gamma);
      return;
    }
    // reference: the right state of a face is the next face's left state
    // (scheme_WENO.f90:54-56)
#pragma unroll
    for (int c = 0; c < NC; ++c) qL[c] = weno_js(V[0][c], V[1][c], V[2][c], V[3][c], V[4][c]);
    cons_to_prim<NC>(qL, L, P.gamma);
#pragma unroll
    for (int c = 0; c < NC; ++c) R[c] = L[c];
  }
};
template <int FLUX, int NC, int D>
__device__ __forceinline__ void face_flux(const double L[NC], const double R[NC], const Params& P,
                                          double F[NC]) {
  if constexpr (FLUX == FLUX_ROE) flux_roe<NC, D>(L, R, P, F);
  else if constexpr (FLUX == FLUX_VANLEER) flux_vanleer<NC, D>(L, R, P, F);
  else if constexpr (FLUX == FLUX_AUSM) flux_ausm<NC, D>(L, R, P, F);
  else flux_ausmplus<NC, D>(L, R, P, F);
}

// Load the reconstruction variables of one vertex (clamped indices must be
// inside the allocation; the caller clamps).
template <class RC, int NC>
__device__ __forceinline__ void load_raw(const double* __restrict__ q, int64_t cs, uint32_t idx,
                                         double v[NC]) {
#pragma unroll
  for (int c = 0; c < NC; ++c) v[c] = ld_comp(q, cs, c, idx);
}

template <class RC, int NC>
__device__ __forceinline__ void to_recon_vars(const double raw[NC], double v[NC], double gamma) {
  if constexpr (RC::CONS) {
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = raw[c];
  } else {
    cons_to_prim<NC>(raw, v, gamma);
  }
}

__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// Index of a stencil point along one direction.  Real vertices map to
// themselves; outside the block, to the vertex its ghost copy holds (the
// periodic image, or the boundary vertex under extrapolation,
// src/base/defs_block_data.f90:226-256), except across a connected face
// (BC_CONN: the neighbour's stage writes the ghost layer; BC_PATCH: a copy
// kernel fills it after every stage), whose ghosts are read.  Reading the source vertex
// instead of a ghost copy gives the same value, so the stage writes no
// ghost images of its own.
__device__ __forceinline__ int bmap(int x, int n, int lo, int hi, int h) {
  x = clampi(x, -h, n + h - 1);
  if (x < 0) return lo == BC_PERIODIC ? x + n : (lo >= BC_CONN ? x : 0);
  if (x >= n) return hi == BC_PERIODIC ? x - n : (hi >= BC_CONN ? x : n - 1);
  return x;
}

// max(a, b) where a NaN operand wins, so a diverged state reaches the CFL
// accumulator (fmax would drop it).
__device__ __forceinline__ double nan_max(double a, double b) { return (b > a || b != b) ? b : a; }

// Wave speed sum_d (|u_d| + c) / dr_d of a conserved state (CFL time step).
template <int NC>
__device__ __forceinline__ double wave_speed(const double q[NC], const BlockGeom& g, double gamma) {
  double p[NC];
  cons_to_prim<NC>(q, p, gamma);
  if constexpr (kFast) {
    const double a = fsqrt(gamma * p[NC - 1] * frcp(p[0]));
    double s = (fabs(p[1]) + a) * g.rdx + (fabs(p[2]) + a) * g.rdy;
    if constexpr (NC == 5) s = s + (fabs(p[3]) + a) * g.rdz;
    return s;
  }
  const double a = esqrt(ediv(gamma * p[NC - 1], p[0]));
  double s = ediv(fabs(p[1]) + a, g.dx) + ediv(fabs(p[2]) + a, g.dy);
  if constexpr (NC == 5) s = s + ediv(fabs(p[3]) + a, g.dz);
  return s;
}

// Index of (i, j, k) in a block whose geometry is read from the descriptor
// table (neighbour writes).
template <class Tab>
__device__ __forceinline__ int64_t gidx_of(Tab tab, int blk, int i, int j, int k) {
  const auto& g = tab[blk].g;
  return (int64_t)(k + g.oz) * g.pxy + (int64_t)(j + g.oy) * g.px + (i + g.ox);
}

// Write one output vertex and, across connected faces, the neighbour block's
// ghosts it donates to (shared interface vertex: vertex n-1-l of this block is
// ghost -l of the next; abutting blocks: vertex n-l is ghost -l).  The block's
// own ghosts are not written: readers map such points with bmap.  A vertex on
// a no-slip wall first loses its momentum and the momentum's kinetic energy
// (v is updated in place, so the caller's CFL wave speed sees the stored
// state).  Both belong to the CONN variant, which launches run whenever a
// block has a connected face or a no-slip wall: the other variant carries
// neither's code.
template <int NC, int NDIR, bool CONN, class Tab, class Desc>
__device__ __forceinline__ void store_with_ghosts(const BlockGeom& g, Tab tab, const Desc& B, int out_buf,
                                                  double* __restrict__ out, int i, int j, int k,
                                                  double v[NC]) {
  if (CONN && B.noslip) {   // uniform over the block
    const int nn[3] = {g.nx, g.ny, g.nz}, xx[3] = {i, j, k};
    bool on = false;
#pragma unroll
    for (int d = 0; d < NDIR; ++d)
      on = on || (xx[d] == 0 && ((B.noslip >> (2 * d)) & 1)) ||
           (xx[d] == nn[d] - 1 && ((B.noslip >> (2 * d + 1)) & 1));
    if (on) {
      double m2 = v[1] * v[1] + v[2] * v[2];
      if constexpr (NC == 5) m2 = m2 + v[3] * v[3];
      const double ke = 0.5 * m2 / v[0];
      v[NC - 1] = v[NC - 1] - ke;
#pragma unroll
      for (int m = 1; m < NC - 1; ++m) v[m] = 0.0;
    }
  }
  const uint32_t base = gidx32(g, i, j, k);
#pragma unroll
  for (int c = 0; c < NC; ++c) st_comp(out, g.cs, c, base, v[c]);
  if constexpr (CONN) {
    const int n[3] = {g.nx, g.ny, g.nz};
    const int x[3] = {i, j, k};
    // only vertices within ng + 1 of a face donate; whole waves of interior
    // vertices skip the rest
    bool near = false;
#pragma unroll
    for (int d = 0; d < NDIR; ++d) near = near || x[d] <= g.ng || x[d] >= n[d] - 1 - g.ng;
    if (!near) return;
#pragma unroll
    for (int d = 0; d < NDIR; ++d) {
      const int blo = g.bc[2 * d], bhi = g.bc[2 * d + 1];
      for (int l = 1; l <= g.ng; ++l) {
        // neighbour across the low face: its ghost n_nb - 1 + l
        if (blo == BC_CONN && x[d] == (B.nbr_shared[2 * d] ? l : l - 1)) {
          const int nb = B.nbr[2 * d];
          int y[3] = {x[0], x[1], x[2]};
          const int nn[3] = {tab[nb].g.nx, tab[nb].g.ny, tab[nb].g.nz};
          y[d] = nn[d] - 1 + l;
          double* dst = tab[nb].q[out_buf];
          const int64_t o = gidx_of(tab, nb, y[0], y[1], y[2]);
          const int64_t cs = tab[nb].g.cs;
#pragma unroll
          for (int c = 0; c < NC; ++c) dst[o + c * cs] = v[c];
        }
        // neighbour across the high face: its ghost -l
        if (bhi == BC_CONN && x[d] == (B.nbr_shared[2 * d + 1] ? n[d] - 1 - l : n[d] - l)) {
          const int nb = B.nbr[2 * d + 1];
          int y[3] = {x[0], x[1], x[2]};
          y[d] = -l;
          double* dst = tab[nb].q[out_buf];
          const int64_t o = gidx_of(tab, nb, y[0], y[1], y[2]);
          const int64_t cs = tab[nb].g.cs;
#pragma unroll
          for (int c = 0; c < NC; ++c) dst[o + c * cs] = v[c];
        }
      }
    }
  }
}

// The descriptor table is read through the constant address space: the
// stage writes only state arrays, never the table, and loads marked
// invariant let the compiler re-load descriptor fields instead of keeping
// them live in registers across the k-march.
typedef const __attribute__((address_space(4))) BlockDesc* DescTable;
__device__ __forceinline__ DescTable as_const_table(const BlockDesc* p) { return (DescTable)p; }

// Member-wise copy out of the constant address space.
template <class G>
__device__ __forceinline__ BlockGeom load_geom(const G& s) {
  BlockGeom g;
  g.nx = s.nx; g.ny = s.ny; g.nz = s.nz;
  g.ox = s.ox; g.oy = s.oy; g.oz = s.oz;
  g.px = s.px; g.pxy = s.pxy; g.cs = s.cs;
  g.dx = s.dx; g.dy = s.dy; g.dz = s.dz;
  g.rdx = s.rdx; g.rdy = s.rdy; g.rdz = s.rdz;
#pragma unroll
  for (int f = 0; f < 6; ++f) g.bc[f] = s.bc[f];
  g.ng = s.ng;
  return g;
}

// The block owning `tile`: the last block whose first tile is at or before
// it (tile bases ascend in table order); a binary search, so a launch over
// hundreds of served solves costs a few table reads per workgroup.
template <class Tab>
__device__ __forceinline__ int block_of_tile(Tab tab, int nblocks, int tile) {
  int lo = 0, hi = nblocks - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab[mid].tile_base <= tile) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Workgroup reduction helpers (deterministic order).
template <int NT>
__device__ __forceinline__ double wg_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < NT / 64; ++w) s += red[w];
  __syncthreads();
  return s;
}

template <int NT>
__device__ __forceinline__ double wg_max(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v = nan_max(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    s = red[0];
    for (int w = 1; w < NT / 64; ++w) {
      s = nan_max(s, red[w]);
    }
  }
  __syncthreads();
  return s;
}

// ============================================================================
// 2-D stage: tile of TX = 62 columns, one wave, marching along j.
// Thread t owns column c = i0 - 1 + t (t = 0..63): it keeps the j-queue of
// its column, reconstructs its vertex along i, and for t in 1..62 updates
// cell c.  A workgroup loops over tiles of the launch's tile list.
// ============================================================================
template <class RC, int FLUX, bool CONN>
__global__ __launch_bounds__(64) void stage2d_kernel(StageArgs a) {
  constexpr int NC = 4, TX = 62, NT = 64;
  constexpr int RV = RC::RV, H = RC::H;
  constexpr int NW = RV + (RV > 1 ? RV : 1) + 1;   // queue: V(c - max(RV,1) .. c + RV)
  constexpr int QOFF = (RV > 1 ? RV : 1);          // queue index of the centre c
  __shared__ double Vs[NC][NT + 2 * RV];
  __shared__ double Lx[NC][NT];
  __shared__ double Rx[NC][NT];
  __shared__ double Fx[NC][NT];
  __shared__ double red[NT / 64];

  const Params& P = a.P;
  const int t = threadIdx.x;
  const bool want_smax = a.smax != nullptr || a.block_smax;

  double dt_step = 0.0;
  bool want_norm = false;
  if (!stage_clock(a, dt_step, want_norm)) return;
  const DescTable tab = as_const_table(a.blocks);
  for (int it = blockIdx.x; it < a.ntiles; it += gridDim.x) {
    const int tile = a.tile0 + xcd_remap(it, a.ntiles);
    const int bi = block_of_tile(tab, a.nblocks, tile);
    const auto& B = tab[bi];
    if (B.halt && *B.halt) continue;   // uniform over the workgroup
    const BlockGeom g = load_geom(B.g);
    const double dt = a.cfl_dt ? dt_step : *B.dt_dev;
    const double* __restrict__ rvp = B.rv;
    const RKCoef kc = rk_coef(a.stage, dt);
    double smax = 0.0;
    const double* __restrict__ qin = B.q[a.in_buf];
    const double* __restrict__ q0p = B.q[a.q0_buf];
    double* __restrict__ qout = B.q[a.out_buf];
    double* __restrict__ acc = a.use_acc ? B.q[3] : nullptr;
    const int lt = tile - B.tile_base;
    const int ti = lt % B.tiles_x, tj = lt / B.tiles_x;
    const int i0 = ti * TX;
    const int j0 = tj * B.chunk;
    const int j1 = min(j0 + B.chunk, g.ny);
    const int c = i0 - 1 + t;                       // owned column
    const int cl = bmap(c, g.nx, g.bc[0], g.bc[1], H);
    const bool owner = (t >= 1) && (t <= TX) && (c < g.nx);

    // j-queue of column c: Vq[w] = V(c, jc - QOFF + w), jc = vertex being reconstructed
    double Vq[NW][NC];
    auto loadV = [&](int j, double v[NC]) {
      double raw[NC];
      load_raw<RC, NC>(qin, g.cs, gidx32(g, cl, bmap(j, g.ny, g.bc[2], g.bc[3], H), 0), raw);
      to_recon_vars<RC, NC>(raw, v, P.gamma);
    };
#pragma unroll
    for (int w = 0; w < NW; ++w) loadV(j0 - 1 - QOFF + w, Vq[w]);
    double Lc[NC], Fc[NC], Lt[NC], Rt[NC];
    // vertex j0-1: carry its L
    RC::template vertex<NC>(&Vq[QOFF - RV], P, Lt, Rt);
#pragma unroll
    for (int m = 0; m < NC; ++m) Lc[m] = Lt[m];
    // vertex j0: face j0 = F(L(j0-1), R(j0))
#pragma unroll
    for (int w = 0; w < NW - 1; ++w)
#pragma unroll
      for (int m = 0; m < NC; ++m) Vq[w][m] = Vq[w + 1][m];
    loadV(j0 + RV, Vq[NW - 1]);
    RC::template vertex<NC>(&Vq[QOFF - RV], P, Lt, Rt);
    face_flux<FLUX, NC, 1>(Lc, Rt, P, Fc);
#pragma unroll
    for (int m = 0; m < NC; ++m) Lc[m] = Lt[m];

    double nrm[NC];
#pragma unroll
    for (int m = 0; m < NC; ++m) nrm[m] = 0.0;

    for (int j = j0; j < j1; ++j) {
      // shift: centre becomes j + 1; Vq[QOFF - 1] is V(j)
#pragma unroll
      for (int w = 0; w < NW - 1; ++w)
#pragma unroll
        for (int m = 0; m < NC; ++m) Vq[w][m] = Vq[w + 1][m];
      loadV(j + 1 + RV, Vq[NW - 1]);
      // row j into LDS: owned columns plus RV far-halo columns on each side
#pragma unroll
      for (int m = 0; m < NC; ++m) Vs[m][t + RV] = Vq[QOFF - 1][m];
      if (t < 2 * RV) {
        const int hc = t < RV ? i0 - 1 - RV + t : i0 + TX + 1 + (t - RV);
        double raw[NC], v[NC];
        load_raw<RC, NC>(qin, g.cs, gidx32(g, bmap(hc, g.nx, g.bc[0], g.bc[1], H), j, 0), raw);
        to_recon_vars<RC, NC>(raw, v, P.gamma);
        const int slot = t < RV ? t : NT + RV + (t - RV);
#pragma unroll
        for (int m = 0; m < NC; ++m) Vs[m][slot] = v[m];
      }
      __syncthreads();
      // i-reconstruction at the owned vertex
      {
        double W[2 * RV + 1][NC];
#pragma unroll
        for (int s = 0; s < 2 * RV + 1; ++s)
#pragma unroll
          for (int m = 0; m < NC; ++m) W[s][m] = Vs[m][t + s];
        double L[NC], R[NC];
        RC::template vertex<NC>(W, P, L, R);
#pragma unroll
        for (int m = 0; m < NC; ++m) { Lx[m][t] = L[m]; Rx[m][t] = R[m]; }
      }
      __syncthreads();
      // i-faces f = i0 + t, t = 0..62: F(L(f-1), R(f))
      if (t <= TX) {
        double L[NC], R[NC], F[NC];
#pragma unroll
        for (int m = 0; m < NC; ++m) { L[m] = Lx[m][t]; R[m] = Rx[m][t + 1]; }
        face_flux<FLUX, NC, 0>(L, R, P, F);
#pragma unroll
        for (int m = 0; m < NC; ++m) Fx[m][t] = F[m];
      }
      // j-face j + 1 of the owned column
      double Fn[NC];
      RC::template vertex<NC>(&Vq[QOFF - RV], P, Lt, Rt);
      face_flux<FLUX, NC, 1>(Lc, Rt, P, Fn);
      __syncthreads();
      if (owner && j < g.ny) {
        const int64_t ci = gidx(g, c, j, 0);
        const Recip ix = kFast ? Recip{} : erecip(g.dx), iy = kFast ? Recip{} : erecip(g.dy);
        double out[NC];
#pragma unroll
        for (int m = 0; m < NC; ++m) {
          double r;
          if constexpr (kFast) {
            r = -(Fx[m][t] - Fx[m][t - 1]) * g.rdx;
            r = r - (Fn[m] - Fc[m]) * g.rdy;
          } else {
            r = 0.0 - ediv(Fx[m][t] - Fx[m][t - 1], ix);
            r = r - ediv(Fn[m] - Fc[m], iy);
          }
          if (m > 0 && rvp) r = r + rvp[(int64_t)(m - 1) * g.cs + ci];
          const double q0 = q0p[ci + m * g.cs];
          double* accp = acc ? acc + ci + m * g.cs : nullptr;
          out[m] = rk_apply(a.stage, kc, q0, qin[ci + m * g.cs], accp, r);
          nrm[m] += r * r;
        }
        store_with_ghosts<NC, 2, CONN>(g, tab, B, a.out_buf, qout, c, j, 0, out);
        if (want_smax) smax = nan_max(smax, wave_speed<NC>(out, g, P.gamma));
      }
#pragma unroll
      for (int m = 0; m < NC; ++m) { Fc[m] = Fn[m]; Lc[m] = Lt[m]; }
      __syncthreads();
    }

    if (want_norm) {
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        const double s = wg_sum<NT>(nrm[m], red);
        if (t == 0) a.norm_part[(int64_t)tile * NC + m] = s;
      }
    }
    if (want_smax) {
      const double s = wg_max<NT>(smax, red);
      if (t == 0) smax_add(a.block_smax ? B.smax : a.smax, s);
    }
  }
}

// ============================================================================
// 2-D tiled stage: a workgroup owns a TX x TY tile of vertices of the plane.
//   A  the tile and its halo (H vertices) into LDS as reconstruction variables
//   B  vertex reconstructions along i and j (owned vertices, then the halo
//      column pair and row pair the tile's faces need)
//   C  face fluxes along i and j into LDS (aliasing A's plane)
//   D  owner update: residual, RK combine, store
// Same per-vertex and per-face arithmetic as the row kernel above (same
// results bit for bit); fewer halo reconstructions per vertex and
// multi-wave workgroups, for planes too small to fill the GPU with rows.
// ============================================================================
template <class RC, int FLUX, int TX, int TY, bool CONN>
__global__ __launch_bounds__(TX * TY) void stage2d_tile_kernel(StageArgs a) {
  constexpr int NC = 4, NT = TX * TY;
  constexpr int RV = RC::RV, H = RC::H;
  constexpr int VW = TX + 2 * H, VR = TY + 2 * H, VSN = VW * VR;
  constexpr int NLR = RC::SAME ? 1 : 2;
  constexpr int XRW = TX + 2, LRX = TY * XRW;      // i-recon columns i0-1 .. i0+TX
  constexpr int LRY = (TY + 2) * TX;               // j-recon rows j0-1 .. j0+TY
  constexpr int FXW = TX + 1, FXN = TY * FXW, FYN = (TY + 1) * TX;
  constexpr int AREA = VSN > FXN + FYN ? VSN : FXN + FYN;
  constexpr int RT = 2 * NT + 2 * TY + 2 * TX;     // reconstruction tasks
  constexpr int FT = 2 * NT + TY + TX;             // face tasks
  __shared__ double Vs[NC * AREA];
  __shared__ double LRx[NLR][NC * LRX];
  __shared__ double LRy[NLR][NC * LRY];
  __shared__ double red[NT / 64];
  double* Fx = Vs;                 // [NC][FXN] after phase B
  double* Fy = Vs + NC * FXN;      // [NC][FYN]

  const Params& P = a.P;
  const int t = threadIdx.x;
  const int tx = t % TX, ty = t / TX;
  const bool want_smax = a.smax != nullptr || a.block_smax;

  double dt_step = 0.0;
  bool want_norm = false;
  if (!stage_clock(a, dt_step, want_norm)) return;
  const DescTable tab = as_const_table(a.blocks);
  for (int it = blockIdx.x; it < a.ntiles; it += gridDim.x) {
    const int tile = a.tile0 + xcd_remap(it, a.ntiles);
    const int bi = block_of_tile(tab, a.nblocks, tile);
    const auto& B = tab[bi];
    if (B.halt && *B.halt) continue;   // uniform over the workgroup
    const BlockGeom g = load_geom(B.g);
    const double dt = a.cfl_dt ? dt_step : *B.dt_dev;
    const double* __restrict__ rvp = B.rv;
    const RKCoef kc = rk_coef(a.stage, dt);
    const double* __restrict__ qin = B.q[a.in_buf];
    const double* __restrict__ q0p = B.q[a.q0_buf];
    double* __restrict__ qout = B.q[a.out_buf];
    double* __restrict__ acc = a.use_acc ? B.q[3] : nullptr;
    const int lt = tile - B.tile_base;
    const int i0 = (lt % B.tiles_x) * TX, j0 = (lt / B.tiles_x) * TY;
    const int c = i0 + tx, r = j0 + ty;
    const bool owner = c < g.nx && r < g.ny;

    // ---- A: the plane (tile + halo) into LDS
    for (int e = t; e < VSN; e += NT) {
      const int x = i0 - H + e % VW, y = j0 - H + e / VW;
      double raw[NC], v[NC];
      load_raw<RC, NC>(qin, g.cs, gidx32(g, bmap(x, g.nx, g.bc[0], g.bc[1], H),
                                         bmap(y, g.ny, g.bc[2], g.bc[3], H), 0), raw);
      to_recon_vars<RC, NC>(raw, v, P.gamma);
#pragma unroll
      for (int m = 0; m < NC; ++m) Vs[m * VSN + e] = v[m];
    }
    // operands of the RK combine, loaded now so their latency hides behind B and C
    double pq0[NC], px[NC];
    rk_prefetch<NC, 3>(a.stage, owner, gidx32(g, c, r, 0), g.cs, q0p, qin, acc, pq0, px);
    __syncthreads();

    // ---- B: vertex reconstructions (task e: own i, own j, then halo)
    for (int e = t; e < RT; e += NT) {
      int src, stride, dst;
      bool xdir;
      if (e < NT) {                       // own vertex, i-direction
        const int x = e % TX, y = e / TX;
        src = (y + H) * VW + x + H; stride = 1; xdir = true; dst = y * XRW + x + 1;
      } else if (e < 2 * NT) {            // own vertex, j-direction
        const int e1 = e - NT, x = e1 % TX, y = e1 / TX;
        src = (y + H) * VW + x + H; stride = VW; xdir = false; dst = (y + 1) * TX + x;
      } else if (e < 2 * NT + 2 * TY) {   // i-direction at columns i0-1 and i0+TX
        const int e2 = e - 2 * NT, y = e2 >> 1, side = e2 & 1;
        const int x = side ? TX : -1;
        src = (y + H) * VW + x + H; stride = 1; xdir = true; dst = y * XRW + x + 1;
      } else {                            // j-direction at rows j0-1 and j0+TY
        const int e3 = e - 2 * NT - 2 * TY, x = e3 % TX, side = e3 / TX;
        const int y = side ? TY : -1;
        src = (y + H) * VW + x + H; stride = VW; xdir = false; dst = (y + 1) * TX + x;
      }
      double W[2 * RV + 1][NC], L[NC], R[NC];
#pragma unroll
      for (int s = 0; s < 2 * RV + 1; ++s)
#pragma unroll
        for (int m = 0; m < NC; ++m) W[s][m] = Vs[m * VSN + src + (s - RV) * stride];
      RC::template vertex<NC>(W, P, L, R);
      if (xdir) {
#pragma unroll
        for (int m = 0; m < NC; ++m) {
          LRx[0][m * LRX + dst] = L[m];
          if constexpr (NLR == 2) LRx[NLR - 1][m * LRX + dst] = R[m];
        }
      } else {
#pragma unroll
        for (int m = 0; m < NC; ++m) {
          LRy[0][m * LRY + dst] = L[m];
          if constexpr (NLR == 2) LRy[NLR - 1][m * LRY + dst] = R[m];
        }
      }
    }
    __syncthreads();

    // ---- C: face fluxes (into the plane buffer, dead after B)
    for (int e = t; e < FT; e += NT) {
      double L[NC], R[NC], F[NC];
      if (e < NT || (e >= 2 * NT && e < 2 * NT + TY)) {   // i-face s of row y
        int y, s;
        if (e < NT) { s = e % TX; y = e / TX; } else { s = TX; y = e - 2 * NT; }
#pragma unroll
        for (int m = 0; m < NC; ++m) {
          L[m] = LRx[0][m * LRX + y * XRW + s];
          R[m] = LRx[NLR - 1][m * LRX + y * XRW + s + 1];
        }
        face_flux<FLUX, NC, 0>(L, R, P, F);
#pragma unroll
        for (int m = 0; m < NC; ++m) Fx[m * FXN + y * FXW + s] = F[m];
      } else {                                            // j-face s of column x
        int x, s;
        if (e < 2 * NT) { const int e1 = e - NT; x = e1 % TX; s = e1 / TX; }
        else { x = e - 2 * NT - TY; s = TY; }
#pragma unroll
        for (int m = 0; m < NC; ++m) {
          L[m] = LRy[0][m * LRY + s * TX + x];
          R[m] = LRy[NLR - 1][m * LRY + (s + 1) * TX + x];
        }
        face_flux<FLUX, NC, 1>(L, R, P, F);
#pragma unroll
        for (int m = 0; m < NC; ++m) Fy[m * FYN + s * TX + x] = F[m];
      }
    }
    __syncthreads();

    // ---- D: update
    double nrm[NC], smax = 0.0;
#pragma unroll
    for (int m = 0; m < NC; ++m) nrm[m] = 0.0;
    if (owner) {
      const uint32_t cidx = gidx32(g, c, r, 0);
      const Recip ix = kFast ? Recip{} : erecip(g.dx), iy = kFast ? Recip{} : erecip(g.dy);
      double rr[NC], out[NC];
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        const double fxd = Fx[m * FXN + ty * FXW + tx + 1] - Fx[m * FXN + ty * FXW + tx];
        const double fyd = Fy[m * FYN + (ty + 1) * TX + tx] - Fy[m * FYN + ty * TX + tx];
        if constexpr (kFast) {
          rr[m] = -(fxd * g.rdx) - fyd * g.rdy;
        } else {
          rr[m] = 0.0 - ediv(fxd, ix);
          rr[m] = rr[m] - ediv(fyd, iy);
        }
      }
      if (rvp) {
#pragma unroll
        for (int m = 1; m < NC; ++m) rr[m] = rr[m] + ld_comp(rvp, g.cs, m - 1, cidx);
      }
      rk_apply_all<NC>(a.stage, kc, pq0, px, rr, out, acc, g.cs, cidx);
#pragma unroll
      for (int m = 0; m < NC; ++m) nrm[m] = rr[m] * rr[m];
      store_with_ghosts<NC, 2, CONN>(g, tab, B, a.out_buf, qout, c, r, 0, out);
      if (want_smax) smax = wave_speed<NC>(out, g, P.gamma);
    }
    if (want_norm) {
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        const double s = wg_sum<NT>(nrm[m], red);
        if (t == 0) a.norm_part[(int64_t)tile * NC + m] = s;
      }
    }
    if (want_smax) {
      const double s = wg_max<NT>(smax, red);
      if (t == 0) smax_add(a.block_smax ? B.smax : a.smax, s);
    }
    __syncthreads();   // the next tile's plane overwrites the fluxes
  }
}

// ============================================================================
// 3-D stage: tile of TX (32 or 64) x TY columns, NT = TX * TY threads, every
// thread owns the column (i0 + tx, j0 + ty) and marches along k.  A
// workgroup loops over tiles of the launch's tile list (all blocks of the
// context in one launch).
//
// Per plane k:
//   A  convert the prefetched raw values: plane k (owned vertex + i/j halo)
//      into the LDS plane buffer, V(k+1+RV) into the owner's k-queue; issue
//      the loads for plane k+1 and V(k+2+RV) so they fly during the compute
//   B  vertex reconstructions: i-direction at columns i0-1..i0+TX of the TY
//      rows, j-direction at rows j0-1..j0+TY of the TX columns (flat task
//      list: each thread's own two vertices first, then the halo ring)
//   C  face fluxes: TX+1 i-faces per row, TY+1 j-faces per column (own faces
//      first, then the far edge), and the owner's k-face k+1 from its queue
//   D  owner update: residual, RK combine, store with ghost images
// ============================================================================
#ifndef FVM_WPE
#define FVM_WPE 2
#endif
template <class RC, int FLUX, int TX, int TY, bool CONN>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(FVM_WPE)))
void stage3d_kernel(StageArgs a) {
  constexpr int NC = 5, NT = TX * TY;
  constexpr int RV = RC::RV, H = RC::H;
  constexpr int NW = RV + (RV > 1 ? RV : 1) + 1;   // k-queue V(c - max(RV,1) .. c + RV)
  constexpr int QOFF = (RV > 1 ? RV : 1);          // queue index of the centre c
  constexpr int VW = TX + 2 * H, VR = TY + 2 * H, VSN = VW * VR;
  constexpr int NLR = RC::SAME ? 1 : 2;
  constexpr int XRW = TX + 2, LRX = TY * XRW;      // i-recon columns i0-1 .. i0+TX
  constexpr int LRY = (TY + 2) * TX;               // j-recon rows j0-1 .. j0+TY
  constexpr int FXW = TX + 1, FXN = TY * FXW, FYN = (TY + 1) * TX;
  constexpr int AREA = VSN > FXN + FYN ? VSN : FXN + FYN;
  constexpr int NXH = 2 * H * TY, NYH = 2 * H * TX, NHALO = NXH + NYH;
  constexpr int NHL = (NHALO + NT - 1) / NT;       // halo loads per thread
  constexpr int RT = 2 * NT + 2 * TY + 2 * TX;     // reconstruction tasks
  constexpr int FT = 2 * NT + TY + TX;             // face tasks
  // Balanced schedule (256-thread tiles): the halo reconstructions go to the
  // first waves while the others reconstruct along k, and the halo faces to
  // the last wave while the first ones reconstruct along k, so no wave runs a
  // third round of i/j work while the rest wait at the barrier.
  constexpr int kHaloWavesB = (RT - 2 * NT + 63) / 64;
  // (WENO's five-plane queue has no registers to spare for it.)
  constexpr bool kBalanced = RV <= 1 && NT == 256 && kHaloWavesB < NT / 64 &&
                             FT - 2 * NT <= 64 && kHaloWavesB < NT / 64 - 1;
  __shared__ double Vs[NC * AREA];
  __shared__ double LRx[NLR][NC * LRX];
  __shared__ double LRy[NLR][NC * LRY];
  __shared__ double red[NT / 64];
  __shared__ double NrmW[(NT / 64) * NC];   // per-wave rhs^2 sums (norm stage only)
  __shared__ double RcpD[3];       // exact family: reciprocals of the tile's dx, dy, dz
  double* Fx = Vs;                 // [NC][FXN] after phase B
  double* Fy = Vs + NC * FXN;      // [NC][FYN]

  const Params& P = a.P;
  const int t = threadIdx.x;
  const int tx = t % TX, ty = t / TX;
  const int wave = t / 64, lane = t % 64;
  const bool want_smax = a.smax != nullptr || a.block_smax;

  double dt_step = 0.0;
  bool want_norm = false;
  if (!stage_clock(a, dt_step, want_norm)) return;
  const DescTable tab = as_const_table(a.blocks);
  for (int it = blockIdx.x; it < a.ntiles; it += gridDim.x) {
    const int tile = a.tile0 + xcd_remap(it, a.ntiles);
    const int bi = block_of_tile(tab, a.nblocks, tile);
    const auto& B = tab[bi];
    if (B.halt && *B.halt) continue;   // uniform over the workgroup
    const BlockGeom g = load_geom(B.g);
    const double dt = a.cfl_dt ? dt_step : *B.dt_dev;
    const double* __restrict__ rvp = B.rv;
    const RKCoef kc = rk_coef(a.stage, dt);
    double smax = 0.0;
    const double* __restrict__ qin = B.q[a.in_buf];
    const double* __restrict__ q0p = B.q[a.q0_buf];
    double* __restrict__ qout = B.q[a.out_buf];
    double* __restrict__ acc = a.use_acc ? B.q[3] : nullptr;
    const int lt = tile - B.tile_base;
    const int ti = lt % B.tiles_x;
    const int tj = (lt / B.tiles_x) % B.tiles_y;
    const int tk = lt / (B.tiles_x * B.tiles_y);
    const int i0 = ti * TX, j0 = tj * TY;
    const int k0 = tk * B.chunk, k1 = min(k0 + B.chunk, g.nz);
    const int c = i0 + tx, r = j0 + ty;
    const bool owner = c < g.nx && r < g.ny;
    const int ci_ = bmap(c, g.nx, g.bc[0], g.bc[1], H), rj_ = bmap(r, g.ny, g.bc[2], g.bc[3], H);

    auto gaddr = [&](int i, int j, int k) {
      return gidx32(g, bmap(i, g.nx, g.bc[0], g.bc[1], H), bmap(j, g.ny, g.bc[2], g.bc[3], H),
                    bmap(k, g.nz, g.bc[4], g.bc[5], H));
    };
    // halo slot e (0..NHALO-1) of a plane: global (i, j) and LDS offset
    auto halo_pos = [&](int e, int& hi, int& hj, int& slot) {
      if (e < NXH) {
        const int row = e / (2 * H), s = e % (2 * H);
        hi = s < H ? i0 - H + s : i0 + TX + (s - H);
        hj = j0 + row;
        slot = (row + H) * VW + (s < H ? s : TX + H + (s - H));
      } else {
        const int e2 = e - NXH;
        const int col = e2 % TX, s = e2 / TX;
        hi = i0 + col;
        hj = s < H ? j0 - H + s : j0 + TY + (s - H);
        slot = (s < H ? s : TY + H + (s - H)) * VW + col + H;
      }
    };

    // ---- k-queue warm-up: vertices k0-1 and k0, face k0
    double Vq[NW][NC];
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      double raw[NC];
      load_raw<RC, NC>(qin, g.cs, gaddr(ci_, rj_, k0 - 1 - QOFF + w), raw);
      to_recon_vars<RC, NC>(raw, Vq[w], P.gamma);
    }
    double Lc[NC], Fc[NC], Lt[NC], Rt[NC];
    RC::template vertex<NC>(&Vq[QOFF - RV], P, Lt, Rt);
#pragma unroll
    for (int m = 0; m < NC; ++m) Lc[m] = Lt[m];
    {
      double raw[NC];
#pragma unroll
      for (int w = 0; w < NW - 1; ++w)
#pragma unroll
        for (int m = 0; m < NC; ++m) Vq[w][m] = Vq[w + 1][m];
      load_raw<RC, NC>(qin, g.cs, gaddr(ci_, rj_, k0 + RV), raw);
      to_recon_vars<RC, NC>(raw, Vq[NW - 1], P.gamma);
    }
    RC::template vertex<NC>(&Vq[QOFF - RV], P, Lt, Rt);
    face_flux<FLUX, NC, 2>(Lc, Rt, P, Fc);
#pragma unroll
    for (int m = 0; m < NC; ++m) Lc[m] = Lt[m];

    // ---- prefetch for plane k0
    double oraw[NC];
    load_raw<RC, NC>(qin, g.cs, gaddr(ci_, rj_, k0 + 1 + RV), oraw);
    double hraw[NHL][NC];
    int hslot[NHL];
#pragma unroll
    for (int h = 0; h < NHL; ++h) {
      const int e = t + h * NT;
      hslot[h] = -1;
      if (e < NHALO) {
        int hi, hj;
        halo_pos(e, hi, hj, hslot[h]);
        load_raw<RC, NC>(qin, g.cs, gaddr(hi, hj, k0), hraw[h]);
      }
    }

    if (want_norm && t < (NT / 64) * NC) NrmW[t] = 0.0;
    // the divisors' reciprocals, formed once per tile instead of per plane
    // (read in D, after this plane's first barrier)
    if (!kFast && t < 3) RcpD[t] = erecip(t == 0 ? g.dx : (t == 1 ? g.dy : g.dz)).r;

    for (int k = k0; k < k1; ++k) {
      // ---- A: queue shift (centre k+1), plane k into LDS, next loads
#pragma unroll
      for (int w = 0; w < NW - 1; ++w)
#pragma unroll
        for (int m = 0; m < NC; ++m) Vq[w][m] = Vq[w + 1][m];
      to_recon_vars<RC, NC>(oraw, Vq[NW - 1], P.gamma);
#pragma unroll
      for (int m = 0; m < NC; ++m) Vs[m * VSN + (ty + H) * VW + tx + H] = Vq[QOFF - 1][m];
#pragma unroll
      for (int h = 0; h < NHL; ++h) {
        if (hslot[h] >= 0) {
          double v[NC];
          to_recon_vars<RC, NC>(hraw[h], v, P.gamma);
#pragma unroll
          for (int m = 0; m < NC; ++m) Vs[m * VSN + hslot[h]] = v[m];
        }
      }
      if (k + 1 < k1) {
        load_raw<RC, NC>(qin, g.cs, gaddr(ci_, rj_, k + 2 + RV), oraw);
#pragma unroll
        for (int h = 0; h < NHL; ++h) {
          const int e = t + h * NT;
          if (e < NHALO) {
            int hi, hj, sl;
            halo_pos(e, hi, hj, sl);
            load_raw<RC, NC>(qin, g.cs, gaddr(hi, hj, k + 1), hraw[h]);
          }
        }
      }
      __syncthreads();

      // ---- B: vertex reconstructions (task e: own i, own j, then halo)
      auto recon_task = [&](int e) {
        int src, stride, dst;
        bool xdir;
        if (e < NT) {                       // own vertex, i-direction
          const int x = e % TX, y = e / TX;
          src = (y + H) * VW + x + H; stride = 1; xdir = true; dst = y * XRW + x + 1;
        } else if (e < 2 * NT) {            // own vertex, j-direction
          const int e1 = e - NT, x = e1 % TX, y = e1 / TX;
          src = (y + H) * VW + x + H; stride = VW; xdir = false; dst = (y + 1) * TX + x;
        } else if (e < 2 * NT + 2 * TY) {   // i-direction at columns i0-1 and i0+TX
          const int e2 = e - 2 * NT, y = e2 >> 1, side = e2 & 1;
          const int x = side ? TX : -1;
          src = (y + H) * VW + x + H; stride = 1; xdir = true; dst = y * XRW + x + 1;
        } else {                            // j-direction at rows j0-1 and j0+TY
          const int e3 = e - 2 * NT - 2 * TY, x = e3 % TX, side = e3 / TX;
          const int y = side ? TY : -1;
          src = (y + H) * VW + x + H; stride = VW; xdir = false; dst = (y + 1) * TX + x;
        }
        double W[2 * RV + 1][NC], L[NC], R[NC];
#pragma unroll
        for (int s = 0; s < 2 * RV + 1; ++s)
#pragma unroll
          for (int m = 0; m < NC; ++m) W[s][m] = Vs[m * VSN + src + (s - RV) * stride];
        RC::template vertex<NC>(W, P, L, R);
        if (xdir) {
#pragma unroll
          for (int m = 0; m < NC; ++m) {
            LRx[0][m * LRX + dst] = L[m];
            if constexpr (NLR == 2) LRx[NLR - 1][m * LRX + dst] = R[m];
          }
        } else {
#pragma unroll
          for (int m = 0; m < NC; ++m) {
            LRy[0][m * LRY + dst] = L[m];
            if constexpr (NLR == 2) LRy[NLR - 1][m * LRY + dst] = R[m];
          }
        }
      };
      // k-direction vertex k+1 and face k+1 of the owned column (registers
      // only); afterwards Lc holds L(k+1) for the next plane's face
      double Fn[NC];
      auto zstep = [&]() {
        RC::template vertex<NC>(&Vq[QOFF - RV], P, Lt, Rt);
        face_flux<FLUX, NC, 2>(Lc, Rt, P, Fn);
#pragma unroll
        for (int m = 0; m < NC; ++m) Lc[m] = Lt[m];
      };
      if constexpr (kBalanced) {
        // the waves with the longer share of a phase (here the k-direction
        // work, in C the halo faces and the k-work) issue first, so the
        // other workgroup's waves fill in while the short ones wait at the
        // barrier (measured 1.5-2 % per launch)
        if (wave >= kHaloWavesB) __builtin_amdgcn_s_setprio(3);
        recon_task(t);
        recon_task(t + NT);
        if (wave < kHaloWavesB) {
          if (t < RT - 2 * NT) recon_task(2 * NT + t);
        } else {
          zstep();
        }
        __builtin_amdgcn_s_setprio(0);
      } else {
        for (int e = t; e < RT; e += NT) recon_task(e);
      }
      __syncthreads();

      // operands of the RK combine of vertex (c, r, k), loaded now so their
      // latency hides behind the face fluxes
      // (all of them where registers allow; q0 only in the balanced schedule;
      // none with WENO's five-plane queue: the rest is loaded in D)
      constexpr int kPrefetchRK = RC::RV > 1 ? 0 : 3;
      double pq0[NC], px[NC];
      if constexpr (kPrefetchRK != 0)
        rk_prefetch<NC, kPrefetchRK>(a.stage, owner, gidx32(g, c, r, k), g.cs, q0p, qin, acc, pq0, px);

      // ---- C: face fluxes (into the plane buffer, dead after B)
      auto face_task = [&](int e) {
        double L[NC], R[NC], F[NC];
        if (e < NT || (e >= 2 * NT && e < 2 * NT + TY)) {   // i-face s of row y
          int y, s;
          if (e < NT) { s = e % TX; y = e / TX; } else { s = TX; y = e - 2 * NT; }
#pragma unroll
          for (int m = 0; m < NC; ++m) {
            L[m] = LRx[0][m * LRX + y * XRW + s];
            R[m] = LRx[NLR - 1][m * LRX + y * XRW + s + 1];
          }
          face_flux<FLUX, NC, 0>(L, R, P, F);
#pragma unroll
          for (int m = 0; m < NC; ++m) Fx[m * FXN + y * FXW + s] = F[m];
        } else {                                            // j-face s of column x
          int x, s;
          if (e < 2 * NT) { const int e1 = e - NT; x = e1 % TX; s = e1 / TX; }
          else { x = e - 2 * NT - TY; s = TY; }
#pragma unroll
          for (int m = 0; m < NC; ++m) {
            L[m] = LRy[0][m * LRY + s * TX + x];
            R[m] = LRy[NLR - 1][m * LRY + (s + 1) * TX + x];
          }
          face_flux<FLUX, NC, 1>(L, R, P, F);
#pragma unroll
          for (int m = 0; m < NC; ++m) Fy[m * FYN + s * TX + x] = F[m];
        }
      };
      if constexpr (kBalanced) {
        if (wave < kHaloWavesB || wave == NT / 64 - 1) __builtin_amdgcn_s_setprio(3);
        face_task(t);
        face_task(t + NT);
        if (wave < kHaloWavesB) {
          zstep();
        } else if (wave == NT / 64 - 1) {
          if (t - (NT - 64) < FT - 2 * NT) face_task(2 * NT + t - (NT - 64));
        }
        __builtin_amdgcn_s_setprio(0);
      } else {
        for (int e = t; e < FT; e += NT) face_task(e);
        zstep();
      }
      __syncthreads();

      // ---- D: update
      double rr[NC];
#pragma unroll
      for (int m = 0; m < NC; ++m) rr[m] = 0.0;
      if (owner) {
        const uint32_t cidx = gidx32(g, c, r, k);
        const Recip ix = kFast ? Recip{} : Recip{g.dx, RcpD[0]};
        const Recip iy = kFast ? Recip{} : Recip{g.dy, RcpD[1]};
        const Recip iz = kFast ? Recip{} : Recip{g.dz, RcpD[2]};
        double out[NC];
#pragma unroll
        for (int m = 0; m < NC; ++m) {
          const double fxd = Fx[m * FXN + ty * FXW + tx + 1] - Fx[m * FXN + ty * FXW + tx];
          const double fyd = Fy[m * FYN + (ty + 1) * TX + tx] - Fy[m * FYN + ty * TX + tx];
          if constexpr (kFast) {
            rr[m] = -(fxd * g.rdx) - fyd * g.rdy - (Fn[m] - Fc[m]) * g.rdz;
          } else {
            rr[m] = 0.0 - ediv(fxd, ix);
            rr[m] = rr[m] - ediv(fyd, iy);
            rr[m] = rr[m] - ediv(Fn[m] - Fc[m], iz);
          }
          if constexpr (kPrefetchRK != 3) {
            if (m == 0) rk_prefetch<NC, 3 - kPrefetchRK>(a.stage, true, cidx, g.cs, q0p, qin, acc, pq0, px);
          }
        }
        if (rvp) {
#pragma unroll
          for (int m = 1; m < NC; ++m) rr[m] = rr[m] + ld_comp(rvp, g.cs, m - 1, cidx);
        }
        rk_apply_all<NC>(a.stage, kc, pq0, px, rr, out, acc, g.cs, cidx);
        store_with_ghosts<NC, 3, CONN>(g, tab, B, a.out_buf, qout, c, r, k, out);
        if (want_smax) smax = nan_max(smax, wave_speed<NC>(out, g, P.gamma));
      }
      if (want_norm) {   // uniform branch: each wave's sum of rhs^2 over its plane
#pragma unroll
        for (int m = 0; m < NC; ++m) {
          double v = rr[m] * rr[m];
          for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
          if (lane == 0) NrmW[wave * NC + m] += v;
        }
      }
#pragma unroll
      for (int m = 0; m < NC; ++m) Fc[m] = Fn[m];
      __syncthreads();
    }

    if (want_norm) {
      if (t == 0) {
#pragma unroll
        for (int m = 0; m < NC; ++m) {
          double s = 0.0;
          for (int w = 0; w < NT / 64; ++w) s += NrmW[w * NC + m];
          a.norm_part[(int64_t)tile * NC + m] = s;
        }
      }
      __syncthreads();   // the next tile clears the sums
    }
    if (want_smax) {
      const double s = wg_max<NT>(smax, red);
      if (t == 0) smax_add(a.block_smax ? B.smax : a.smax, s);
    }
  }
}


// ============================================================================
// 3-D stage, register-reconstruction layout (simple and MUSCL stencils): the
// i/j vertex reconstructions never pass through LDS, so a workgroup holds the
// plane and the two flux planes only (about half of stage3d_kernel's LDS).
// Per plane k:
//   A  as stage3d_kernel: plane k (owned vertex + i/j halo) into LDS as
//      reconstruction variables, k-queue shift, next plane's loads in flight
//   BC face tasks: the thread of an i-face (j-face) reads the 2 RV + 2 plane
//      values across it, reconstructs the face's two vertices in registers
//      and writes the flux; each owner also takes its k-face from its queue
//   D  owner update as stage3d_kernel; norm partials summed per wave
// Same per-vertex and per-face arithmetic as stage3d_kernel (results bit for
// bit), about twice its i/j reconstruction work.
// ============================================================================
#ifndef FVM_WPE_REG
#define FVM_WPE_REG 3
#endif
template <class RC, int FLUX, int TX, int TY, bool CONN>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(FVM_WPE_REG)))
void stage3d_reg_kernel(StageArgs a) {
  constexpr int NC = 5, NT = TX * TY, NWAVE = NT / 64;
  constexpr int RV = RC::RV, H = RC::H;
  static_assert(RV <= 1, "register reconstruction takes the simple and MUSCL stencils");
  constexpr int NW = RV + 2;                       // k-queue V(c - 1 .. c + RV)
  constexpr int QOFF = 1;                          // queue index of the centre c
  constexpr int VW = TX + 2 * H, VR = TY + 2 * H, VSN = VW * VR;
  constexpr int FXW = TX + 1, FXN = TY * FXW, FYN = (TY + 1) * TX;
  constexpr int NXH = 2 * H * TY, NYH = 2 * H * TX, NHALO = NXH + NYH;
  constexpr int NHL = (NHALO + NT - 1) / NT;       // halo loads per thread
  constexpr int FT = FXN + FYN;                    // i- then j-face tasks
  constexpr int NWIN = 2 * RV + 2;                 // plane values across a face
  __shared__ double Vs[NC * VSN];
  __shared__ double Fx[NC * FXN];
  __shared__ double Fy[NC * FYN];
  __shared__ double red[NWAVE];
  __shared__ double NrmW[NWAVE * NC];              // per-wave rhs^2 sums (norm stage)
  __shared__ double Fk[NC * NT];                   // the owner's k-face flux k (own slot)

  const Params& P = a.P;
  const int t = threadIdx.x;
  const int tx = t % TX, ty = t / TX;
  const int wave = t / 64, lane = t % 64;
  const bool want_smax = a.smax != nullptr || a.block_smax;

  double dt_step = 0.0;
  bool want_norm = false;
  if (!stage_clock(a, dt_step, want_norm)) return;
  const DescTable tab = as_const_table(a.blocks);
  for (int it = blockIdx.x; it < a.ntiles; it += gridDim.x) {
    const int tile = a.tile0 + xcd_remap(it, a.ntiles);
    const int bi = block_of_tile(tab, a.nblocks, tile);
    const auto& B = tab[bi];
    if (B.halt && *B.halt) continue;   // uniform over the workgroup
    const BlockGeom g = load_geom(B.g);
    const double dt = a.cfl_dt ? dt_step : *B.dt_dev;
    const double* __restrict__ rvp = B.rv;
    const RKCoef kc = rk_coef(a.stage, dt);
    double smax = 0.0;
    const double* __restrict__ qin = B.q[a.in_buf];
    const double* __restrict__ q0p = B.q[a.q0_buf];
    double* __restrict__ qout = B.q[a.out_buf];
    double* __restrict__ acc = a.use_acc ? B.q[3] : nullptr;
    const int lt = tile - B.tile_base;
    const int ti = lt % B.tiles_x;
    const int tj = (lt / B.tiles_x) % B.tiles_y;
    const int tk = lt / (B.tiles_x * B.tiles_y);
    const int i0 = ti * TX, j0 = tj * TY;
    const int k0 = tk * B.chunk, k1 = min(k0 + B.chunk, g.nz);
    const int c = i0 + tx, r = j0 + ty;
    const bool owner = c < g.nx && r < g.ny;
    const int ci_ = bmap(c, g.nx, g.bc[0], g.bc[1], H), rj_ = bmap(r, g.ny, g.bc[2], g.bc[3], H);

    auto gaddr = [&](int i, int j, int k) {
      return gidx32(g, bmap(i, g.nx, g.bc[0], g.bc[1], H), bmap(j, g.ny, g.bc[2], g.bc[3], H),
                    bmap(k, g.nz, g.bc[4], g.bc[5], H));
    };
    // halo slot e (0..NHALO-1) of a plane: global (i, j) and LDS offset
    auto halo_pos = [&](int e, int& hi, int& hj, int& slot) {
      if (e < NXH) {
        const int row = e / (2 * H), s = e % (2 * H);
        hi = s < H ? i0 - H + s : i0 + TX + (s - H);
        hj = j0 + row;
        slot = (row + H) * VW + (s < H ? s : TX + H + (s - H));
      } else {
        const int e2 = e - NXH;
        const int col = e2 % TX, s = e2 / TX;
        hi = i0 + col;
        hj = s < H ? j0 - H + s : j0 + TY + (s - H);
        slot = (s < H ? s : TY + H + (s - H)) * VW + col + H;
      }
    };

    // ---- k-queue warm-up: vertices k0-1 and k0, face k0
    double Vq[NW][NC];
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      double raw[NC];
      load_raw<RC, NC>(qin, g.cs, gaddr(ci_, rj_, k0 - 1 - QOFF + w), raw);
      to_recon_vars<RC, NC>(raw, Vq[w], P.gamma);
    }
    double Lc[NC], Lt[NC], Rt[NC];
    RC::template vertex<NC>(&Vq[QOFF - RV], P, Lt, Rt);
#pragma unroll
    for (int m = 0; m < NC; ++m) Lc[m] = Lt[m];
    {
      double raw[NC];
#pragma unroll
      for (int w = 0; w < NW - 1; ++w)
#pragma unroll
        for (int m = 0; m < NC; ++m) Vq[w][m] = Vq[w + 1][m];
      load_raw<RC, NC>(qin, g.cs, gaddr(ci_, rj_, k0 + RV), raw);
      to_recon_vars<RC, NC>(raw, Vq[NW - 1], P.gamma);
    }
    RC::template vertex<NC>(&Vq[QOFF - RV], P, Lt, Rt);
    {
      double Fc[NC];
      face_flux<FLUX, NC, 2>(Lc, Rt, P, Fc);
#pragma unroll
      for (int m = 0; m < NC; ++m) Fk[m * NT + t] = Fc[m];
    }
#pragma unroll
    for (int m = 0; m < NC; ++m) Lc[m] = Lt[m];

    // ---- prefetch for plane k0
    double oraw[NC];
    load_raw<RC, NC>(qin, g.cs, gaddr(ci_, rj_, k0 + 1 + RV), oraw);
    double hraw[NHL][NC];
    int hslot[NHL];
#pragma unroll
    for (int h = 0; h < NHL; ++h) {
      const int e = t + h * NT;
      hslot[h] = -1;
      if (e < NHALO) {
        int hi, hj;
        halo_pos(e, hi, hj, hslot[h]);
        load_raw<RC, NC>(qin, g.cs, gaddr(hi, hj, k0), hraw[h]);
      }
    }
    if (want_norm && t < NWAVE * NC) NrmW[t] = 0.0;

    for (int k = k0; k < k1; ++k) {
      // ---- A: queue shift (centre k+1), plane k into LDS, next loads
#pragma unroll
      for (int w = 0; w < NW - 1; ++w)
#pragma unroll
        for (int m = 0; m < NC; ++m) Vq[w][m] = Vq[w + 1][m];
      to_recon_vars<RC, NC>(oraw, Vq[NW - 1], P.gamma);
#pragma unroll
      for (int m = 0; m < NC; ++m) Vs[m * VSN + (ty + H) * VW + tx + H] = Vq[QOFF - 1][m];
#pragma unroll
      for (int h = 0; h < NHL; ++h) {
        if (hslot[h] >= 0) {
          double v[NC];
          to_recon_vars<RC, NC>(hraw[h], v, P.gamma);
#pragma unroll
          for (int m = 0; m < NC; ++m) Vs[m * VSN + hslot[h]] = v[m];
        }
      }
      __syncthreads();

      // ---- BC: face tasks.  i-face s of row y lies between vertices s - 1
      // and s of the row (s = 0..TX); j-face s of column x likewise.
      auto face_task = [&](int e) {
        int base, stride, out;
        const bool xdir = e < FXN;
        if (xdir) {
          const int y = e / FXW, s = e % FXW;
          base = (y + H) * VW + s - 1 - RV + H; stride = 1; out = e;
        } else {
          const int e1 = e - FXN, x = e1 % TX, s = e1 / TX;
          base = (s - 1 - RV + H) * VW + x + H; stride = VW; out = e1;
        }
        double W[NWIN][NC], La[NC], Ra[NC], Lb[NC], Rb[NC], F[NC];
#pragma unroll
        for (int w = 0; w < NWIN; ++w)
#pragma unroll
          for (int m = 0; m < NC; ++m) W[w][m] = Vs[m * VSN + base + w * stride];
        RC::template vertex<NC>(&W[0], P, La, Ra);    // vertex s - 1
        RC::template vertex<NC>(&W[1], P, Lb, Rb);    // vertex s
        if (xdir) {
          face_flux<FLUX, NC, 0>(La, Rb, P, F);
#pragma unroll
          for (int m = 0; m < NC; ++m) Fx[m * FXN + out] = F[m];
        } else {
          face_flux<FLUX, NC, 1>(La, Rb, P, F);
#pragma unroll
          for (int m = 0; m < NC; ++m) Fy[m * FYN + out] = F[m];
        }
      };
      for (int e = t; e < FT; e += NT) face_task(e);
      // the next plane's loads, in flight behind the k-face and the update
      // (not live across the face tasks, where registers are the limit)
      if (k + 1 < k1) {
        load_raw<RC, NC>(qin, g.cs, gaddr(ci_, rj_, k + 2 + RV), oraw);
#pragma unroll
        for (int h = 0; h < NHL; ++h) {
          const int e = t + h * NT;
          if (e < NHALO) {
            int hi, hj, sl;
            halo_pos(e, hi, hj, sl);
            load_raw<RC, NC>(qin, g.cs, gaddr(hi, hj, k + 1), hraw[h]);
          }
        }
      }
      // k-direction vertex k+1 and face k+1 of the owned column (registers)
      double Fn[NC];
      RC::template vertex<NC>(&Vq[QOFF - RV], P, Lt, Rt);
      face_flux<FLUX, NC, 2>(Lc, Rt, P, Fn);
#pragma unroll
      for (int m = 0; m < NC; ++m) Lc[m] = Lt[m];
      __syncthreads();

      // ---- D: update
      double rr[NC];
#pragma unroll
      for (int m = 0; m < NC; ++m) rr[m] = 0.0;
      if (owner) {
        const uint32_t cidx = gidx32(g, c, r, k);
        const Recip ix = kFast ? Recip{} : erecip(g.dx), iy = kFast ? Recip{} : erecip(g.dy);
        const Recip iz = kFast ? Recip{} : erecip(g.dz);
        // operands of the RK combine (loaded here: registers are the limit
        // during the face tasks, and the other waves cover the latency)
        double pq0[NC], px[NC], out[NC];
        rk_prefetch<NC, 3>(a.stage, true, cidx, g.cs, q0p, qin, acc, pq0, px);
#pragma unroll
        for (int m = 0; m < NC; ++m) {
          const double fxd = Fx[m * FXN + ty * FXW + tx + 1] - Fx[m * FXN + ty * FXW + tx];
          const double fyd = Fy[m * FYN + (ty + 1) * TX + tx] - Fy[m * FYN + ty * TX + tx];
          const double fzd = Fn[m] - Fk[m * NT + t];
          if constexpr (kFast) {
            rr[m] = -(fxd * g.rdx) - fyd * g.rdy - fzd * g.rdz;
          } else {
            rr[m] = 0.0 - ediv(fxd, ix);
            rr[m] = rr[m] - ediv(fyd, iy);
            rr[m] = rr[m] - ediv(fzd, iz);
          }
        }
        if (rvp) {
#pragma unroll
          for (int m = 1; m < NC; ++m) rr[m] = rr[m] + ld_comp(rvp, g.cs, m - 1, cidx);
        }
        rk_apply_all<NC>(a.stage, kc, pq0, px, rr, out, acc, g.cs, cidx);
        store_with_ghosts<NC, 3, CONN>(g, tab, B, a.out_buf, qout, c, r, k, out);
        if (want_smax) smax = nan_max(smax, wave_speed<NC>(out, g, P.gamma));
      }
      if (want_norm) {   // uniform branch
#pragma unroll
        for (int m = 0; m < NC; ++m) {
          double v = rr[m] * rr[m];
          for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
          if (lane == 0) NrmW[wave * NC + m] += v;
        }
      }
#pragma unroll
      for (int m = 0; m < NC; ++m) Fk[m * NT + t] = Fn[m];
      __syncthreads();
    }

    if (want_norm && t == 0) {
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        double s = 0.0;
        for (int w = 0; w < NWAVE; ++w) s += NrmW[w * NC + m];
        a.norm_part[(int64_t)tile * NC + m] = s;
      }
    }
    if (want_smax) {
      const double s = wg_max<NT>(smax, red);
      if (t == 0) smax_add(a.block_smax ? B.smax : a.smax, s);
    }
    __syncthreads();   // the next tile's plane and norm sums
  }
}

}  // namespace exact / fast
}  // namespace fvm
