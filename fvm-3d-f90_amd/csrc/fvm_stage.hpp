// fvm_stage.hpp -- fused RK-stage kernels for gfx950.
//
// One launch computes, for every real vertex of a block, the inviscid
// residual of all swept directions and the Runge-Kutta combine of that stage:
//   rhs = ((0 - dFx/dx) - dFy/dy) - dFz/dz          (class_space_discret.f90:678-732)
//   out = RK combine(q0, qin, acc, rhs)             (scheme_Runge_Kutta.f90:30-63)
// and scatters the new values into the output's ghost layers, so the next
// stage reads valid ghosts without a separate launch.
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

// ---------------------------------------------------------------- geometry
struct BlockGeom {
  int nx, ny, nz;        // real vertices (nz = 1 in 2-D)
  int ox, oy, oz;        // array offset of real vertex 0 along i, j, k
  int64_t px, pxy, cs;   // row pitch, plane pitch, component stride (elements)
  double dx, dy, dz;
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

__device__ __forceinline__ int64_t gidx(const BlockGeom& g, int i, int j, int k) {
  return (int64_t)(k + g.oz) * g.pxy + (int64_t)(j + g.oy) * g.px + (i + g.ox);
}

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of tiles so neighbouring tiles share an L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
  const int nx = 8;
  const int per = (nblocks + nx - 1) / nx;
  const int xcd = bid % nx, local = bid / nx;
  const int full = nblocks % nx == 0 ? nx : nblocks % nx;   // XCDs with `per` blocks
  const int base = xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1);
  return base + local;
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

// ---------------------------------------------------------------- recon policies
// RV: vertex reconstruction radius; H = RV + 1: stencil radius of a cell.
// CONS: the reconstructed variables are conserved (converted at the face).
// SAME: both face values of a vertex are the same number.
struct ReconSimple {   // scheme_reconstruct_simple.f90:42-85
  static constexpr int RV = 0, H = 1;
  static constexpr bool CONS = false, SAME = true;
  template <int NC>
  __device__ static void vertex(const double (*V)[NC], const Params&, double L[NC], double R[NC]) {
#pragma unroll
    for (int c = 0; c < NC; ++c) { L[c] = V[0][c]; R[c] = V[0][c]; }
  }
};

struct ReconMuscl {    // scheme_reconstruct_MUSCL.f90:47-226
  static constexpr int RV = 1, H = 2;
  static constexpr bool CONS = false, SAME = false;
  template <int NC>
  __device__ static void vertex(const double (*V)[NC], const Params& P, double L[NC], double R[NC]) {
#pragma unroll
    for (int c = 0; c < NC; ++c) muscl_vertex(V[0][c], V[1][c], V[2][c], P.kappa, P.limiter, L[c], R[c]);
  }
};

struct ReconWeno {     // scheme_WENO.f90:28-121, then q -> p at the face
  static constexpr int RV = 2, H = 3;
  static constexpr bool CONS = true, SAME = false;
  template <int NC>
  __device__ static void vertex(const double (*V)[NC], const Params& P, double L[NC], double R[NC]) {
    double qL[NC], qR[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      qL[c] = weno_js(V[0][c], V[1][c], V[2][c], V[3][c], V[4][c], P.weno_textbook);
      // reference: the right state of a face is the next face's left state
      // (scheme_WENO.f90:54-56); textbook: mirror image of the stencil
      qR[c] = P.weno_textbook ? weno_js(V[4][c], V[3][c], V[2][c], V[1][c], V[0][c], 1) : qL[c];
    }
    cons_to_prim<NC>(qL, L, P.gamma);
    if (P.weno_textbook) cons_to_prim<NC>(qR, R, P.gamma);
    else {
#pragma unroll
      for (int c = 0; c < NC; ++c) R[c] = L[c];
    }
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
__device__ __forceinline__ void load_raw(const double* __restrict__ q, int64_t cs, int64_t idx,
                                         double v[NC]) {
#pragma unroll
  for (int c = 0; c < NC; ++c) v[c] = q[idx + c * cs];
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

// Wave speed sum_d (|u_d| + c) / dr_d of a conserved state (CFL time step).
template <int NC>
__device__ __forceinline__ double wave_speed(const double q[NC], const BlockGeom& g, double gamma) {
  double p[NC];
  cons_to_prim<NC>(q, p, gamma);
  const double a = sqrt(gamma * p[NC - 1] / p[0]);
  double s = (fabs(p[1]) + a) / g.dx + (fabs(p[2]) + a) / g.dy;
  if constexpr (NC == 5) s = s + (fabs(p[3]) + a) / g.dz;
  return s;
}

// Write one output vertex and its ghost images along every swept direction.
template <int NC, int NDIR>
__device__ __forceinline__ void store_with_ghosts(const BlockGeom& g, double* __restrict__ out,
                                                  int i, int j, int k, const double v[NC]) {
  const int64_t base = gidx(g, i, j, k);
#pragma unroll
  for (int c = 0; c < NC; ++c) out[base + c * g.cs] = v[c];
  const int n[3] = {g.nx, g.ny, g.nz};
  const int x[3] = {i, j, k};
#pragma unroll
  for (int d = 0; d < NDIR; ++d) {
    const int64_t stride = d == 0 ? 1 : (d == 1 ? g.px : g.pxy);
    for (int l = 1; l <= g.ng; ++l) {
      // low side
      int tgt = -1000000;
      if (g.bc[2 * d] == BC_PERIODIC) { if (x[d] == n[d] - l) tgt = -l; }
      else if (x[d] == 0) tgt = -l;
      if (tgt != -1000000) {
        const int64_t o = base + (int64_t)(tgt - x[d]) * stride;
#pragma unroll
        for (int c = 0; c < NC; ++c) out[o + c * g.cs] = v[c];
      }
      // high side
      tgt = -1000000;
      if (g.bc[2 * d + 1] == BC_PERIODIC) { if (x[d] == l - 1) tgt = n[d] - 1 + l; }
      else if (x[d] == n[d] - 1) tgt = n[d] - 1 + l;
      if (tgt != -1000000) {
        const int64_t o = base + (int64_t)(tgt - x[d]) * stride;
#pragma unroll
        for (int c = 0; c < NC; ++c) out[o + c * g.cs] = v[c];
      }
    }
  }
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
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    s = red[0];
    for (int w = 1; w < NT / 64; ++w) {
      const double x = red[w];
      s = (x > s || x != x) ? x : s;   // NaN wins
    }
  }
  __syncthreads();
  return s;
}

// ============================================================================
// 2-D stage: tile of TX = 62 columns, one wave, marching along j.
// Thread t owns column c = i0 - 1 + t (t = 0..63): it keeps the j-queue of
// its column, reconstructs its vertex along i, and for t in 1..62 updates
// cell c.
// ============================================================================
template <class RC, int FLUX>
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

  const BlockGeom& g = a.g;
  const Params& P = a.P;
  const int t = threadIdx.x;
  const int bid = xcd_remap(blockIdx.x, gridDim.x);
  const int ti = bid % a.tiles_x, tj = bid / a.tiles_x;
  const int i0 = ti * TX;
  const int j0 = tj * a.chunk;
  const int j1 = min(j0 + a.chunk, g.ny);
  const int c = i0 - 1 + t;                       // owned column
  const int cl = clampi(c, -H, g.nx + H - 1);
  const bool owner = (t >= 1) && (t <= TX) && (c < g.nx);
  const double dt = a.dt_dev ? *a.dt_dev : a.dt;
  const RKCoef kc = rk_coef(a.stage, dt);

  // j-queue of column c: Vq[w] = V(c, jc - QOFF + w), jc = vertex being reconstructed
  double Vq[NW][NC];
  auto loadV = [&](int j, double v[NC]) {
    double raw[NC];
    load_raw<RC, NC>(a.qin, g.cs, gidx(g, cl, clampi(j, -H, g.ny + H - 1), 0), raw);
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
  double smax = 0.0;

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
      load_raw<RC, NC>(a.qin, g.cs, gidx(g, clampi(hc, -H, g.nx + H - 1), j, 0), raw);
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
      double out[NC];
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        double r = 0.0 - (Fx[m][t] - Fx[m][t - 1]) / g.dx;
        r = r - (Fn[m] - Fc[m]) / g.dy;
        const double q0 = a.q0[ci + m * g.cs];
        double* accp = a.acc ? a.acc + ci + m * g.cs : nullptr;
        out[m] = rk_apply(a.stage, kc, q0, a.qin[ci + m * g.cs], accp, r);
        nrm[m] += r * r;
      }
      store_with_ghosts<NC, 2>(g, a.qout, c, j, 0, out);
      if (a.smax) smax = fmax(smax, wave_speed<NC>(out, g, P.gamma));
    }
#pragma unroll
    for (int m = 0; m < NC; ++m) { Fc[m] = Fn[m]; Lc[m] = Lt[m]; }
    __syncthreads();
  }

  if (a.norm_part) {
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      const double s = wg_sum<NT>(nrm[m], red);
      if (t == 0) a.norm_part[(int64_t)blockIdx.x * NC + m] = s;
    }
  }
  if (a.smax) {
    const double s = wg_max<NT>(smax, red);
    if (t == 0) atomicMax(a.smax, (unsigned long long)__double_as_longlong(s));
  }
}

// ============================================================================
// 3-D stage: tile TX = 62 by TY rows; NT = 64 * TY threads; march along k.
// Thread (tx, ty) owns column (i0 - 1 + tx, j0 + ty).
// ============================================================================
template <class RC, int FLUX, int TY>
__global__ __launch_bounds__(64 * TY) void stage3d_kernel(StageArgs a) {
  constexpr int NC = 5, TX = 62, NT = 64 * TY;
  constexpr int RV = RC::RV, H = RC::H;
  constexpr int NW = RV + (RV > 1 ? RV : 1) + 1;
  constexpr int QOFF = (RV > 1 ? RV : 1);
  constexpr int VW = 64 + 2 * RV;            // Vs row width: columns i0-1-RV .. i0+TX+RV
  constexpr int VR = TY + 2 * H;             // Vs rows: j0-H .. j0+TY+H-1
  constexpr int NLR = RC::SAME ? 1 : 2;
  // x-reconstructions [TY][64], y-reconstructions [TY+2][TX]; the face fluxes
  // reuse the plane buffer once both reconstructions are done.
  constexpr int LRX = TY * 64, LRY = (TY + 2) * TX;
  constexpr int FXN = TY * (TX + 1), FYN = (TY + 1) * TX;
  constexpr int VSN = VR * VW;
  constexpr int AREA = VSN > FXN + FYN ? VSN : FXN + FYN;
  __shared__ double Vs[NC * AREA];
  __shared__ double LRx[NLR][NC * LRX];
  __shared__ double LRy[NLR][NC * LRY];
  __shared__ double red[NT / 64];
  double* Fx = Vs;                 // [NC][FXN], valid between the face and update phases
  double* Fy = Vs + NC * FXN;      // [NC][FYN]

  const BlockGeom& g = a.g;
  const Params& P = a.P;
  const int t = threadIdx.x;
  const int tx = t & 63, ty = t >> 6;
  const int bid = xcd_remap(blockIdx.x, gridDim.x);
  const int ti = bid % a.tiles_x;
  const int tj = (bid / a.tiles_x) % a.tiles_y;
  const int tk = bid / (a.tiles_x * a.tiles_y);
  const int i0 = ti * TX, j0 = tj * TY;
  const int k0 = tk * a.chunk, k1 = min(k0 + a.chunk, g.nz);
  const int c = i0 - 1 + tx, r = j0 + ty;
  const int cl = clampi(c, -H, g.nx + H - 1), rl = clampi(r, -H, g.ny + H - 1);
  const bool owner = tx >= 1 && tx <= TX && c < g.nx && r < g.ny;
  const double dt = a.dt_dev ? *a.dt_dev : a.dt;
  const RKCoef kc = rk_coef(a.stage, dt);

  double Vq[NW][NC];
  auto loadV = [&](int i, int j, int k, double v[NC]) {
    double raw[NC];
    load_raw<RC, NC>(a.qin, g.cs,
                     gidx(g, clampi(i, -H, g.nx + H - 1), clampi(j, -H, g.ny + H - 1),
                          clampi(k, -H, g.nz + H - 1)), raw);
    to_recon_vars<RC, NC>(raw, v, P.gamma);
  };
#pragma unroll
  for (int w = 0; w < NW; ++w) loadV(cl, rl, k0 - 1 - QOFF + w, Vq[w]);
  double Lc[NC], Fc[NC], Lt[NC], Rt[NC];
  RC::template vertex<NC>(&Vq[QOFF - RV], P, Lt, Rt);
#pragma unroll
  for (int m = 0; m < NC; ++m) Lc[m] = Lt[m];
#pragma unroll
  for (int w = 0; w < NW - 1; ++w)
#pragma unroll
    for (int m = 0; m < NC; ++m) Vq[w][m] = Vq[w + 1][m];
  loadV(cl, rl, k0 + RV, Vq[NW - 1]);
  RC::template vertex<NC>(&Vq[QOFF - RV], P, Lt, Rt);
  face_flux<FLUX, NC, 2>(Lc, Rt, P, Fc);
#pragma unroll
  for (int m = 0; m < NC; ++m) Lc[m] = Lt[m];

  double nrm[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) nrm[m] = 0.0;
  double smax = 0.0;

  // extra plane loads: x far halo (2*RV columns x TY rows), y halo (2*H rows x TX columns)
  constexpr int NXH = 2 * RV * TY, NYH = 2 * H * TX;

  for (int k = k0; k < k1; ++k) {
#pragma unroll
    for (int w = 0; w < NW - 1; ++w)
#pragma unroll
      for (int m = 0; m < NC; ++m) Vq[w][m] = Vq[w + 1][m];
    loadV(cl, rl, k + 1 + RV, Vq[NW - 1]);
    // plane k into LDS
#pragma unroll
    for (int m = 0; m < NC; ++m) Vs[m * VSN + (ty + H) * VW + tx + RV] = Vq[QOFF - 1][m];
    for (int e = t; e < NXH + NYH; e += NT) {
      int hi, hj, slot;
      if (e < NXH) {
        constexpr int XW = 2 * RV > 0 ? 2 * RV : 1;
        const int row = e / XW, s = e % XW;
        hi = s < RV ? i0 - 1 - RV + s : i0 + TX + 1 + (s - RV);
        hj = j0 + row;
        slot = (row + H) * VW + (s < RV ? s : 64 + RV + (s - RV));
      } else {
        const int e2 = e - NXH;
        const int col = e2 % TX, s = e2 / TX;
        hi = i0 + col;
        hj = s < H ? j0 - H + s : j0 + TY + (s - H);
        slot = (s < H ? s : TY + H + (s - H)) * VW + col + 1 + RV;
      }
      double v[NC];
      loadV(hi, hj, k, v);
#pragma unroll
      for (int m = 0; m < NC; ++m) Vs[m * VSN + slot] = v[m];
    }
    __syncthreads();
    // reconstructions: x at the owned vertex (row ty); y at rows j0-1 .. j0+TY
    // of columns i0 .. i0+TX-1
    {
      double W[2 * RV + 1][NC], L[NC], R[NC];
#pragma unroll
      for (int s = 0; s < 2 * RV + 1; ++s)
#pragma unroll
        for (int m = 0; m < NC; ++m) W[s][m] = Vs[m * VSN + (ty + H) * VW + tx + s];
      RC::template vertex<NC>(W, P, L, R);
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        LRx[0][m * LRX + ty * 64 + tx] = L[m];
        if constexpr (NLR == 2) LRx[NLR - 1][m * LRX + ty * 64 + tx] = R[m];
      }
    }
    for (int e = t; e < LRY; e += NT) {
      const int y = e / TX, col = e % TX;
      double W[2 * RV + 1][NC], L[NC], R[NC];
#pragma unroll
      for (int s = 0; s < 2 * RV + 1; ++s)
#pragma unroll
        for (int m = 0; m < NC; ++m) W[s][m] = Vs[m * VSN + (y - 1 + H - RV + s) * VW + col + 1 + RV];
      RC::template vertex<NC>(W, P, L, R);
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        LRy[0][m * LRY + e] = L[m];
        if constexpr (NLR == 2) LRy[NLR - 1][m * LRY + e] = R[m];
      }
    }
    __syncthreads();
    // x-faces f = i0 + s, s = 0..TX: F(L(f-1), R(f))
    for (int e = t; e < FXN; e += NT) {
      const int row = e / (TX + 1), s = e % (TX + 1);
      double L[NC], R[NC], F[NC];
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        L[m] = LRx[0][m * LRX + row * 64 + s];
        R[m] = LRx[NLR - 1][m * LRX + row * 64 + s + 1];
      }
      face_flux<FLUX, NC, 0>(L, R, P, F);
#pragma unroll
      for (int m = 0; m < NC; ++m) Fx[m * FXN + e] = F[m];
    }
    // y-faces fj = j0 + y, y = 0..TY: F(L(row y-1), R(row y)) = recon rows (y, y+1)
    for (int e = t; e < FYN; e += NT) {
      const int y = e / TX, col = e % TX;
      double L[NC], R[NC], F[NC];
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        L[m] = LRy[0][m * LRY + y * TX + col];
        R[m] = LRy[NLR - 1][m * LRY + (y + 1) * TX + col];
      }
      face_flux<FLUX, NC, 1>(L, R, P, F);
#pragma unroll
      for (int m = 0; m < NC; ++m) Fy[m * FYN + e] = F[m];
    }
    // z-face k + 1 of the owned column
    double Fn[NC];
    RC::template vertex<NC>(&Vq[QOFF - RV], P, Lt, Rt);
    face_flux<FLUX, NC, 2>(Lc, Rt, P, Fn);
    __syncthreads();
    if (owner && k < g.nz) {
      const int64_t ci = gidx(g, c, r, k);
      const int col = tx - 1;
      double out[NC];
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        double rr = 0.0 - (Fx[m * FXN + ty * (TX + 1) + col + 1] - Fx[m * FXN + ty * (TX + 1) + col]) / g.dx;
        rr = rr - (Fy[m * FYN + (ty + 1) * TX + col] - Fy[m * FYN + ty * TX + col]) / g.dy;
        rr = rr - (Fn[m] - Fc[m]) / g.dz;
        const double q0 = a.q0[ci + m * g.cs];
        double* accp = a.acc ? a.acc + ci + m * g.cs : nullptr;
        out[m] = rk_apply(a.stage, kc, q0, a.qin[ci + m * g.cs], accp, rr);
        nrm[m] += rr * rr;
      }
      store_with_ghosts<NC, 3>(g, a.qout, c, r, k, out);
      if (a.smax) smax = fmax(smax, wave_speed<NC>(out, g, P.gamma));
    }
#pragma unroll
    for (int m = 0; m < NC; ++m) { Fc[m] = Fn[m]; Lc[m] = Lt[m]; }
    __syncthreads();
  }

  if (a.norm_part) {
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      const double s = wg_sum<NT>(nrm[m], red);
      if (t == 0) a.norm_part[(int64_t)blockIdx.x * NC + m] = s;
    }
  }
  if (a.smax) {
    const double s = wg_max<NT>(smax, red);
    if (t == 0) atomicMax(a.smax, (unsigned long long)__double_as_longlong(s));
  }
}

}  // namespace fvm
