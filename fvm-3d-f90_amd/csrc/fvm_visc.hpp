// fvm_visc.hpp -- central second-order viscous residual (exact or fast
// numerics, selected by FVM_FAST as for the stage kernels).
//
// The reference reserves viscous_flux_sub (src/solve/class_space_discret.f90:
// 671-675) and leaves it empty; this is the library's own definition of the
// term, the same one oracle/fvm_oracle.c restates on the CPU:
//   - per vertex: u, v, w, T = p / (rho R), mu = Sutherland(T);
//   - a stencil point outside the block maps into it per direction
//     (periodic: wrap by the nodal period n; otherwise clamp to the boundary
//     vertex), so no edge or corner ghost is read; across a face connected
//     to another block the point is that block's vertex (directions taken in
//     the order i, j, k, each in the block the previous one reached), so a
//     cut grid sees the values of the grid it was cut from; across a wall
//     face the point is the mirror vertex (vertex l for ghost -l) with the
//     normal velocity negated (slip wall, symmetry) or every velocity
//     component negated (no-slip wall), each crossing flipping its
//     components again; across a connection patch it is the donor vertex
//     the patch's index map gives, its velocity rotated into this block's
//     axes;
//   - face between a and b = a + e_d: means of the two vertices, normal
//     derivative (phi_b - phi_a) / dr_d, tangential derivatives the mean of
//     the central differences at a and b, /(4 dr_t);
//   - tau_ij = mu (d_j u_i + d_i u_j) - 2/3 mu div(u) delta_ij,
//     F_d = (tau_d1, tau_d2, tau_d3, u_i tau_di + mu cp/Pr dT/dx_d);
//   - rv = sum_d (F_d(+) - F_d(-)) / dr_d for momentum and energy.
// The stage kernel adds rv to its inviscid residual.  Each workgroup stages
// the primitive quantities of its tile (+1 halo) in LDS: a ring of three
// k-planes in 3-D, one plane in 2-D.
#pragma once
#include "fvm_numerics.hpp"
#include "fvm_types.hpp"

namespace fvm {
#if FVM_FAST
namespace fast {
#else
namespace exact {
#endif

namespace visc {

#if FVM_FAST
constexpr bool kFastV = true;
#else
constexpr bool kFastV = false;
#endif

// Reciprocals of the divisors every vertex and face shares (exact family):
// the spacings, four times the spacings and T_ref, formed once per thread.
struct Rec {
  Recip dr[3], dr4[3], tref;
};
__device__ __forceinline__ Rec make_rec(const BlockGeom& g, const ViscParams& V) {
  Rec R;
  if constexpr (!kFastV) {
    const double dr[3] = {g.dx, g.dy, g.dz};
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      R.dr[e] = erecip(dr[e]);
      R.dr4[e] = erecip(4.0 * dr[e]);
    }
    R.tref = erecip(V.t_ref);
  }
  return R;
}

// index of a stencil point along one direction: periodic wrap or clamp
__device__ __forceinline__ int vmap(int x, int n, int lo_bc, int hi_bc) {
  if (x < 0) return lo_bc == BC_PERIODIC ? ((x % n) + n) % n : 0;
  if (x >= n) return hi_bc == BC_PERIODIC ? x % n : n - 1;
  return x;
}

// A stencil point resolved to the vertex it reads: state array, element
// offset of the vertex, component stride, and the velocity's frame: the
// launch block's velocity component r is sign(r) x the read vertex's
// component axis(r) (axis(r) in bits 2r..2r+1 of `axes`, sign(r) negative
// where bit r of `neg` is set) -- crossed patches permute and reverse the
// axes, walls negate components.
struct VPoint {
  const double* q;
  int64_t c, cs;
  int axes, neg;
};
constexpr int kAxesIdentity = 0 | (1 << 2) | (2 << 4);

__device__ __forceinline__ int64_t vidx(const BlockGeom& g, int i, int j, int k) {
  return (int64_t)(k + g.oz) * g.pxy + (int64_t)(j + g.oy) * g.px + (i + g.ox);
}

__device__ __forceinline__ int ext(const BlockGeom& g, int d) { return d == 0 ? g.nx : (d == 1 ? g.ny : g.nz); }

// The vertex stencil point (x, y, z) of the launch's block reads.  Per
// direction out of range: a connected face continues in the neighbour (the
// directions are taken again there), a wall mirrors, a periodic face wraps,
// any other face clamps; a connection-patch face is crossed last, through
// the patch covering the point's (clamped) position on the face, whose
// linear index map d = dmin + S (r - rmin) carries the whole point into the
// donor.
template <int DIM>
__device__ __forceinline__ VPoint resolve(const ViscArgs& a, int x, int y, int z) {
  const BlockGeom& g = a.g;
  if (!a.tab) {
    const int mi = vmap(x, g.nx, g.bc[0], g.bc[1]), mj = vmap(y, g.ny, g.bc[2], g.bc[3]);
    const int mk = DIM == 3 ? vmap(z, g.nz, g.bc[4], g.bc[5]) : 0;
    return VPoint{a.q, vidx(g, mi, mj, mk), g.cs, kAxesIdentity, 0};
  }
  int b = a.self, neg = 0;
  int c[3] = {x, y, DIM == 3 ? z : 0};
  int axis[3] = {0, 1, 2};          // launch-block velocity component r = current block's axis[r]
  for (int hop = 0; hop <= DIM; ++hop) {
    const BlockDesc& B = a.tab[b];
    int moved = 0, patch_dir = -1;
    for (int d = 0; d < DIM && !moved; ++d) {
      const int n = ext(B.g, d);
      if (c[d] >= 0 && c[d] < n) continue;
      const int f = 2 * d + (c[d] < 0 ? 0 : 1);
      if (B.g.bc[f] == BC_CONN) {
        // shared interface vertex: vertex n-1 of the low block is vertex 0 of
        // the high one; abutting blocks: one spacing apart
        const int nb = B.nbr[f], sh = B.nbr_shared[f];
        const int nn = ext(a.tab[nb].g, d);
        c[d] = c[d] < 0 ? c[d] + nn - sh : c[d] - n + sh;
        b = nb;
        moved = 1;
      } else if (B.wall[f] != WALL_NONE) {
        c[d] = c[d] < 0 ? -c[d] : 2 * (n - 1) - c[d];
#pragma unroll
        for (int r = 0; r < 3; ++r)
          if (B.wall[f] == WALL_NOSLIP ? axis[r] < DIM : axis[r] == d) neg ^= 1 << r;
      } else if (B.g.bc[f] == BC_PATCH) {
        if (patch_dir < 0) patch_dir = d;
      } else {
        c[d] = vmap(c[d], n, B.g.bc[2 * d], B.g.bc[2 * d + 1]);
      }
    }
    if (moved) continue;
    if (patch_dir < 0) break;
    // cross the patch face of direction patch_dir
    const int d = patch_dir, side = c[d] < 0 ? 0 : 1;
    int pi = -1;
    for (int e = 0; e < a.npatch && pi < 0; ++e) {
      const PatchDesc& P = a.patches[e];
      if (P.recv != b || P.dn != d || P.side != side) continue;
      const int ca = min(max(c[P.ta], 0), ext(B.g, P.ta) - 1);
      const int cb = min(max(c[P.tb], 0), ext(B.g, P.tb) - 1);
      if (ca >= P.rmin[P.ta] && ca < P.rmin[P.ta] + P.na && cb >= P.rmin[P.tb] && cb < P.rmin[P.tb] + P.nb)
        pi = e;
    }
    if (pi < 0) break;                // (the step checks that patches tile their faces)
    const PatchDesc& P = a.patches[pi];
    int dcoord[3];
#pragma unroll
    for (int e = 0; e < 3; ++e)
      dcoord[P.map_dir[e]] = (int)P.dmin[P.map_dir[e]] + P.map_sign[e] * (c[e] - (int)P.rmin[e]);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      // the current block's component axis[r] is map_sign x the donor's map_dir
      if (P.map_sign[axis[r]] < 0) neg ^= 1 << r;
      axis[r] = P.map_dir[axis[r]];
    }
#pragma unroll
    for (int e = 0; e < 3; ++e) c[e] = dcoord[e];
    b = P.donor;
  }
  const BlockDesc& B = a.tab[b];
  return VPoint{B.q[a.in_buf], vidx(B.g, c[0], c[1], c[2]), B.g.cs,
                axis[0] | (axis[1] << 2) | (axis[2] << 4), neg};
}

// The launch block's velocity from the read vertex's (out[0..2]: u, v, w).
__device__ __forceinline__ void apply_frame(const VPoint& pt, double out[5]) {
  if (pt.axes == kAxesIdentity && pt.neg == 0) return;
  const double v3[3] = {out[0], out[1], out[2]};
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double v = v3[(pt.axes >> (2 * r)) & 3];
    out[r] = (pt.neg >> r) & 1 ? -v : v;
  }
}

// The NC conserved components of a resolved vertex.
template <int NC>
__device__ __forceinline__ void load_cons(const VPoint& pt, double q[NC]) {
  const double* __restrict__ p = pt.q;
#pragma unroll
  for (int m = 0; m < NC; ++m) q[m] = p[pt.c + m * pt.cs];
}

// u, v, w, T, mu of a conserved state q[0..NC-1]
template <int NC>
__device__ __forceinline__ void prims_of(const double q[NC], const ViscParams& V, const Rec& R,
                                         double out[5]) {
  const double rho = q[0];
  if constexpr (kFastV) {
    // one reciprocal of rho; reciprocals of R and T_ref from the host
    const double rr = frcp(rho);
    const double u = q[1] * rr, v = q[2] * rr;
    const double w = NC == 5 ? q[3] * rr : 0.0;
    const double e = q[NC - 1];
    const double p = (V.gamma - 1.0) * (e - 0.5 * rho * (u * u + v * v + w * w));
    const double t = p * (rr * V.rr_gas);
    const double tr = t * V.rt_ref;
    out[0] = u;
    out[1] = v;
    out[2] = w;
    out[3] = t;
    out[4] = V.mu_ref * tr * fsqrt(tr) * V.sfac * frcp(t + V.s);
    return;
  }
  const Recip ir = erecip(rho);
  const double u = ediv(q[1], ir), v = ediv(q[2], ir);
  const double w = NC == 5 ? ediv(q[3], ir) : 0.0;
  const double e = q[NC - 1];
  // class_flow_field.f90:383-396 operation order
  const double p = (V.gamma - 1.0) * (e - 0.5 * rho * (u * u + v * v + w * w));
  const double t = ediv(p, rho * V.r_gas);
  const double tr = ediv(t, R.tref);
  out[0] = u;
  out[1] = v;
  out[2] = w;
  out[3] = t;
  out[4] = ediv(V.mu_ref * tr * esqrt(tr) * (V.t_ref + V.s), t + V.s);
}

// u, v, w, T, mu of the conserved state at a resolved vertex
template <int NC>
__device__ __forceinline__ void prims(const VPoint& pt, const ViscParams& V, const Rec& R,
                                      double out[5]) {
  double q[NC];
  load_cons<NC>(pt, q);
  prims_of<NC>(q, V, R, out);
}

// Face flux in direction D between vertex A (offset O along D from the
// centre) and A + e_D.  at(q, dx, dy, dz) reads quantity q (u, v, w, T, mu)
// at an offset from the centre vertex.
template <int DIM, int D, int O, class At>
__device__ __forceinline__ void face(const At& at, const BlockGeom& g, const ViscParams& V,
                                     const Rec& R, double F[4]) {
  constexpr int a0 = D == 0 ? O : 0, a1 = D == 1 ? O : 0, a2 = D == 2 ? O : 0;
  constexpr int b0 = a0 + (D == 0), b1 = a1 + (D == 1), b2 = a2 + (D == 2);
  const double dr[3] = {g.dx, g.dy, g.dz};
  const double rdr[3] = {g.rdx, g.rdy, g.rdz};
  double gq[3][4];
#pragma unroll
  for (int e = 0; e < 3; ++e)
#pragma unroll
    for (int m = 0; m < 4; ++m) gq[e][m] = 0.0;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const double diff = at(m, b0, b1, b2) - at(m, a0, a1, a2);
    gq[D][m] = kFastV ? diff * rdr[D] : ediv(diff, R.dr[D]);
  }
#pragma unroll
  for (int e = 0; e < DIM; ++e) {
    if (e == D) continue;
    const int s0 = e == 0, s1 = e == 1, s2 = e == 2;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const double sum = (at(m, a0 + s0, a1 + s1, a2 + s2) - at(m, a0 - s0, a1 - s1, a2 - s2)) +
                         (at(m, b0 + s0, b1 + s1, b2 + s2) - at(m, b0 - s0, b1 - s1, b2 - s2));
      gq[e][m] = kFastV ? sum * (0.25 * rdr[e]) : ediv(sum, R.dr4[e]);
    }
  }
  const double mu = 0.5 * (at(4, a0, a1, a2) + at(4, b0, b1, b2));
  const double uf0 = 0.5 * (at(0, a0, a1, a2) + at(0, b0, b1, b2));
  const double uf1 = 0.5 * (at(1, a0, a1, a2) + at(1, b0, b1, b2));
  const double uf2 = 0.5 * (at(2, a0, a1, a2) + at(2, b0, b1, b2));
  const double div = gq[0][0] + gq[1][1] + gq[2][2];
  double tau[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double t = mu * (gq[i][D] + gq[D][i]);
    if (i == D) t = t - (2.0 / 3.0) * mu * div;
    tau[i] = t;
  }
  F[0] = tau[0];
  F[1] = tau[1];
  F[2] = tau[2];
  F[3] = uf0 * tau[0] + uf1 * tau[1] + uf2 * tau[2] + mu * V.cpr * gq[D][3];
}

template <int DIM, class At>
__device__ __forceinline__ void residual(const At& at, const BlockGeom& g, const ViscParams& V,
                                         const Rec& R, double acc[4]) {
  double Fp[4], Fm[4];
  face<DIM, 0, 0>(at, g, V, R, Fp);
  face<DIM, 0, -1>(at, g, V, R, Fm);
#pragma unroll
  for (int m = 0; m < 4; ++m) acc[m] = kFastV ? (Fp[m] - Fm[m]) * g.rdx : ediv(Fp[m] - Fm[m], R.dr[0]);
  face<DIM, 1, 0>(at, g, V, R, Fp);
  face<DIM, 1, -1>(at, g, V, R, Fm);
#pragma unroll
  for (int m = 0; m < 4; ++m)
    acc[m] = kFastV ? acc[m] + (Fp[m] - Fm[m]) * g.rdy : acc[m] + ediv(Fp[m] - Fm[m], R.dr[1]);
  if constexpr (DIM == 3) {
    face<DIM, 2, 0>(at, g, V, R, Fp);
    face<DIM, 2, -1>(at, g, V, R, Fm);
#pragma unroll
    for (int m = 0; m < 4; ++m)
      acc[m] = kFastV ? acc[m] + (Fp[m] - Fm[m]) * g.rdz : acc[m] + ediv(Fp[m] - Fm[m], R.dr[2]);
  }
}

}  // namespace visc

// 3-D: tile TX x TY columns marching over `chunk` k-planes, three-plane ring.
// The conserved values of plane k+2 are loaded while plane k is computed
// (registers: NPL entries per thread), and converted into the ring at the
// start of the next plane.  Tiles are dealt to the XCDs in contiguous ranges
// so the halos neighbouring tiles share stay in one L2.
template <int TX, int TY>
__global__ __launch_bounds__(TX * TY) void viscous3d_kernel(ViscArgs a) {
  constexpr int NT = TX * TY, W = TX + 2, HH = TY + 2, PN = W * HH;
  constexpr int NPL = (PN + NT - 1) / NT;   // ring entries per thread and plane
  __shared__ double S[3][5][PN];
  const BlockGeom& g = a.g;
  const ViscParams& V = a.V;
  const int t = threadIdx.x, tx = t % TX, ty = t / TX;
  const int tiles_xy = a.tiles_x * a.tiles_y;
  const int bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tk = bid / tiles_xy, rem = bid % tiles_xy;
  const int i0 = (rem % a.tiles_x) * TX, j0 = (rem / a.tiles_x) * TY;
  const int k0 = tk * a.chunk, k1 = min(k0 + a.chunk, g.nz);
  const int i = i0 + tx, j = j0 + ty;
  const visc::Rec R = visc::make_rec(g, V);

  // conserved values of this thread's ring entries of plane k, with the
  // velocity frame of each resolved point
  double raw[NPL][5];
  int frm[NPL][2];
  auto fetch = [&](int k) {
#pragma unroll
    for (int p = 0; p < NPL; ++p) {
      const int e = t + p * NT;
      if (e < PN) {
        const visc::VPoint pt = visc::resolve<3>(a, i0 - 1 + e % W, j0 - 1 + e / W, k);
        visc::load_cons<5>(pt, raw[p]);
        frm[p][0] = pt.axes;
        frm[p][1] = pt.neg;
      }
    }
  };
  auto commit = [&](int slot) {
#pragma unroll
    for (int p = 0; p < NPL; ++p) {
      const int e = t + p * NT;
      if (e < PN) {
        double v[5];
        visc::prims_of<5>(raw[p], V, R, v);
        visc::apply_frame(visc::VPoint{nullptr, 0, 0, frm[p][0], frm[p][1]}, v);
#pragma unroll
        for (int m = 0; m < 5; ++m) S[slot][m][e] = v[m];
      }
    }
  };
  fetch(k0 - 1);
  commit(0);
  fetch(k0);
  commit(1);
  fetch(k0 + 1);
  int sm = 0, sc = 1, sp = 2;   // ring slots of k-1, k, k+1
  double Fzm[4] = {0.0, 0.0, 0.0, 0.0};   // k- face flux of the owned column
  for (int k = k0; k < k1; ++k) {
    commit(sp);                  // plane k+1
    if (k + 1 < k1) fetch(k + 2);
    __syncthreads();
    if (i < g.nx && j < g.ny) {
      const int ctr = (ty + 1) * W + (tx + 1);
      auto at = [&](int m, int dx, int dy, int dz) -> double {
        const int slot = dz < 0 ? sm : (dz > 0 ? sp : sc);
        return S[slot][m][ctr + dy * W + dx];
      };
      // x and y faces as visc::residual; the k- face is the previous
      // plane's k+ face (same points, same arithmetic), carried in registers
      double acc[4], Fp[4], Fm[4];
      visc::face<3, 0, 0>(at, g, V, R, Fp);
      visc::face<3, 0, -1>(at, g, V, R, Fm);
#pragma unroll
      for (int m = 0; m < 4; ++m)
        acc[m] = visc::kFastV ? (Fp[m] - Fm[m]) * g.rdx : ediv(Fp[m] - Fm[m], R.dr[0]);
      visc::face<3, 1, 0>(at, g, V, R, Fp);
      visc::face<3, 1, -1>(at, g, V, R, Fm);
#pragma unroll
      for (int m = 0; m < 4; ++m)
        acc[m] = visc::kFastV ? acc[m] + (Fp[m] - Fm[m]) * g.rdy : acc[m] + ediv(Fp[m] - Fm[m], R.dr[1]);
      visc::face<3, 2, 0>(at, g, V, R, Fp);
      if (k == k0) visc::face<3, 2, -1>(at, g, V, R, Fzm);
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        acc[m] = visc::kFastV ? acc[m] + (Fp[m] - Fzm[m]) * g.rdz : acc[m] + ediv(Fp[m] - Fzm[m], R.dr[2]);
        Fzm[m] = Fp[m];
      }
      const int64_t c = (int64_t)(k + g.oz) * g.pxy + (int64_t)(j + g.oy) * g.px + (i + g.ox);
#pragma unroll
      for (int m = 0; m < 4; ++m) a.rv[(int64_t)m * g.cs + c] = acc[m];
    }
    __syncthreads();
    const int s0 = sm;
    sm = sc;
    sc = sp;
    sp = s0;
  }
}

// 2-D: tile TX x TY vertices, one plane.
template <int TX, int TY>
__global__ __launch_bounds__(TX * TY) void viscous2d_kernel(ViscArgs a) {
  constexpr int NT = TX * TY, W = TX + 2, HH = TY + 2, PN = W * HH;
  __shared__ double S[5][PN];
  const BlockGeom& g = a.g;
  const ViscParams& V = a.V;
  const int t = threadIdx.x, tx = t % TX, ty = t / TX;
  const int i0 = (blockIdx.x % a.tiles_x) * TX, j0 = (blockIdx.x / a.tiles_x) * TY;
  const int i = i0 + tx, j = j0 + ty;
  const visc::Rec R = visc::make_rec(g, V);
  for (int e = t; e < PN; e += NT) {
    double v[5];
    const visc::VPoint pt = visc::resolve<2>(a, i0 - 1 + e % W, j0 - 1 + e / W, 0);
    visc::prims<4>(pt, V, R, v);
    visc::apply_frame(pt, v);
#pragma unroll
    for (int m = 0; m < 5; ++m) S[m][e] = v[m];
  }
  __syncthreads();
  if (i < g.nx && j < g.ny) {
    const int ctr = (ty + 1) * W + (tx + 1);
    auto at = [&](int m, int dx, int dy, int) -> double { return S[m][ctr + dy * W + dx]; };
    double acc[4];
    visc::residual<2>(at, g, V, R, acc);
    const int64_t c = (int64_t)(j + g.oy) * g.px + (i + g.ox);
    // 2-D state: rho, rho u, rho v, E -> residual components rho u, rho v, E
    a.rv[c] = acc[0];
    a.rv[g.cs + c] = acc[1];
    a.rv[2 * g.cs + c] = acc[3];
  }
}

}  // namespace exact / fast
}  // namespace fvm
