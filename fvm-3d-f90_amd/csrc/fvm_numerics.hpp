// fvm_numerics.hpp -- per-point FP64 numerics of the explicit residual path
// (reconstruction, face fluxes, state conversion) as device functions.
//
// Two families are compiled from this file.  exact (-ffp-contract=off): every
// expression keeps the operation order of the FVM-3D-F90 scheme routines, so
// each result is the same IEEE sequence as the reference (Fortran evaluates
// a*b*c as (a*b)*c, x**2 as x*x and -a*b as -(a*b)).  fast (FVM_FAST=1,
// -ffp-contract=fast): the same formulas with FMA contraction and
// reciprocal multiplies in place of repeated divisions; results differ from
// the reference at round-off and are held to a relative tolerance.  Face normals are the unit axis vectors
// (src/solve/class_space_discret.f90:458-461); products with the zero normal
// components are dropped, which changes at most the sign of an exact zero in
// a face flux and never the residual.
#pragma once
#include <hip/hip_runtime.h>

#include "fvm_types.hpp"

#ifndef FVM_FAST
#define FVM_FAST 0
#endif

namespace fvm {
#if FVM_FAST
namespace fast {
#else
namespace exact {
#endif
constexpr bool kFast = FVM_FAST != 0;

// Reciprocal.  exact: IEEE division.  fast: the hardware reciprocal estimate
// (relative error ~5e-8 on gfx950) refined by one Newton-Raphson step, which
// leaves a relative error below 3e-15 for the positive normal operands it
// sees here (densities, sound speeds, sums of squares); measured on 2^20
// samples spread over 2^-20..2^20.
__device__ __forceinline__ double frcp(double x) {
  if constexpr (kFast) {
    const double r = __builtin_amdgcn_rcp(x);
    const double e = fma(-x, r, 1.0);
    return fma(r, e, r);
  } else {
    return 1.0 / x;
  }
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

// ---- exact family: IEEE FP64 division and square root, shortened -----------
// The compiler expands an IEEE FP64 division a / b into two v_div_scale, a
// v_rcp estimate, four FMAs (two Newton steps on the reciprocal), a multiply,
// an FMA (the remainder), v_div_fmas and v_div_fixup.  When v_div_scale does
// not rescale -- a numerator of 0 or of magnitude at least 2^-969, a normal
// divisor with a normal reciprocal, a normal quotient, operand exponents
// less than 768 apart -- and the divisor is finite and non-zero, the two
// scale steps and the fix-up return their operand unchanged and v_div_fmas is
// a plain FMA, so the other eight instructions alone give the IEEE quotient
// bit for bit.  Every divisor on this path is positive (densities, sound
// speeds, spacings, sums of squares plus an epsilon, Harten's delta, rho_l +
// rho_b, gamma); taking the remainder negated, b q0 - a, gives a zero
// quotient the sign IEEE gives it for a positive divisor.  The reciprocal
// part depends on the divisor alone, so a divisor used several times is
// inverted once (Recip) and each quotient costs three instructions.
struct Recip { double b = 0.0, r = 0.0; };
__device__ __forceinline__ Recip erecip(double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  return Recip{b, r};
}
__device__ __forceinline__ double ediv(double a, const Recip& d) {
  const double q0 = a * d.r;
  const double rem = fma(d.b, q0, -a);
  return fma(-rem, d.r, q0);
}
__device__ __forceinline__ double ediv(double a, double b) { return ediv(a, erecip(b)); }

// IEEE FP64 square root: the compiler's sequence -- the reciprocal square
// root estimate, one Goldschmidt step and two Newton corrections -- without
// its rescaling of operands below 2^-767 and its select for zero and
// infinity, which leave every finite operand from 2^-767 up unchanged: bit
// for bit sqrt there (the operands here are products of densities and
// squared sound speeds).
__device__ __forceinline__ double esqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double h = y * 0.5;
  double g = x * y;
  const double r = fma(-h, g, 0.5);
  g = fma(g, r, g);
  double d = fma(-g, g, x);
  h = fma(h, r, h);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  return fma(d, h, g);
}

// Square root.  exact: IEEE sqrt.  fast: the hardware reciprocal square
// root estimate refined by one Goldschmidt iteration and a final Newton
// correction (relative error near one ulp for the positive normal operands it
// sees: products of densities, squared sound speeds); sqrt(0) = 0.
__device__ __forceinline__ double fsqrt(double x) {
  if constexpr (kFast) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    const double d = fma(-g, g, x);
    g = fma(d, h, g);
    return x == 0.0 ? 0.0 : g;
  } else {
    return sqrt(x);
  }
}

// sqrt(x) and 1/sqrt(x) (fast family only) from one hardware reciprocal
// square root estimate: fsqrt's Goldschmidt iteration and Newton correction
// give s, one more correction of the reciprocal half gives rs.  Measured as
// for frcp (2^20 samples over 2^-20..2^20): s equal to IEEE sqrt on every
// sample, rs within 2.2e-15 relative.
__device__ __forceinline__ void fsqrt_rsqrt(double x, double& s, double& rs) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  const double d = fma(-g, g, x);
  g = fma(d, h, g);
  r = fma(-g, h, 0.5);
  h = fma(h, r, h);
  s = g;
  rs = 2.0 * h;
}

// Reciprocals of N positive values from one frcp (fast family only): prefix
// products, one reciprocal, then back-substitution (Montgomery's trick).  The
// product of the N values must stay a normal double.  Callers pass MUSCL
// limiter denominators, at least 1e-6 each, so this holds while every slope
// stays below ~1e30 in magnitude (van Albada's squared slopes overflow near
// 1e154 in either family).  Relative error below 2.4e-15 (N = 5, measured as
// for frcp).
template <int N>
__device__ __forceinline__ void batch_rcp(const double d[N], double r[N]) {
  double p[N];
  p[0] = d[0];
#pragma unroll
  for (int i = 1; i < N; ++i) p[i] = p[i - 1] * d[i];
  double inv = frcp(p[N - 1]);
#pragma unroll
  for (int i = N - 1; i > 0; --i) {
    r[i] = inv * p[i - 1];
    inv = inv * d[i];
  }
  r[0] = inv;
}

// Fortran SIGN(1.0_dp, x)
__device__ __forceinline__ double fsign1(double x) { return copysign(1.0, x); }
// Fortran MIN(a, b) for finite reals
__device__ __forceinline__ double fmin2(double a, double b) { return (b < a) ? b : a; }

// ---- primitive <-> conserved, src/solve/class_flow_field.f90:353-396 ---------
// NC = 4: 2-D Euler (rho, rho u, rho v, rho E), w == 0; NC = 5: 3-D.
template <int NC>
__device__ __forceinline__ void cons_to_prim(const double q[NC], double p[NC], double gamma) {
  const double rho = q[0];
  double u, v;
  if constexpr (kFast) {
    const double rr = frcp(rho);
    u = q[1] * rr;
    v = q[2] * rr;
    if constexpr (NC == 5) {
      const double w = q[3] * rr;
      p[0] = rho; p[1] = u; p[2] = v; p[3] = w;
      p[4] = (gamma - 1.0) * (q[4] - 0.5 * rho * (u * u + v * v + w * w));
    } else {
      p[0] = rho; p[1] = u; p[2] = v;
      p[3] = (gamma - 1.0) * (q[3] - 0.5 * rho * (u * u + v * v));
    }
    return;
  }
  const Recip ir = erecip(rho);
  u = ediv(q[1], ir);
  v = ediv(q[2], ir);
  p[0] = rho; p[1] = u; p[2] = v;
  if constexpr (NC == 5) {
    const double w = ediv(q[3], ir);
    p[3] = w;
    p[4] = (gamma - 1.0) * (q[4] - 0.5 * rho * (u * u + v * v + w * w));
  } else {
    p[3] = (gamma - 1.0) * (q[3] - 0.5 * rho * (u * u + v * v));
  }
}

// ---- MUSCL, src/scheme/scheme_reconstruct_MUSCL.f90:114-226 ----------------
// Per-vertex form: both face values of vertex p2 share the limited slopes.
//   L = value at the vertex's right face (left state of that face, wind -1)
//   R = value at the vertex's left face  (right state of that face, wind +1)
// LIM is the limiter code (a compile-time constant: the caller switches on it
// once per vertex); codes other than minmod, van Leer and van Albada leave the
// slopes unlimited.
template <int LIM>
__device__ __forceinline__ void muscl_vertex(double p1, double p2, double p3, double kappa,
                                             double& L, double& R) {
  const double D_m = p2 - p1;
  const double D_p = p3 - p2;
  double Db_m, Db_p;
  constexpr bool kLimited = LIM == LIM_MINMOD || LIM == LIM_VANLEER || LIM == LIM_VANALBADA;
  if constexpr (LIM == LIM_MINMOD) {
    // minmod(x, y, 1) is symmetric in x and y: one evaluation serves both.
    Db_m = 0.5 * (fsign1(D_m) + fsign1(D_p)) * fmin2(fabs(D_m), fabs(1.0 * D_p));
    Db_p = Db_m;
  } else if constexpr (LIM == LIM_VANLEER) {
    const double eps = 1.0e-6;
    if constexpr (kFast) Db_m = (fsign1(D_m) + fsign1(D_p)) * D_m * D_p * frcp(fabs(D_m) + fabs(D_p) + eps);
    else Db_m = ediv((fsign1(D_m) + fsign1(D_p)) * D_m * D_p, fabs(D_m) + fabs(D_p) + eps);
    Db_p = Db_m;
  } else if constexpr (LIM == LIM_VANALBADA) {
    const double eps = 1.0e-6;
    if constexpr (kFast) Db_m = (D_m * (D_p * D_p + eps) + D_p * (D_m * D_m + eps)) * frcp(D_m * D_m + D_p * D_p + eps);
    else Db_m = ediv(D_m * (D_p * D_p + eps) + D_p * (D_m * D_m + eps), D_m * D_m + D_p * D_p + eps);
    Db_p = Db_m;
  } else {
    Db_m = D_m;
    Db_p = D_p;
  }
  if constexpr (kFast && kLimited) {
    // limited slopes are equal (Db_m == Db_p): (1-k)Db + (1+k)Db = 2 Db
    L = p2 + 0.5 * Db_m;
    R = p2 - 0.5 * Db_m;
    return;
  }
  // p2 +/- 0.25 s as one FMA: the product by 0.25 is exact (no slope is
  // below 2^-1020 in magnitude unless it is zero), so the single rounding of
  // the FMA is the rounding of the sum, as in the reference
  L = fma(0.25, (1.0 - kappa) * Db_m + (1.0 + kappa) * Db_p, p2);
  R = fma(-0.25, (1.0 - kappa) * Db_p + (1.0 + kappa) * Db_m, p2);
}

// muscl_vertex over the NC components of one vertex (V[0..2]: the vertex and
// its two neighbours).  Fast family with a rational limiter: the NC limiter
// denominators share one reciprocal.
template <int LIM, int NC>
__device__ __forceinline__ void muscl_vertex_all(const double (*V)[NC], double kappa, double L[NC],
                                                 double R[NC]) {
  if constexpr (kFast && (LIM == LIM_VANALBADA || LIM == LIM_VANLEER)) {
    const double eps = 1.0e-6;
    double num[NC], den[NC], inv[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const double D_m = V[1][c] - V[0][c];
      const double D_p = V[2][c] - V[1][c];
      if constexpr (LIM == LIM_VANALBADA) {
        // D_m (D_p^2 + eps) + D_p (D_m^2 + eps) = (D_m + D_p)(D_m D_p + eps)
        num[c] = (D_m + D_p) * (D_m * D_p + eps);
        den[c] = D_m * D_m + (D_p * D_p + eps);
      } else {
        num[c] = (fsign1(D_m) + fsign1(D_p)) * D_m * D_p;
        den[c] = fabs(D_m) + fabs(D_p) + eps;
      }
    }
    batch_rcp<NC>(den, inv);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const double Db = num[c] * inv[c];
      L[c] = V[1][c] + 0.5 * Db;
      R[c] = V[1][c] - 0.5 * Db;
    }
  } else {
#pragma unroll
    for (int c = 0; c < NC; ++c) muscl_vertex<LIM>(V[0][c], V[1][c], V[2][c], kappa, L[c], R[c]);
  }
}

// ---- WENO-JS, src/scheme/scheme_WENO.f90:87-121 ------------------------------
// The reference's variant: value at the right face of u2, with
// IS_1 = 13/12 (u1-2u2+u3)^2 + 1/4 (u1 + u3)^2 (scheme_WENO.f90:107-108), in
// its operation order.
__device__ __forceinline__ double weno_js(double u0, double u1, double u2, double u3,
                                          double u4) {
  const double c_0 = 0.1, c_1 = 0.6, c_2 = 0.3, eps = 1.0e-6;
  const double c13 = 1.0 / 3.0, c76 = 7.0 / 6.0, c116 = 11.0 / 6.0;
  const double c16 = 1.0 / 6.0, c56 = 5.0 / 6.0, c1312 = 13.0 / 12.0, c14 = 1.0 / 4.0;
  const double q0 = c13 * u0 - c76 * u1 + c116 * u2;
  const double q1 = -(c16 * u1) + c56 * u2 + c13 * u3;
  const double q2 = c13 * u2 + c56 * u3 - c16 * u4;
  // a - 2b and a - 4b as FMAs: products by 2 and 4 are exact, so each FMA
  // rounds once, as the separate subtraction does
  const double t0 = fma(-2.0, u1, u0) + u2;
  const double s0 = fma(-4.0, u1, u0) + 3.0 * u2;
  const double t1 = fma(-2.0, u2, u1) + u3;
  const double s1 = u1 + u3;
  const double t2 = fma(-2.0, u3, u2) + u4;
  const double s2 = fma(-4.0, u3, 3.0 * u2) + u4;
  const double e0 = c1312 * (t0 * t0) + c14 * (s0 * s0) + eps;
  const double e1 = c1312 * (t1 * t1) + c14 * (s1 * s1) + eps;
  const double e2 = c1312 * (t2 * t2) + c14 * (s2 * s2) + eps;
  if constexpr (kFast) {
    // alpha_k ~ c_k / e_k^2: scale all three by e0^2 e1^2 e2^2, one division
    const double E0 = e0 * e0, E1 = e1 * e1, E2 = e2 * e2;
    const double b0 = c_0 * (E1 * E2), b1 = c_1 * (E0 * E2), b2 = c_2 * (E0 * E1);
    return (b0 * q0 + b1 * q1 + b2 * q2) * frcp(b0 + b1 + b2);
  }
  const double a0 = c_0 / (e0 * e0);
  const double a1 = c_1 / (e1 * e1);
  const double a2 = c_2 / (e2 * e2);
  const Recip ras = erecip(a0 + a1 + a2);
  const double w0 = ediv(a0, ras), w1 = ediv(a1, ras), w2 = ediv(a2, ras);
  return w0 * q0 + w1 * q1 + w2 * q2;
}

// Textbook WENO5-JS (Jiang & Shu, JCP 126 (1996); IS_1 with (u1 - u3)^2) at
// both faces of u2, exact family: the value at the right face (stencil
// u0..u4) and at the left face (the mirrored stencil u4..u0).  The indicator
// terms are formed in an order symmetric under stencil reversal,
// t = (a + c) - 2 b and s = (3 c + a) - 4 b (the weight-3 point first), so the
// mirrored stencil's indicators are the forward ones reversed, bit for bit,
// and are formed once.  The weights alpha_k = c_k / e_k^2,
// w_k = alpha_k / sum alpha, are taken in product form, scaled by E0 E1 E2
// (E_k = e_k^2): w_k = b_k / (b0 + b1 + b2), b0 = c_0 E1 E2, b1 = c_1 E0 E2,
// b2 = c_2 E0 E1, one IEEE division per face; the mirrored face's b are the
// same products in reverse order (b1 is shared).  Every e_k is at least
// 1e-6, and the products stay normal doubles while the e_k stay below ~1e76.
// Products by 2, 4 and 1/4 are exact; the other sums of products are FMAs.
__device__ __forceinline__ void weno_js_pair_exact(double u0, double u1, double u2, double u3,
                                                   double u4, double& right, double& left) {
  const double c_0 = 0.1, c_1 = 0.6, c_2 = 0.3, eps = 1.0e-6;
  const double c13 = 1.0 / 3.0, c76 = 7.0 / 6.0, c116 = 11.0 / 6.0;
  const double c16 = 1.0 / 6.0, c56 = 5.0 / 6.0, c1312 = 13.0 / 12.0;
  const double t0 = fma(-2.0, u1, u0 + u2), s0 = fma(-4.0, u1, 3.0 * u2 + u0);
  const double t1 = fma(-2.0, u2, u1 + u3), s1 = u1 - u3;
  const double t2 = fma(-2.0, u3, u2 + u4), s2 = fma(-4.0, u3, 3.0 * u2 + u4);
  const double e0 = fma(c1312, t0 * t0, fma(0.25 * s0, s0, eps));
  const double e1 = fma(c1312, t1 * t1, fma(0.25 * s1, s1, eps));
  const double e2 = fma(c1312, t2 * t2, fma(0.25 * s2, s2, eps));
  const double E0 = e0 * e0, E1 = e1 * e1, E2 = e2 * e2;
  const double P12 = E1 * E2, P01 = E0 * E1;
  const double b1 = c_1 * (E0 * E2);
  {
    const double q0 = fma(c116, u2, fma(-c76, u1, c13 * u0));
    const double q1 = fma(c13, u3, fma(c56, u2, -(c16 * u1)));
    const double q2 = fma(-c16, u4, fma(c56, u3, c13 * u2));
    const double b0 = c_0 * P12, b2 = c_2 * P01;
    right = ediv(fma(b2, q2, fma(b1, q1, b0 * q0)), (b0 + b1) + b2);
  }
  {
    const double q0 = fma(c116, u2, fma(-c76, u3, c13 * u4));
    const double q1 = fma(c13, u1, fma(c56, u2, -(c16 * u3)));
    const double q2 = fma(-c16, u0, fma(c56, u1, c13 * u2));
    const double b0 = c_0 * P01, b2 = c_2 * P12;
    left = ediv(fma(b2, q2, fma(b1, q1, b0 * q0)), (b0 + b1) + b2);
  }
}

// Textbook WENO5-JS at both faces of u2 (fast family): the value at the
// right face (stencil u0..u4) and the mirrored value at the left face
// (stencil u4..u0).  The mirrored smoothness indicators are the forward ones
// in reverse order (each indicator is symmetric under stencil reversal), so
// they and the pairwise weight products are computed once.
__device__ __forceinline__ void weno_js_pair(double u0, double u1, double u2, double u3, double u4,
                                             double& right, double& left) {
  const double c_0 = 0.1, c_1 = 0.6, c_2 = 0.3, eps = 1.0e-6;
  const double c13 = 1.0 / 3.0, c76 = 7.0 / 6.0, c116 = 11.0 / 6.0;
  const double c16 = 1.0 / 6.0, c56 = 5.0 / 6.0, c1312 = 13.0 / 12.0, c14 = 1.0 / 4.0;
  const double t0 = u0 - 2.0 * u1 + u2, s0 = u0 - 4.0 * u1 + 3.0 * u2;
  const double t1 = u1 - 2.0 * u2 + u3, s1 = u1 - u3;
  const double t2 = u2 - 2.0 * u3 + u4, s2 = 3.0 * u2 - 4.0 * u3 + u4;
  const double e0 = c1312 * (t0 * t0) + c14 * (s0 * s0) + eps;
  const double e1 = c1312 * (t1 * t1) + c14 * (s1 * s1) + eps;
  const double e2 = c1312 * (t2 * t2) + c14 * (s2 * s2) + eps;
  const double E0 = e0 * e0, E1 = e1 * e1, E2 = e2 * e2;
  const double P12 = E1 * E2, P02 = E0 * E2, P01 = E0 * E1;
  double nr, dr, nl, dl;
  {
    const double q0 = c13 * u0 - c76 * u1 + c116 * u2;
    const double q1 = -(c16 * u1) + c56 * u2 + c13 * u3;
    const double q2 = c13 * u2 + c56 * u3 - c16 * u4;
    const double b0 = c_0 * P12, b1 = c_1 * P02, b2 = c_2 * P01;
    nr = b0 * q0 + b1 * q1 + b2 * q2;
    dr = b0 + b1 + b2;
  }
  {
    const double q0 = c13 * u4 - c76 * u3 + c116 * u2;
    const double q1 = -(c16 * u3) + c56 * u2 + c13 * u1;
    const double q2 = c13 * u2 + c56 * u1 - c16 * u0;
    const double b0 = c_0 * P01, b1 = c_1 * P02, b2 = c_2 * P12;
    nl = b0 * q0 + b1 * q1 + b2 * q2;
    dl = b0 + b1 + b2;
  }
  // one reciprocal for both weight sums (their product stays a normal double
  // while each smoothness indicator is below ~1e38)
  const double inv = frcp(dr * dl);
  right = nr * (dl * inv);
  left = nl * (dr * inv);
}

// ---- Harten fix, src/scheme/scheme_FDS_Roe.f90:167-177 ----------------------
__device__ __forceinline__ double entropy_correct(double lambda, double delta) {
  if constexpr (kFast) return (lambda > delta) ? lambda : 0.5 * (lambda * lambda + delta * delta) * frcp(delta);
  return (lambda > delta) ? lambda : ediv(0.5 * (lambda * lambda + delta * delta), delta);
}
// The same with delta's reciprocal formed by the caller (exact family).
__device__ __forceinline__ double entropy_correct(double lambda, const Recip& delta) {
  return (lambda > delta.b) ? lambda : ediv(0.5 * (lambda * lambda + delta.b * delta.b), delta);
}

// Face state as primitive components along the sweep direction D (0,1,2):
// un = normal velocity, ut1/ut2 = the other two components in (u,v,w) order.
// Fluxes are produced in (mass, mom_u, mom_v, mom_w, energy) order.

// ---- Roe FDS, src/scheme/scheme_FDS_Roe.f90:25-138 ---------------------------
template <int NC, int D>
__device__ __forceinline__ void flux_roe(const double L[NC], const double R[NC],
                                         const Params& P, double F[NC]) {
  const double gamma = P.gamma;
  const double rho_l = L[0], rho_r = R[0];
  const double u_l = L[1], v_l = L[2], w_l = (NC == 5) ? L[3] : 0.0;
  const double u_r = R[1], v_r = R[2], w_r = (NC == 5) ? R[3] : 0.0;
  const double p_l = L[NC - 1], p_r = R[NC - 1];
  const double Vc_l = (D == 0) ? u_l : (D == 1) ? v_l : w_l;
  const double Vc_r = (D == 0) ? u_r : (D == 1) ? v_r : w_r;
  double ke_l, ke_r;
  if constexpr (NC == 5) { ke_l = u_l * u_l + v_l * v_l + w_l * w_l; ke_r = u_r * u_r + v_r * v_r + w_r * w_r; }
  else { ke_l = u_l * u_l + v_l * v_l; ke_r = u_r * u_r + v_r * v_r; }
  double H_l, H_r, rho_b;
  if constexpr (kFast) {
    // one reciprocal square root of rho_l rho_r gives rho_b and both 1/rho
    double rs;
    fsqrt_rsqrt(rho_l * rho_r, rho_b, rs);
    const double rs2 = rs * rs;
    H_l = P.g_gm1 * p_l * (rho_r * rs2) + 0.5 * ke_l;
    H_r = P.g_gm1 * p_r * (rho_l * rs2) + 0.5 * ke_r;
  } else {
    // x + 0.5 y as one FMA (exact product, one rounding)
    H_l = fma(0.5, ke_l, ediv(P.g_gm1 * p_l, rho_l));
    H_r = fma(0.5, ke_r, ediv(P.g_gm1 * p_r, rho_r));
    rho_b = esqrt(rho_l * rho_r);
  }
  const double drho = rho_r - rho_l, du = u_r - u_l, dv = v_r - v_l, dw = w_r - w_l;
  const double dpress = p_r - p_l, dVl = Vc_r - Vc_l;
  const double den = rho_l + rho_b;
  double u_b, v_b, w_b, H_b;
  if constexpr (kFast) {
    // (rho_l a_l + rho_b a_r) / (rho_l + rho_b) = a_l + wr (a_r - a_l),
    // wr = rho_b / (rho_l + rho_b): one FMA per averaged quantity
    const double wr = rho_b * frcp(den);
    u_b = u_l + wr * du;
    v_b = v_l + wr * dv;
    w_b = (NC == 5) ? w_l + wr * dw : 0.0;
    H_b = H_l + wr * (H_r - H_l);
  } else {
    const Recip rden = erecip(den);
    u_b = ediv(rho_l * u_l + rho_b * u_r, rden);
    v_b = ediv(rho_l * v_l + rho_b * v_r, rden);
    w_b = (NC == 5) ? ediv(rho_l * w_l + rho_b * w_r, rden) : 0.0;
    H_b = ediv(rho_l * H_l + rho_b * H_r, rden);
  }
  const double Vl_b = (D == 0) ? u_b : (D == 1) ? v_b : w_b;
  double q2_b;
  if constexpr (NC == 5) q2_b = u_b * u_b + v_b * v_b + w_b * w_b;
  else q2_b = u_b * u_b + v_b * v_b;
  const double c2_b = (gamma - 1.0) * fma(-0.5, q2_b, H_b);   // H_b - 0.5 q2_b, one rounding
  double c_b, rc_b = 0.0;
  if constexpr (kFast) fsqrt_rsqrt(c2_b, c_b, rc_b);
  else c_b = esqrt(c2_b);
  const double delta = P.entfix * fabs(Vl_b + c_b);
  double lam1, lam2, lam5;
  if constexpr (kFast) {
    // one reciprocal of delta serves the three Harten corrections
    const double hrd = 0.5 * frcp(delta), dd = delta * delta;
    const double a1 = fabs(Vl_b - c_b), a2 = fabs(Vl_b), a5 = fabs(Vl_b + c_b);
    lam1 = a1 > delta ? a1 : (a1 * a1 + dd) * hrd;
    lam2 = a2 > delta ? a2 : (a2 * a2 + dd) * hrd;
    lam5 = a5 > delta ? a5 : (a5 * a5 + dd) * hrd;
  } else {
    const Recip rdelta = erecip(delta);
    lam1 = entropy_correct(fabs(Vl_b - c_b), rdelta);
    lam2 = entropy_correct(fabs(Vl_b), rdelta);
    lam5 = entropy_correct(fabs(Vl_b + c_b), rdelta);
  }
  double d1, d2, d5;
  if constexpr (kFast) {
    const double ic2 = rc_b * rc_b, hic2 = 0.5 * ic2;
    d1 = lam1 * (dpress - rho_b * c_b * dVl) * hic2;
    d2 = lam2 * (drho - dpress * ic2);
    d5 = lam5 * (dpress + rho_b * c_b * dVl) * hic2;
  } else {
    const Recip r2c2 = erecip(2.0 * c_b * c_b);
    d1 = ediv(lam1 * (dpress - rho_b * c_b * dVl), r2c2);
    d2 = lam2 * (drho - ediv(dpress, c_b * c_b));
    d5 = ediv(lam5 * (dpress + rho_b * c_b * dVl), r2c2);
  }
  const double lr = lam2 * rho_b;
  // mass
  const double dsum = d1 + d2 + d5;
  F[0] = 0.5 * ((rho_l * Vc_l + rho_r * Vc_r) - dsum);
  // momentum components; the normal one carries -/+ c and the pressure term
  const double ub[3] = {u_b, v_b, w_b};
  const double dd[3] = {du, dv, dw};
  const double vl[3] = {u_l, v_l, w_l};
  const double vr[3] = {u_r, v_r, w_r};
  if constexpr (kFast) {
    // the three waves' sums regrouped: d1 (u - c) + d2 u + d5 (u + c)
    // = (d1 + d2 + d5) u + (d5 - d1) c, likewise for the energy; the
    // convective fluxes reuse the mass fluxes rho V of each side
    const double d51 = d5 - d1;
    const double ml = rho_l * Vc_l, mr = rho_r * Vc_r;
#pragma unroll
    for (int m = 0; m < NC - 2; ++m) {
      double es, fl, fr;
      if (m == D) {
        es = dsum * ub[m] + d51 * c_b + lr * (dd[m] - dVl);
        fl = ml * vl[m] + p_l;
        fr = mr * vr[m] + p_r;
      } else {
        es = dsum * ub[m] + lr * dd[m];
        fl = ml * vl[m];
        fr = mr * vr[m];
      }
      F[1 + m] = 0.5 * ((fl + fr) - es);
    }
    double udu;
    if constexpr (NC == 5) udu = u_b * du + v_b * dv + w_b * dw;
    else udu = u_b * du + v_b * dv;
    const double es = (d1 + d5) * H_b + d51 * (c_b * Vl_b) + d2 * (0.5 * q2_b) +
                      lr * (udu - Vl_b * dVl);
    F[NC - 1] = 0.5 * ((ml * H_l + mr * H_r) - es);
    return;
  }
#pragma unroll
  for (int m = 0; m < NC - 2; ++m) {
    double e1, e2, e5, fl, fr;
    if (m == D) {
      e1 = d1 * (ub[m] - c_b);
      e2 = d2 * ub[m] + lr * (dd[m] - dVl);
      e5 = d5 * (ub[m] + c_b);
      fl = rho_l * vl[m] * Vc_l + p_l;
      fr = rho_r * vr[m] * Vc_r + p_r;
    } else {
      e1 = d1 * ub[m];
      e2 = d2 * ub[m] + lr * dd[m];
      e5 = d5 * ub[m];
      fl = rho_l * vl[m] * Vc_l;
      fr = rho_r * vr[m] * Vc_r;
    }
    F[1 + m] = 0.5 * ((fl + fr) - (e1 + e2 + e5));
  }
  // energy
  double udu;
  if constexpr (NC == 5) udu = u_b * du + v_b * dv + w_b * dw;
  else udu = u_b * du + v_b * dv;
  const double e1 = d1 * (H_b - c_b * Vl_b);
  const double e2 = d2 * 0.5 * q2_b + lr * (udu - Vl_b * dVl);
  const double e5 = d5 * (H_b + c_b * Vl_b);
  F[NC - 1] = 0.5 * ((rho_l * H_l * Vc_l + rho_r * H_r * Vc_r) - (e1 + e2 + e5));
}

// ---- Van Leer FVS, src/scheme/scheme_FVS_VanLeer.f90:25-126 ------------------
// The split keeps the subsonic momentum/energy factors in the supersonic
// branch, exactly as the reference does.
template <int NC, int D>
__device__ __forceinline__ void flux_vanleer(const double L[NC], const double R[NC],
                                             const Params& P, double F[NC]) {
  const double gamma = P.gamma;
  const double rho_l = L[0], rho_r = R[0];
  const double vl[3] = {L[1], L[2], (NC == 5) ? L[3] : 0.0};
  const double vr[3] = {R[1], R[2], (NC == 5) ? R[3] : 0.0};
  const double p_l = L[NC - 1], p_r = R[NC - 1];
  const double Vc_l = vl[D], Vc_r = vr[D];
  double q2_l, q2_r;
  if constexpr (NC == 5) { q2_l = vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2]; q2_r = vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]; }
  else { q2_l = vl[0] * vl[0] + vl[1] * vl[1]; q2_r = vr[0] * vr[0] + vr[1] * vr[1]; }
  const double c_l = kFast ? fsqrt(gamma * p_l * frcp(rho_l)) : esqrt(ediv(gamma * p_l, rho_l));
  const double c_r = kFast ? fsqrt(gamma * p_r * frcp(rho_r)) : esqrt(ediv(gamma * p_r, rho_r));
  double M_l = kFast ? Vc_l * frcp(c_l) : ediv(Vc_l, c_l);
  double M_r = kFast ? Vc_r * frcp(c_r) : ediv(Vc_r, c_r);
  M_l = (M_l >= 1.0) ? M_l : (M_l <= -1.0) ? 0.0 : 0.25 * (M_l + 1.0) * (M_l + 1.0);
  M_r = (M_r >= 1.0) ? 0.0 : (M_r <= -1.0) ? M_r : -(0.25 * (M_r - 1.0) * (M_r - 1.0));
  const double fp0 = rho_l * c_l * M_l;
  const double fm0 = rho_r * c_r * M_r;
  F[0] = fp0 + fm0;
  const double gp = fma(2.0, c_l, -Vc_l), gm = fma(-2.0, c_r, -Vc_r);   // exact products
  const double rg = kFast ? frcp(gamma) : 0.0;
  const Recip rgam = kFast ? Recip{gamma, 0.0} : erecip(gamma);
  // (0.0 / gamma is the reference's +0 for the tangential components)
#pragma unroll
  for (int m = 0; m < NC - 2; ++m) {
    double a, b;
    if (m == D && kFast) { a = fp0 * (gp * rg + vl[m]); b = fm0 * (gm * rg + vr[m]); }
    else if (m == D) { a = fp0 * (ediv(gp, rgam) + vl[m]); b = fm0 * (ediv(gm, rgam) + vr[m]); }
    else { a = fp0 * (0.0 + vl[m]); b = fm0 * (0.0 + vr[m]); }
    F[1 + m] = a + b;
  }
  const double tl = fma(2.0, c_l, (gamma - 1.0) * Vc_l);
  const double tr = fma(-2.0, c_r, (gamma - 1.0) * Vc_r);
  const double den = 2.0 * (gamma * gamma - 1.0);
  if constexpr (kFast) {
    const double rden = frcp(den);
    F[NC - 1] = fp0 * (tl * tl * rden + (q2_l - Vc_l * Vc_l) * 0.5)
              + fm0 * (tr * tr * rden + (q2_r - Vc_r * Vc_r) * 0.5);
    return;
  }
  const Recip rden = erecip(den);
  // x + y / 2 = fma(0.5, y, x): halving is exact
  F[NC - 1] = fp0 * fma(0.5, q2_l - Vc_l * Vc_l, ediv(tl * tl, rden))
            + fm0 * fma(0.5, q2_r - Vc_r * Vc_r, ediv(tr * tr, rden));
}

// ---- "AUSM" as the reference implements it, src/scheme/scheme_AUSM.f90 -------
// beta = 0, alpha = 0 (:128-129,146-147), per-side c and M (:80-83),
// Harten-type fix on |Mn| with delta = 0.1 (:160-171).
template <int NC, int D>
__device__ __forceinline__ void flux_ausm(const double L[NC], const double R[NC],
                                          const Params& P, double F[NC]) {
  const double gamma = P.gamma;
  const double rho_l = L[0], rho_r = R[0];
  const double vl[3] = {L[1], L[2], (NC == 5) ? L[3] : 0.0};
  const double vr[3] = {R[1], R[2], (NC == 5) ? R[3] : 0.0};
  const double p_l = L[NC - 1], p_r = R[NC - 1];
  const double Vc_l = vl[D], Vc_r = vr[D];
  double ke_l, ke_r;
  if constexpr (NC == 5) { ke_l = vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2]; ke_r = vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]; }
  else { ke_l = vl[0] * vl[0] + vl[1] * vl[1]; ke_r = vr[0] * vr[0] + vr[1] * vr[1]; }
  const double ir_l = kFast ? frcp(rho_l) : 0.0, ir_r = kFast ? frcp(rho_r) : 0.0;
  const Recip er_l = kFast ? Recip{rho_l, 0.0} : erecip(rho_l);
  const Recip er_r = kFast ? Recip{rho_r, 0.0} : erecip(rho_r);
  const double H_l = kFast ? P.g_gm1 * p_l * ir_l + 0.5 * ke_l : fma(0.5, ke_l, ediv(P.g_gm1 * p_l, er_l));
  const double H_r = kFast ? P.g_gm1 * p_r * ir_r + 0.5 * ke_r : fma(0.5, ke_r, ediv(P.g_gm1 * p_r, er_r));
  const double c_l = kFast ? fsqrt(gamma * p_l * ir_l) : esqrt(ediv(gamma * p_l, er_l));
  const double c_r = kFast ? fsqrt(gamma * p_r * ir_r) : esqrt(ediv(gamma * p_r, er_r));
  const double M_l = kFast ? Vc_l * frcp(c_l) : ediv(Vc_l, c_l);
  const double M_r = kFast ? Vc_r * frcp(c_r) : ediv(Vc_r, c_r);
  // Mach split, beta = 0: the beta term adds an exact +0 for |M| < 1
  const double mp = (fabs(M_l) >= 1.0) ? 0.5 * (M_l + 1.0 * fabs(M_l))
                                       : 1.0 * (0.25 * ((M_l + 1.0) * (M_l + 1.0)));
  const double mm = (fabs(M_r) >= 1.0) ? 0.5 * (M_r + -1.0 * fabs(M_r))
                                       : -1.0 * (0.25 * ((M_r + -1.0) * (M_r + -1.0)));
  const double Mn = mp + mm;
  const double aMn = fabs(Mn);
  const double delta = 0.1;
  const double Mabs = (aMn > delta) ? aMn
                    : (kFast ? 5.0 * (aMn * aMn + delta * delta)
                             : ediv(0.5 * (aMn * aMn + delta * delta), delta));
  // pressure split, alpha = 0: the alpha term adds an exact zero for |M| < 1
  const double pp = ((fabs(M_l) >= 1.0) ? 0.5 * (1.0 + 1.0 * fsign1(M_l))
                                        : 0.25 * ((M_l + 1.0) * (M_l + 1.0)) * (2.0 - 1.0 * M_l)) * p_l;
  const double pm = ((fabs(M_r) >= 1.0) ? 0.5 * (1.0 + -1.0 * fsign1(M_r))
                                        : 0.25 * ((M_r + -1.0) * (M_r + -1.0)) * (2.0 - -1.0 * M_r)) * p_r;
  const double ph = pp + pm;
  const double a = 0.5 * (Mn + Mabs), b = 0.5 * (Mn - Mabs);
  const double phl = rho_l * c_l, phr = rho_r * c_r;
  F[0] = a * phl + b * phr;
#pragma unroll
  for (int m = 0; m < NC - 2; ++m) {
    const double f = a * (phl * vl[m]) + b * (phr * vr[m]);
    F[1 + m] = (m == D) ? f + ph : f;
  }
  F[NC - 1] = a * (phl * H_l) + b * (phr * H_r);
}

// ---- AUSM+ (Liou 1996), no reference counterpart ------------------------------
// J. Comput. Phys. 129 (1996) 364-382: beta = 1/8, alpha = 3/16,
// c_1/2 = min(c~_L, c~_R), c~ = c*^2 / max(c*, |V|), c*^2 = 2(g-1)/(g+1) H.
template <int NC, int D>
__device__ __forceinline__ void flux_ausmplus(const double L[NC], const double R[NC],
                                              const Params& P, double F[NC]) {
  const double gamma = P.gamma;
  const double beta = 0.125, alpha = 0.1875;
  const double rho_l = L[0], rho_r = R[0];
  const double vl[3] = {L[1], L[2], (NC == 5) ? L[3] : 0.0};
  const double vr[3] = {R[1], R[2], (NC == 5) ? R[3] : 0.0};
  const double p_l = L[NC - 1], p_r = R[NC - 1];
  const double Vc_l = vl[D], Vc_r = vr[D];
  double ke_l, ke_r;
  if constexpr (NC == 5) { ke_l = vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2]; ke_r = vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]; }
  else { ke_l = vl[0] * vl[0] + vl[1] * vl[1]; ke_r = vr[0] * vr[0] + vr[1] * vr[1]; }
  double H_l, H_r, c12, M_l, M_r;
  if constexpr (kFast) {
    // 1/rho_l, 1/rho_r from one reciprocal; c*, 1/c*^2 from one reciprocal
    // square root per side; 1/c~ = max(c*, |V|) / c*^2, so the two
    // max(c*, |V|) share the last reciprocal
    const double rlr = frcp(rho_l * rho_r);
    H_l = P.g_gm1 * p_l * (rho_r * rlr) + 0.5 * ke_l;
    H_r = P.g_gm1 * p_r * (rho_l * rlr) + 0.5 * ke_r;
    const double k2 = 2.0 * (gamma - 1.0) * frcp(gamma + 1.0);
    const double cs2_l = k2 * H_l, cs2_r = k2 * H_r;
    double cs_l, cs_r, rcs_l, rcs_r;
    fsqrt_rsqrt(cs2_l, cs_l, rcs_l);
    fsqrt_rsqrt(cs2_r, cs_r, rcs_r);
    const double m_l = fmax(cs_l, Vc_l), m_r = fmax(cs_r, -Vc_r);
    const double rm = frcp(m_l * m_r);
    const double ct_l = cs2_l * (m_r * rm), ct_r = cs2_r * (m_l * rm);
    const bool left = !(ct_r < ct_l);
    c12 = left ? ct_l : ct_r;
    const double rc = left ? m_l * (rcs_l * rcs_l) : m_r * (rcs_r * rcs_r);
    M_l = Vc_l * rc;
    M_r = Vc_r * rc;
  } else {
    // x + 0.5 y as one FMA (exact product, one rounding)
    H_l = fma(0.5, ke_l, ediv(P.g_gm1 * p_l, rho_l));
    H_r = fma(0.5, ke_r, ediv(P.g_gm1 * p_r, rho_r));
    const double k2 = ediv(2.0 * (gamma - 1.0), gamma + 1.0);
    const double cs2_l = k2 * H_l;
    const double cs2_r = k2 * H_r;
    const double cs_l = esqrt(cs2_l), cs_r = esqrt(cs2_r);
    const double ct_l = ediv(cs2_l, fmax(cs_l, Vc_l));
    const double ct_r = ediv(cs2_r, fmax(cs_r, -Vc_r));
    c12 = fmin(ct_l, ct_r);
    const Recip rc12 = erecip(c12);
    M_l = ediv(Vc_l, rc12);
    M_r = ediv(Vc_r, rc12);
  }
  double m4p, m4m, p5p, p5m;
  if (fabs(M_l) >= 1.0) {
    m4p = 0.5 * (M_l + fabs(M_l));
    p5p = 0.5 * (1.0 + fsign1(M_l));
  } else {
    const double t = M_l * M_l - 1.0;
    m4p = 0.25 * (M_l + 1.0) * (M_l + 1.0) + beta * t * t;
    p5p = 0.25 * (M_l + 1.0) * (M_l + 1.0) * (2.0 - M_l) + alpha * M_l * t * t;
  }
  if (fabs(M_r) >= 1.0) {
    m4m = 0.5 * (M_r - fabs(M_r));
    p5m = 0.5 * (1.0 - fsign1(M_r));
  } else {
    const double t = M_r * M_r - 1.0;
    m4m = -0.25 * (M_r - 1.0) * (M_r - 1.0) - beta * t * t;
    p5m = 0.25 * (M_r - 1.0) * (M_r - 1.0) * (2.0 + M_r) - alpha * M_r * t * t;
  }
  const double m12 = m4p + m4m;
  const double mpos = 0.5 * (m12 + fabs(m12)), mneg = 0.5 * (m12 - fabs(m12));
  const double p12 = p5p * p_l + p5m * p_r;
  if constexpr (kFast) {
    // the interface speed times each side's density, once for all components
    const double al = c12 * (mpos * rho_l), ar = c12 * (mneg * rho_r);
    F[0] = al + ar;
#pragma unroll
    for (int m = 0; m < NC - 2; ++m) F[1 + m] = (m == D) ? al * vl[m] + ar * vr[m] + p12 : al * vl[m] + ar * vr[m];
    F[NC - 1] = al * H_l + ar * H_r;
    return;
  }
  F[0] = c12 * (mpos * rho_l * 1.0 + mneg * rho_r * 1.0);
#pragma unroll
  for (int m = 0; m < NC - 2; ++m) {
    const double f = c12 * (mpos * rho_l * vl[m] + mneg * rho_r * vr[m]);
    F[1 + m] = (m == D) ? f + p12 * 1.0 : f + p12 * 0.0;
  }
  F[NC - 1] = c12 * (mpos * rho_l * H_l + mneg * rho_r * H_r) + p12 * 0.0;
}

}  // namespace exact / fast
}  // namespace fvm
