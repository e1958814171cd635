/*
 * fvm_oracle.c -- CPU restatement of FVM-3D-F90's explicit residual + RK
 * stage path, against which the HIP implementation is tested.
 *
 * TEST INFRASTRUCTURE ONLY (see fvm_oracle.h).  Compiled with
 * -ffp-contract=off: every expression below keeps the reference's operation
 * order (Fortran evaluates a*b*c as (a*b)*c, x**2 as x*x, and -a*b as
 * -(a*b)), so on the same inputs the results are bit-identical to the
 * reference modules compiled with -fdefault-real-8 and no contraction.
 *
 * Layout: primitive states are AoS (rho,u,v,w,p) per point, the memory
 * image of Primitive_Variable (src/base/defs_fluid_props.f90:64-70);
 * conserved arrays are Fortran column-major q(i,j,k,eq).
 */
#include "fvm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NEQ 5
#define NG  3   /* ghost layers per side, src/base/defs_ghost_cell.f90:42 */

/* Fortran SIGN(1.0_dp, x): magnitude 1 with the sign bit of x. */
static inline double fsign1(double x) { return copysign(1.0, x); }
/* Fortran MIN for two finite reals. */
static inline double fmin2(double a, double b) { return (b < a) ? b : a; }

/* ======================================================================== */
/*  Reconstruction                                                           */
/* ======================================================================== */

/* src/scheme/scheme_reconstruct_simple.f90:42-85: first order, the left
 * state of face f is vertex f+dghost-1, the right state vertex f+dghost
 * (1-based face and ghost-mesh indices). */
void orc_recon_simple(int64_t nface, int64_t dghost, int64_t npg,
                      const double *pg, double *pr, int wind)
{
    int64_t dw = (wind == -1) ? -1 : 0;
    for (int64_t f = 1; f <= nface; ++f) {
        int64_t src = f + dghost + dw;                 /* 1-based */
        memcpy(pr + 5 * (f - 1), pg + 5 * (src - 1), 5 * sizeof(double));
    }
}

/* minmod with coefficient b, scheme_reconstruct_MUSCL.f90:177-184 */
static inline double lim_minmod(double x, double y, double b)
{
    return 0.5 * (fsign1(x) + fsign1(y)) * fmin2(fabs(x), fabs(b * y));
}

/* scheme_reconstruct_MUSCL.f90:198-205 */
static inline double lim_vanleer(double x, double y)
{
    const double eps = 1.0e-6;
    return (fsign1(x) + fsign1(y)) * x * y / (fabs(x) + fabs(y) + eps);
}

/* scheme_reconstruct_MUSCL.f90:219-226 */
static inline double lim_vanalbada(double x, double y)
{
    const double eps = 1.0e-6;
    return (x * (y * y + eps) + y * (x * x + eps)) / (x * x + y * y + eps);
}

/* One primitive component, scheme_reconstruct_MUSCL.f90:114-159.
 * wind -1: value at the right face of p2 (left state of that face);
 * wind +1: value at the left face of p2 (right state of that face). */
static double muscl_one(double p1, double p2, double p3, int wind,
                        double kappa, int limiter)
{
    double D_m = p2 - p1;
    double D_p = p3 - p2;
    double Db_m = 0.0, Db_p = 0.0;
    if (limiter == ORC_LIM_NONE || limiter == -1) {
        Db_m = D_m;
        Db_p = D_p;
    } else if (limiter == ORC_LIM_MINMOD) {
        const double b = 1.0;
        Db_m = lim_minmod(D_m, D_p, b);
        Db_p = lim_minmod(D_p, D_m, b);
    } else if (limiter == ORC_LIM_VANLEER) {
        Db_m = lim_vanleer(D_m, D_p);
        Db_p = Db_m;
    } else if (limiter == ORC_LIM_VANALBADA) {
        Db_m = lim_vanalbada(D_m, D_p);
        Db_p = Db_m;
    }
    if (wind == -1)
        return p2 + 0.25 * ((1.0 - kappa) * Db_m + (1.0 + kappa) * Db_p);
    return p2 - 0.25 * ((1.0 - kappa) * Db_p + (1.0 + kappa) * Db_m);
}

/* scheme_reconstruct_MUSCL.f90:47-104: face f uses the ghost-mesh stencil
 * ig-1..ig+1 (left) or ig..ig+2 (right), ig = f + dghost - 1. */
void orc_recon_muscl(int64_t nface, int64_t dghost, double kappa, int limiter,
                     int64_t npg, const double *pg, double *pr, int wind)
{
    int64_t s0 = (wind == -1) ? -1 : 0;
    for (int64_t f = 1; f <= nface; ++f) {
        int64_t ig = f + dghost - 1;
        const double *a = pg + 5 * (ig + s0 - 1);
        const double *b = pg + 5 * (ig + s0);
        const double *c = pg + 5 * (ig + s0 + 1);
        for (int m = 0; m < 5; ++m)
            pr[5 * (f - 1) + m] = muscl_one(a[m], b[m], c[m], wind, kappa, limiter);
    }
}

/* Jiang-Shu weights on a 5-point stencil, value at the right face of u[2]:
 * the reference's variant, scheme_WENO.f90:87-121, with its IS_1 term
 * (u2 + u4)**2 (scheme_WENO.f90:107-108) and its operation order. */
static double weno_js_ref(const double u[5])
{
    const double c_0 = 0.1, c_1 = 0.6, c_2 = 0.3, eps = 1.0e-6;
    const double c13 = 1.0 / 3.0, c76 = 7.0 / 6.0, c116 = 11.0 / 6.0;
    const double c16 = 1.0 / 6.0, c56 = 5.0 / 6.0, c1312 = 13.0 / 12.0;
    const double c14 = 1.0 / 4.0;
    double q0 = c13 * u[0] - c76 * u[1] + c116 * u[2];
    double q1 = -(c16 * u[1]) + c56 * u[2] + c13 * u[3];
    double q2 = c13 * u[2] + c56 * u[3] - c16 * u[4];
    double t0 = u[0] - 2.0 * u[1] + u[2];
    double s0 = u[0] - 4.0 * u[1] + 3.0 * u[2];
    double t1 = u[1] - 2.0 * u[2] + u[3];
    double s1 = u[1] + u[3];
    double t2 = u[2] - 2.0 * u[3] + u[4];
    double s2 = 3.0 * u[2] - 4.0 * u[3] + u[4];
    double IS_0 = c1312 * (t0 * t0) + c14 * (s0 * s0);
    double IS_1 = c1312 * (t1 * t1) + c14 * (s1 * s1);
    double IS_2 = c1312 * (t2 * t2) + c14 * (s2 * s2);
    double e0 = IS_0 + eps, e1 = IS_1 + eps, e2 = IS_2 + eps;
    double alpha_0 = c_0 / (e0 * e0);
    double alpha_1 = c_1 / (e1 * e1);
    double alpha_2 = c_2 / (e2 * e2);
    double alpha_sum = alpha_0 + alpha_1 + alpha_2;
    double omega_0 = alpha_0 / alpha_sum;
    double omega_1 = alpha_1 / alpha_sum;
    double omega_2 = alpha_2 / alpha_sum;
    return omega_0 * q0 + omega_1 * q1 + omega_2 * q2;
}

/* Textbook WENO5-JS (Jiang & Shu, JCP 126 (1996)), value at the right face
 * of u[2]; this library's variant beside the reference's, with
 * IS_1 = 13/12 (u1 - 2 u2 + u3)^2 + 1/4 (u1 - u3)^2.  The indicator terms are
 * formed in an order symmetric under stencil reversal, t = (a + c) - 2 b and
 * s = (3 c + a) - 4 b (the weight-3 point first), so the mirrored stencil's
 * indicators are the forward ones reversed, bit for bit (the GPU forms them
 * once for a vertex's two faces).  The weights alpha_k = c_k / e_k^2,
 * w_k = alpha_k / sum alpha, are taken in product form, scaled by
 * E_0 E_1 E_2 (E_k = e_k^2): w_k = b_k / (b_0 + b_1 + b_2) with
 * b_0 = c_0 E_1 E_2, b_1 = c_1 E_0 E_2, b_2 = c_2 E_0 E_1, one division per
 * face.  Products by 2, 4 and 1/4 are exact; the other sums of products are
 * single-rounding fma's. */
static double weno_js_textbook(const double u[5])
{
    const double c_0 = 0.1, c_1 = 0.6, c_2 = 0.3, eps = 1.0e-6;
    const double c13 = 1.0 / 3.0, c76 = 7.0 / 6.0, c116 = 11.0 / 6.0;
    const double c16 = 1.0 / 6.0, c56 = 5.0 / 6.0, c1312 = 13.0 / 12.0;
    double q0 = fma(c116, u[2], fma(-c76, u[1], c13 * u[0]));
    double q1 = fma(c13, u[3], fma(c56, u[2], -(c16 * u[1])));
    double q2 = fma(-c16, u[4], fma(c56, u[3], c13 * u[2]));
    double t0 = (u[0] + u[2]) - 2.0 * u[1];
    double s0 = (3.0 * u[2] + u[0]) - 4.0 * u[1];
    double t1 = (u[1] + u[3]) - 2.0 * u[2];
    double s1 = u[1] - u[3];
    double t2 = (u[2] + u[4]) - 2.0 * u[3];
    double s2 = (3.0 * u[2] + u[4]) - 4.0 * u[3];
    double e0 = fma(c1312, t0 * t0, fma(0.25 * s0, s0, eps));
    double e1 = fma(c1312, t1 * t1, fma(0.25 * s1, s1, eps));
    double e2 = fma(c1312, t2 * t2, fma(0.25 * s2, s2, eps));
    double E0 = e0 * e0, E1 = e1 * e1, E2 = e2 * e2;
    double b0 = c_0 * (E1 * E2), b1 = c_1 * (E0 * E2), b2 = c_2 * (E0 * E1);
    return fma(b2, q2, fma(b1, q1, b0 * q0)) / ((b0 + b1) + b2);
}

static double weno_js(const double u[5], int textbook)
{
    return textbook ? weno_js_textbook(u) : weno_js_ref(u);
}

/* scheme_WENO.f90:28-72 on conserved variables qg(npg, neq) column-major.
 * Reference mode: left state on ig-2..ig+2, right state on ig-1..ig+3 with
 * the same left-biased formula (scheme_WENO.f90:54-56,69).  Textbook mode:
 * the right state is the mirror image on ig-1..ig+3 reversed. */
void orc_recon_weno(int64_t nface, int neq, int64_t dghost, int64_t npg,
                    const double *qg, double *qr, int wind, int textbook)
{
    int64_t lo = (wind == -1) ? -2 : -1;
    for (int64_t f = 1; f <= nface; ++f) {
        int64_t ig = f + dghost - 1;
        for (int e = 0; e < neq; ++e) {
            double u[5];
            for (int s = 0; s < 5; ++s)
                u[s] = qg[(int64_t)e * npg + (ig + lo + s - 1)];
            if (textbook && wind == 1) {
                double r[5] = {u[4], u[3], u[2], u[1], u[0]};
                qr[(int64_t)e * nface + (f - 1)] = weno_js(r, 1);
            } else {
                qr[(int64_t)e * nface + (f - 1)] = weno_js(u, textbook);
            }
        }
    }
}

/* ======================================================================== */
/*  Face fluxes                                                              */
/* ======================================================================== */

/* Harten fix, scheme_FDS_Roe.f90:167-177 */
static inline double entropy_correct(double lambda, double delta)
{
    if (lambda > delta) return lambda;
    return 0.5 * (lambda * lambda + delta * delta) / delta;
}

/* scheme_FDS_Roe.f90:25-138; ql/qr AoS primitives, norm(nface,3),
 * flux(nface,neq) column-major. */
void orc_flux_roe(int64_t nface, int neq, double gamma, const double *norm,
                  const double *ql, const double *qr, double *flux, double entfix)
{
    for (int64_t f = 0; f < nface; ++f) {
        const double n1 = norm[f], n2 = norm[nface + f], n3 = norm[2 * nface + f];
        const double *L = ql + 5 * f, *R = qr + 5 * f;
        double rho_l = L[0], u_l = L[1], v_l = L[2], w_l = L[3], p_l = L[4];
        double rho_r = R[0], u_r = R[1], v_r = R[2], w_r = R[3], p_r = R[4];
        double Vc_l = u_l * n1 + v_l * n2 + w_l * n3;
        double Vc_r = u_r * n1 + v_r * n2 + w_r * n3;
        double H_l = gamma / (gamma - 1.0) * p_l / rho_l + 0.5 * (u_l * u_l + v_l * v_l + w_l * w_l);
        double H_r = gamma / (gamma - 1.0) * p_r / rho_r + 0.5 * (u_r * u_r + v_r * v_r + w_r * w_r);
        double drho = rho_r - rho_l, du = u_r - u_l, dv = v_r - v_l, dw = w_r - w_l;
        double dpress = p_r - p_l, dVl = Vc_r - Vc_l;
        /* Roe averages, scheme_FDS_Roe.f90:151-159 */
        double rho_b = sqrt(rho_l * rho_r);
        double u_b = (rho_l * u_l + rho_b * u_r) / (rho_l + rho_b);
        double v_b = (rho_l * v_l + rho_b * v_r) / (rho_l + rho_b);
        double w_b = (rho_l * w_l + rho_b * w_r) / (rho_l + rho_b);
        double H_b = (rho_l * H_l + rho_b * H_r) / (rho_l + rho_b);
        double Vl_b = u_b * n1 + v_b * n2 + w_b * n3;
        double q2_b = u_b * u_b + v_b * v_b + w_b * w_b;
        double c_b = sqrt((gamma - 1.0) * (H_b - 0.5 * q2_b));
        double delta = entfix * fabs(Vl_b + c_b);
        double lam1 = entropy_correct(fabs(Vl_b - c_b), delta);
        double lam2 = entropy_correct(fabs(Vl_b), delta);
        double lam5 = entropy_correct(fabs(Vl_b + c_b), delta);
        double d1[5], d2[5], d5[5], fl[5], fr[5];
        d1[0] = lam1 * (dpress - rho_b * c_b * dVl) / (2.0 * c_b * c_b);
        d1[1] = d1[0] * (u_b - c_b * n1);
        d1[2] = d1[0] * (v_b - c_b * n2);
        d1[3] = d1[0] * (w_b - c_b * n3);
        d1[4] = d1[0] * (H_b - c_b * Vl_b);
        d2[0] = lam2 * (drho - dpress / (c_b * c_b));
        d2[1] = d2[0] * u_b + lam2 * rho_b * (du - dVl * n1);
        d2[2] = d2[0] * v_b + lam2 * rho_b * (dv - dVl * n2);
        d2[3] = d2[0] * w_b + lam2 * rho_b * (dw - dVl * n3);
        d2[4] = d2[0] * 0.5 * q2_b + lam2 * rho_b * (u_b * du + v_b * dv + w_b * dw - Vl_b * dVl);
        d5[0] = lam5 * (dpress + rho_b * c_b * dVl) / (2.0 * c_b * c_b);
        d5[1] = d5[0] * (u_b + c_b * n1);
        d5[2] = d5[0] * (v_b + c_b * n2);
        d5[3] = d5[0] * (w_b + c_b * n3);
        d5[4] = d5[0] * (H_b + c_b * Vl_b);
        fl[0] = rho_l * Vc_l;
        fl[1] = rho_l * u_l * Vc_l + p_l * n1;
        fl[2] = rho_l * v_l * Vc_l + p_l * n2;
        fl[3] = rho_l * w_l * Vc_l + p_l * n3;
        fl[4] = rho_l * H_l * Vc_l;
        fr[0] = rho_r * Vc_r;
        fr[1] = rho_r * u_r * Vc_r + p_r * n1;
        fr[2] = rho_r * v_r * Vc_r + p_r * n2;
        fr[3] = rho_r * w_r * Vc_r + p_r * n3;
        fr[4] = rho_r * H_r * Vc_r;
        for (int e = 0; e < neq && e < 5; ++e)
            flux[(int64_t)e * nface + f] = 0.5 * ((fl[e] + fr[e]) - (d1[e] + d2[e] + d5[e]));
    }
}

/* scheme_FVS_VanLeer.f90:101-126; the reference passes the side's own Mach
 * number as the branch selector (:82-83). */
static inline double vl_split(double Ma, int wind)
{
    if (wind == -1) {
        if (Ma >= 1.0) return Ma;
        if (Ma <= -1.0) return 0.0;
        return 0.25 * (Ma + 1.0) * (Ma + 1.0);
    }
    if (Ma >= 1.0) return 0.0;
    if (Ma <= -1.0) return Ma;
    return -(0.25 * (Ma - 1.0) * (Ma - 1.0));
}

/* scheme_FVS_VanLeer.f90:25-98 */
void orc_flux_vanleer(int64_t nface, int neq, double gamma, const double *norm,
                      const double *ql, const double *qr, double *flux)
{
    for (int64_t f = 0; f < nface; ++f) {
        const double n[3] = {norm[f], norm[nface + f], norm[2 * nface + f]};
        const double *L = ql + 5 * f, *R = qr + 5 * f;
        double rho_l = L[0], u_l = L[1], v_l = L[2], w_l = L[3], p_l = L[4];
        double rho_r = R[0], u_r = R[1], v_r = R[2], w_r = R[3], p_r = R[4];
        double Vc_l = u_l * n[0] + v_l * n[1] + w_l * n[2];
        double Vc_r = u_r * n[0] + v_r * n[1] + w_r * n[2];
        double q2_l = u_l * u_l + v_l * v_l + w_l * w_l;
        double q2_r = u_r * u_r + v_r * v_r + w_r * w_r;
        double c_l = sqrt(gamma * p_l / rho_l);
        double c_r = sqrt(gamma * p_r / rho_r);
        double M_l = vl_split(Vc_l / c_l, -1);
        double M_r = vl_split(Vc_r / c_r, 1);
        double fp[5], fm[5];
        fp[0] = rho_l * c_l * M_l;
        fp[1] = fp[0] * (n[0] * (-Vc_l + 2.0 * c_l) / gamma + u_l);
        fp[2] = fp[0] * (n[1] * (-Vc_l + 2.0 * c_l) / gamma + v_l);
        fp[3] = fp[0] * (n[2] * (-Vc_l + 2.0 * c_l) / gamma + w_l);
        double tl = (gamma - 1.0) * Vc_l + 2.0 * c_l;
        fp[4] = fp[0] * (tl * tl / (2.0 * (gamma * gamma - 1.0)) + (q2_l - Vc_l * Vc_l) / 2.0);
        fm[0] = rho_r * c_r * M_r;
        fm[1] = fm[0] * (n[0] * (-Vc_r - 2.0 * c_r) / gamma + u_r);
        fm[2] = fm[0] * (n[1] * (-Vc_r - 2.0 * c_r) / gamma + v_r);
        fm[3] = fm[0] * (n[2] * (-Vc_r - 2.0 * c_r) / gamma + w_r);
        double tr = (gamma - 1.0) * Vc_r - 2.0 * c_r;
        fm[4] = fm[0] * (tr * tr / (2.0 * (gamma * gamma - 1.0)) + (q2_r - Vc_r * Vc_r) / 2.0);
        for (int e = 0; e < neq && e < 5; ++e)
            flux[(int64_t)e * nface + f] = fp[e] + fm[e];
    }
}

/* scheme_AUSM.f90:123-140 with beta = 0 */
static inline double ausm_msplit(double Ma, double wind)
{
    const double beta = 0.0;
    if (fabs(Ma) >= 1.0) return 0.5 * (Ma + wind * fabs(Ma));
    double t = Ma * Ma - 1.0;
    return wind * (0.25 * ((Ma + wind) * (Ma + wind)) + beta * (t * t));
}

/* scheme_AUSM.f90:142-160 with alpha = 0 */
static inline double ausm_psplit(double p, double Ma, double wind)
{
    const double alpha = 0.0;
    double p_s;
    if (fabs(Ma) >= 1.0) {
        p_s = 0.5 * (1.0 + wind * fsign1(Ma));
    } else {
        double t = Ma * Ma - 1.0;
        p_s = 0.25 * ((Ma + wind) * (Ma + wind)) * (2.0 - wind * Ma) + wind * alpha * Ma * (t * t);
    }
    return p_s * p;
}

/* scheme_AUSM.f90:162-171, delta = 0.1 */
static inline double ausm_dissip(double Mn_abs)
{
    const double delta = 0.1;
    if (Mn_abs > delta) return Mn_abs;
    return 0.5 * (Mn_abs * Mn_abs + delta * delta) / delta;
}

/* scheme_AUSM.f90:27-112 */
void orc_flux_ausm(int64_t nface, int neq, double gamma, const double *norm,
                   const double *ql, const double *qr, double *flux)
{
    for (int64_t f = 0; f < nface; ++f) {
        const double n[3] = {norm[f], norm[nface + f], norm[2 * nface + f]};
        const double *L = ql + 5 * f, *R = qr + 5 * f;
        double rho_l = L[0], u_l = L[1], v_l = L[2], w_l = L[3], p_l = L[4];
        double rho_r = R[0], u_r = R[1], v_r = R[2], w_r = R[3], p_r = R[4];
        double Vc_l = u_l * n[0] + v_l * n[1] + w_l * n[2];
        double Vc_r = u_r * n[0] + v_r * n[1] + w_r * n[2];
        double H_l = gamma / (gamma - 1.0) * p_l / rho_l + 0.5 * (u_l * u_l + v_l * v_l + w_l * w_l);
        double H_r = gamma / (gamma - 1.0) * p_r / rho_r + 0.5 * (u_r * u_r + v_r * v_r + w_r * w_r);
        double c_l = sqrt(gamma * p_l / rho_l);
        double c_r = sqrt(gamma * p_r / rho_r);
        double M_l = Vc_l / c_l, M_r = Vc_r / c_r;
        double Mn = ausm_msplit(M_l, 1.0) + ausm_msplit(M_r, -1.0);
        double Mn_abs = ausm_dissip(fabs(Mn));
        double phil[5], phir[5];
        phil[0] = rho_l * c_l; phil[1] = phil[0] * u_l; phil[2] = phil[0] * v_l;
        phil[3] = phil[0] * w_l; phil[4] = phil[0] * H_l;
        phir[0] = rho_r * c_r; phir[1] = phir[0] * u_r; phir[2] = phir[0] * v_r;
        phir[3] = phir[0] * w_r; phir[4] = phir[0] * H_r;
        double ph = ausm_psplit(p_l, M_l, 1.0) + ausm_psplit(p_r, M_r, -1.0);
        const double nn[5] = {0.0, n[0], n[1], n[2], 0.0};
        for (int e = 0; e < neq && e < 5; ++e)
            flux[(int64_t)e * nface + f] = 0.5 * (Mn + Mn_abs) * phil[e]
                                         + 0.5 * (Mn - Mn_abs) * phir[e] + ph * nn[e];
    }
}

/* Liou's AUSM+ (J. Comput. Phys. 129 (1996) 364-382), eqs. (19)-(21),
 * (26), (32), (40); beta = 1/8, alpha = 3/16, interface speed of sound
 * c_1/2 = min(c~_L, c~_R), c~ = c*^2 / max(c*, |V|), c*^2 = 2(g-1)/(g+1) H.
 * No reference counterpart. */
void orc_flux_ausmplus(int64_t nface, int neq, double gamma, const double *norm,
                       const double *ql, const double *qr, double *flux)
{
    const double beta = 0.125, alpha = 0.1875;
    for (int64_t f = 0; f < nface; ++f) {
        const double n[3] = {norm[f], norm[nface + f], norm[2 * nface + f]};
        const double *L = ql + 5 * f, *R = qr + 5 * f;
        double rho_l = L[0], u_l = L[1], v_l = L[2], w_l = L[3], p_l = L[4];
        double rho_r = R[0], u_r = R[1], v_r = R[2], w_r = R[3], p_r = R[4];
        double Vc_l = u_l * n[0] + v_l * n[1] + w_l * n[2];
        double Vc_r = u_r * n[0] + v_r * n[1] + w_r * n[2];
        double H_l = gamma / (gamma - 1.0) * p_l / rho_l + 0.5 * (u_l * u_l + v_l * v_l + w_l * w_l);
        double H_r = gamma / (gamma - 1.0) * p_r / rho_r + 0.5 * (u_r * u_r + v_r * v_r + w_r * w_r);
        double cs2_l = 2.0 * (gamma - 1.0) / (gamma + 1.0) * H_l;
        double cs2_r = 2.0 * (gamma - 1.0) / (gamma + 1.0) * H_r;
        double cs_l = sqrt(cs2_l), cs_r = sqrt(cs2_r);
        double ct_l = cs2_l / fmax(cs_l, Vc_l);
        double ct_r = cs2_r / fmax(cs_r, -Vc_r);
        double c12 = fmin(ct_l, ct_r);
        double M_l = Vc_l / c12, M_r = Vc_r / c12;
        double m4p, m4m, p5p, p5m;
        if (fabs(M_l) >= 1.0) {
            m4p = 0.5 * (M_l + fabs(M_l));
            p5p = 0.5 * (1.0 + fsign1(M_l));
        } else {
            double t = M_l * M_l - 1.0;
            m4p = 0.25 * (M_l + 1.0) * (M_l + 1.0) + beta * t * t;
            p5p = 0.25 * (M_l + 1.0) * (M_l + 1.0) * (2.0 - M_l) + alpha * M_l * t * t;
        }
        if (fabs(M_r) >= 1.0) {
            m4m = 0.5 * (M_r - fabs(M_r));
            p5m = 0.5 * (1.0 - fsign1(M_r));
        } else {
            double t = M_r * M_r - 1.0;
            m4m = -0.25 * (M_r - 1.0) * (M_r - 1.0) - beta * t * t;
            p5m = 0.25 * (M_r - 1.0) * (M_r - 1.0) * (2.0 + M_r) - alpha * M_r * t * t;
        }
        double m12 = m4p + m4m;
        double mp = 0.5 * (m12 + fabs(m12)), mm = 0.5 * (m12 - fabs(m12));
        double p12 = p5p * p_l + p5m * p_r;
        double psl[5] = {1.0, u_l, v_l, w_l, H_l};
        double psr[5] = {1.0, u_r, v_r, w_r, H_r};
        const double nn[5] = {0.0, n[0], n[1], n[2], 0.0};
        for (int e = 0; e < neq && e < 5; ++e)
            flux[(int64_t)e * nface + f] = c12 * (mp * rho_l * psl[e] + mm * rho_r * psr[e]) + p12 * nn[e];
    }
}

/* ======================================================================== */
/*  Runge-Kutta combine                                                      */
/* ======================================================================== */

/* scheme_Runge_Kutta.f90:30-63 (coefficients as doubles: -fdefault-real-8) */
void orc_runge_kutta(int m, double dt, int64_t n, double *u0, double *um,
                     const double *rm)
{
    const double a2[3] = {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0};
    const double a3[3] = {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0};
    for (int64_t i = 0; i < n; ++i) {
        switch (m) {
        case 0: u0[i] = um[i]; break;
        case 1: um[i] = u0[i] + dt * rm[i]; break;
        case 2: um[i] = a2[0] * u0[i] + a2[1] * um[i] + a2[2] * dt * rm[i]; break;
        case 3: um[i] = a3[0] * u0[i] + a3[1] * um[i] + a3[2] * dt * rm[i]; break;
        default: break;
        }
    }
}

/* Classical RK4 with an accumulator; no reference counterpart (the
 * reference fixes total_RK_stage = 3, src/solve/class_time_discret.f90:29). */
void orc_rk4_stage(int s, double dt, int64_t n, const double *u0, double *um,
                   double *acc, const double *rm)
{
    const double dt2 = dt / 2.0, dt3 = dt / 3.0, dt6 = dt / 6.0;
    for (int64_t x = 0; x < n; ++x) {
        double r = rm[x];
        switch (s) {
        case 1: acc[x] = u0[x] + dt6 * r; um[x] = u0[x] + dt2 * r; break;
        case 2: acc[x] = acc[x] + dt3 * r; um[x] = u0[x] + dt2 * r; break;
        case 3: acc[x] = acc[x] + dt3 * r; um[x] = u0[x] + dt * r; break;
        default: um[x] = acc[x] + dt6 * r; break;
        }
    }
}

/* ======================================================================== */
/*  Whole-block solver                                                       */
/* ======================================================================== */

typedef struct {
    int64_t n[3];    /* real vertices */
    int64_t gz;      /* ghost layers along k (0 for 2-D) */
    int64_t ng[3];   /* ghost-inclusive extents */
    double *q;       /* (ng0, ng1, ng2, 5) */
    double *p;       /* AoS (5, ng0, ng1, ng2) */
} blk_t;

static inline int64_t gidx(const blk_t *b, int64_t i, int64_t j, int64_t k)
{
    return i + b->ng[0] * (j + b->ng[1] * k);
}

static inline int64_t goff(const blk_t *b, int d)
{
    return d == 2 ? b->gz : NG;
}

/* class_flow_field.f90:353-377 */
static void p_to_q_point(double gamma, const double *p, double *q, int64_t stride)
{
    q[0] = p[0];
    q[stride] = p[0] * p[1];
    q[2 * stride] = p[0] * p[2];
    q[3 * stride] = p[0] * p[3];
    q[4 * stride] = p[4] / (gamma - 1.0) + 0.5 * p[0] * (p[1] * p[1] + p[2] * p[2] + p[3] * p[3]);
}

/* class_flow_field.f90:383-396 */
static void q_to_p_point(double gamma, const double *q, int64_t stride, double *p)
{
    double rho = q[0];
    double u = q[stride] / rho, v = q[2 * stride] / rho, w = q[3 * stride] / rho;
    p[0] = rho;
    p[1] = u;
    p[2] = v;
    p[3] = w;
    p[4] = (gamma - 1.0) * (q[4 * stride] - 0.5 * rho * (u * u + v * v + w * w));
}

/* Wall kind of a face code: 0 none, 1 slip (normal velocity negated),
 * 2 no-slip (every velocity component negated) */
static int wall_of(int kind)
{
    if (kind == ORC_BC_WALL || kind == ORC_BC_WALL_INVISCID || kind == ORC_BC_SYMMETRY) return 1;
    if (kind == ORC_BC_NOSLIP) return 2;
    return 0;
}

/* Ghost fill of p and q on the six face regions (edges/corners untouched).
 * Extrapolation copies the boundary vertex (defs_block_data.f90:226-256);
 * periodic copies vertex (r mod n) for real index r outside [0, n); a wall
 * copies the mirror vertex (ghost -l <- vertex l) and negates the normal
 * velocity (slip) or the first dim velocity components (no-slip). */
static void fill_ghosts(blk_t *b, const int bc[6], int dim)
{
    int64_t cs = b->ng[0] * b->ng[1] * b->ng[2];
    for (int d = 0; d < dim; ++d) {
        int64_t g = goff(b, d);
        for (int side = 0; side < 2; ++side) {
            int kind = bc[2 * d + side];
            int64_t lo[3], hi[3];
            for (int e = 0; e < 3; ++e) {
                lo[e] = goff(b, e);
                hi[e] = goff(b, e) + b->n[e] - 1;
            }
            for (int64_t l = 1; l <= g; ++l) {
                int64_t gpos = side == 0 ? g - l : g + b->n[d] - 1 + l;
                int64_t spos;
                int wall = wall_of(kind);
                if (kind == ORC_BC_PERIODIC)
                    spos = side == 0 ? g + b->n[d] - l : g + l - 1;
                else if (wall)
                    spos = side == 0 ? g + l : g + b->n[d] - 1 - l;
                else
                    spos = side == 0 ? g : g + b->n[d] - 1;
                int64_t l2[3] = {lo[0], lo[1], lo[2]}, h2[3] = {hi[0], hi[1], hi[2]};
                l2[d] = h2[d] = gpos;
                for (int64_t k = l2[2]; k <= h2[2]; ++k)
                    for (int64_t j = l2[1]; j <= h2[1]; ++j)
                        for (int64_t i = l2[0]; i <= h2[0]; ++i) {
                            int64_t s[3] = {i, j, k};
                            s[d] = spos;
                            int64_t dst = gidx(b, i, j, k), src = gidx(b, s[0], s[1], s[2]);
                            memcpy(b->p + 5 * dst, b->p + 5 * src, 5 * sizeof(double));
                            for (int e = 0; e < NEQ; ++e) b->q[e * cs + dst] = b->q[e * cs + src];
                            for (int e = 0; e < dim; ++e)
                                if (wall == 2 || (wall == 1 && e == d)) {
                                    b->p[5 * dst + 1 + e] = -b->p[5 * dst + 1 + e];
                                    b->q[(1 + e) * cs + dst] = -b->q[(1 + e) * cs + dst];
                                }
                        }
            }
        }
    }
}

static void update_primitives(blk_t *b, double gamma)
{
    int64_t cs = b->ng[0] * b->ng[1] * b->ng[2];
    for (int64_t k = 0; k < b->n[2]; ++k)
        for (int64_t j = 0; j < b->n[1]; ++j)
            for (int64_t i = 0; i < b->n[0]; ++i) {
                int64_t c = gidx(b, i + NG, j + NG, k + b->gz);
                q_to_p_point(gamma, b->q + c, cs, b->p + 5 * c);
            }
}

/* q -> p for a line of face states qr(nface, 5) (class_flow_field.f90:415-428) */
static void q_to_p_line(double gamma, int64_t nface, const double *qr, double *p)
{
    for (int64_t f = 0; f < nface; ++f) q_to_p_point(gamma, qr + f, nface, p + 5 * f);
}

/* Inviscid residual, class_space_discret.f90:120-176 and 678-732: for each
 * swept direction and each pencil, reconstruct face states, evaluate the face
 * flux, then rhs -= (F(f+1) - F(f)) / dr. */
static void inviscid_residual(blk_t *b, const orc_config *c, double *rhs)
{
    int64_t cs = b->ng[0] * b->ng[1] * b->ng[2];
    int64_t nreal = b->n[0] * b->n[1] * b->n[2];
    for (int d = 0; d < c->dim; ++d) {
        int64_t nv = b->n[d], nface = nv + 1, g = goff(b, d), npg = b->ng[d];
        int a_dim = d == 0 ? 1 : 0, b_dim = d == 2 ? 1 : 2;
        int64_t na = b->n[a_dim], nb = b->n[b_dim];
        /* pencils of one direction write disjoint rhs entries: they run in
         * parallel (OpenMP), each thread with its own pencil buffers */
#pragma omp parallel
        {
        double *pg = malloc(sizeof(double) * 5 * npg);
        double *qg = malloc(sizeof(double) * NEQ * npg);
        double *pl = calloc(5 * nface, sizeof(double));
        double *pr = calloc(5 * nface, sizeof(double));
        double *qr = calloc(NEQ * nface, sizeof(double));
        double *fl = calloc(NEQ * nface, sizeof(double));
        double *norm = calloc(3 * nface, sizeof(double));
        for (int64_t f = 0; f < nface; ++f) norm[d * nface + f] = 1.0;
#pragma omp for schedule(static)
        for (int64_t pe = 0; pe < na * nb; ++pe) {
                int64_t aa = pe % na, bb = pe / na;
                int64_t base[3];
                base[a_dim] = aa + goff(b, a_dim);
                base[b_dim] = bb + goff(b, b_dim);
                for (int64_t s = 0; s < npg; ++s) {
                    base[d] = s;
                    int64_t gi = gidx(b, base[0], base[1], base[2]);
                    memcpy(pg + 5 * s, b->p + 5 * gi, 5 * sizeof(double));
                    for (int e = 0; e < NEQ; ++e) qg[e * npg + s] = b->q[e * cs + gi];
                }
                if (c->recon == ORC_SIMPLE) {
                    orc_recon_simple(nface, g, npg, pg, pl, -1);
                    orc_recon_simple(nface, g, npg, pg, pr, 1);
                } else if (c->recon == ORC_MUSCL) {
                    orc_recon_muscl(nface, g, c->kappa, c->limiter, npg, pg, pl, -1);
                    orc_recon_muscl(nface, g, c->kappa, c->limiter, npg, pg, pr, 1);
                } else if (c->recon == ORC_WENO) {
                    orc_recon_weno(nface, NEQ, g, npg, qg, qr, -1, c->weno_textbook);
                    q_to_p_line(c->gamma, nface, qr, pl);
                    orc_recon_weno(nface, NEQ, g, npg, qg, qr, 1, c->weno_textbook);
                    q_to_p_line(c->gamma, nface, qr, pr);
                }
                if (c->flux == ORC_ROE)
                    orc_flux_roe(nface, NEQ, c->gamma, norm, pl, pr, fl, c->entfix);
                else if (c->flux == ORC_VANLEER)
                    orc_flux_vanleer(nface, NEQ, c->gamma, norm, pl, pr, fl);
                else if (c->flux == ORC_AUSM)
                    orc_flux_ausm(nface, NEQ, c->gamma, norm, pl, pr, fl);
                else if (c->flux == ORC_AUSMPLUS)
                    orc_flux_ausmplus(nface, NEQ, c->gamma, norm, pl, pr, fl);
                /* update_residual_sub */
                for (int64_t v = 0; v < nv; ++v) {
                    int64_t r[3];
                    r[a_dim] = aa;
                    r[b_dim] = bb;
                    r[d] = v;
                    int64_t ri = r[0] + b->n[0] * (r[1] + b->n[1] * r[2]);
                    for (int e = 0; e < NEQ; ++e)
                        rhs[e * nreal + ri] = rhs[e * nreal + ri]
                            - (fl[e * nface + v + 1] - fl[e * nface + v]) / c->dr[d];
                }
            }
        free(pg); free(qg); free(pl); free(pr); free(qr); free(fl); free(norm);
        }
    }
}

/* ======================================================================== */
/*  Central viscous fluxes (no reference counterpart)                       */
/* ======================================================================== */
/* The reference reserves viscous_flux_sub (class_space_discret.f90:671-675)
 * but leaves it empty and its call commented out (class_time_discret.f90:
 * 182).  This restatement defines the library's viscous term:
 *   - primitives u, v, w, T = p / (rho R) and mu = Sutherland(T) per vertex;
 *     a stencil point outside the block maps into it per direction (periodic:
 *     wrap by n, the nodal period; otherwise clamp to the boundary vertex),
 *     so no edge or corner ghost is ever read;
 *   - at the face between vertices a and b = a + e_d: face values are the
 *     two-point means, the normal derivative (phi_b - phi_a) / dr_d and each
 *     tangential derivative the mean of the central differences at a and b,
 *     ((phi_a+ - phi_a-) + (phi_b+ - phi_b-)) / (4 dr_t);
 *   - tau_ij = mu (d_j u_i + d_i u_j) - 2/3 mu div(u) delta_ij, heat flux
 *     -k dT/dx_d with k = mu cp / Pr; face flux
 *     F_d = (0, tau_d1, tau_d2, tau_d3, u_i tau_di + k dT/dx_d);
 *   - rv = sum_d (F_d(a + e_d/2) - F_d(a - e_d/2)) / dr_d, added to the
 *     inviscid rhs after its three directions. */

/* Threads the pencil loops use (OpenMP); n <= 0 leaves the default. */
int orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

double orc_sutherland(double t, double mu_ref, double t_ref, double s)
{
    double tr = t / t_ref;
    return mu_ref * tr * sqrt(tr) * (t_ref + s) / (t + s);
}

static inline int64_t vmap(int64_t x, int64_t n, int lo_bc, int hi_bc)
{
    if (x < 0) return lo_bc == ORC_BC_PERIODIC ? x + n : 0;
    if (x >= n) return hi_bc == ORC_BC_PERIODIC ? x - n : n - 1;
    return x;
}

/* primitive quantities per real vertex: u, v, w, T, mu */
typedef struct { double u, v, w, t, mu; } vprim_t;

static void visc_prims(const int64_t n[3], const orc_config *c, const double *p_aos,
                       int64_t stride_pt, vprim_t *out)
{
    int64_t nreal = n[0] * n[1] * n[2];
    for (int64_t v = 0; v < nreal; ++v) {
        const double *p = p_aos + stride_pt * v;
        vprim_t *o = out + v;
        o->u = p[1];
        o->v = p[2];
        o->w = p[3];
        o->t = p[4] / (p[0] * c->visc_r);
        o->mu = orc_sutherland(o->t, c->visc_mu_ref, c->visc_t_ref, c->visc_s);
    }
}

/* The primitive quantities stencil point (x, y, z) reads, per direction:
 * periodic wrap, wall mirror (vertex l for ghost -l; the normal velocity
 * negated for a slip wall, the first dim velocity components for a no-slip
 * wall, each crossing flipping its components again), else clamp to the
 * boundary vertex. */
static vprim_t vget(const vprim_t *P, const int64_t n[3], const orc_config *c, int64_t x,
                    int64_t y, int64_t z)
{
    int64_t r[3] = {x, y, c->dim == 3 ? z : 0};
    int flip = 0;
    for (int d = 0; d < c->dim; ++d) {
        if (r[d] >= 0 && r[d] < n[d]) continue;
        int kind = c->bc[2 * d + (r[d] < 0 ? 0 : 1)];
        int wall = wall_of(kind);
        if (kind == ORC_BC_PERIODIC) {
            r[d] = vmap(r[d], n[d], ORC_BC_PERIODIC, ORC_BC_PERIODIC);
        } else if (wall) {
            r[d] = r[d] < 0 ? -r[d] : 2 * (n[d] - 1) - r[d];
            flip ^= wall == 2 ? (1 << c->dim) - 1 : 1 << d;
        } else {
            r[d] = vmap(r[d], n[d], ORC_BC_EXTRAP, ORC_BC_EXTRAP);
        }
    }
    vprim_t o = P[r[0] + n[0] * (r[1] + n[1] * r[2])];
    if (flip & 1) o.u = -o.u;
    if (flip & 2) o.v = -o.v;
    if (flip & 4) o.w = -o.w;
    return o;
}

/* face flux between vertex a and a + e_d */
static void visc_face(const vprim_t *P, const int64_t n[3], const orc_config *c, int d,
                      const int64_t a[3], double F[4])
{
    int dim = c->dim;
    int64_t b[3] = {a[0], a[1], a[2]};
    b[d] += 1;
    const vprim_t Av = vget(P, n, c, a[0], a[1], a[2]), Bv = vget(P, n, c, b[0], b[1], b[2]);
    const vprim_t *A = &Av, *B = &Bv;
    double g[3][4];   /* g[e][q]: d/dx_e of q in (u, v, w, T) at the face */
    for (int e = 0; e < 3; ++e)
        for (int q = 0; q < 4; ++q) g[e][q] = 0.0;
    {
        const double va[4] = {A->u, A->v, A->w, A->t}, vb[4] = {B->u, B->v, B->w, B->t};
        for (int q = 0; q < 4; ++q) g[d][q] = (vb[q] - va[q]) / c->dr[d];
    }
    for (int e = 0; e < dim; ++e) {
        if (e == d) continue;
        int64_t ap[3] = {a[0], a[1], a[2]}, am[3] = {a[0], a[1], a[2]};
        int64_t bp[3] = {b[0], b[1], b[2]}, bm[3] = {b[0], b[1], b[2]};
        ap[e] += 1; am[e] -= 1; bp[e] += 1; bm[e] -= 1;
        const vprim_t Apv = vget(P, n, c, ap[0], ap[1], ap[2]), Amv = vget(P, n, c, am[0], am[1], am[2]);
        const vprim_t Bpv = vget(P, n, c, bp[0], bp[1], bp[2]), Bmv = vget(P, n, c, bm[0], bm[1], bm[2]);
        const vprim_t *Ap = &Apv, *Am = &Amv, *Bp = &Bpv, *Bm = &Bmv;
        const double vap[4] = {Ap->u, Ap->v, Ap->w, Ap->t}, vam[4] = {Am->u, Am->v, Am->w, Am->t};
        const double vbp[4] = {Bp->u, Bp->v, Bp->w, Bp->t}, vbm[4] = {Bm->u, Bm->v, Bm->w, Bm->t};
        for (int q = 0; q < 4; ++q)
            g[e][q] = ((vap[q] - vam[q]) + (vbp[q] - vbm[q])) / (4.0 * c->dr[e]);
    }
    const double mu = 0.5 * (A->mu + B->mu);
    const double uf[3] = {0.5 * (A->u + B->u), 0.5 * (A->v + B->v), 0.5 * (A->w + B->w)};
    const double div = g[0][0] + g[1][1] + g[2][2];
    double tau[3];
    for (int i = 0; i < 3; ++i) {
        /* tau_{d i} = mu (d_i u_d + d_d u_i) - 2/3 mu div delta_{di} */
        double t = mu * (g[i][d] + g[d][i]);
        if (i == d) t = t - (2.0 / 3.0) * mu * div;
        tau[i] = t;
    }
    const double cpr = c->gamma * c->visc_r / ((c->gamma - 1.0) * c->visc_pr);
    F[0] = tau[0];
    F[1] = tau[1];
    F[2] = tau[2];
    F[3] = uf[0] * tau[0] + uf[1] * tau[1] + uf[2] * tau[2] + mu * cpr * g[d][3];
}

static void visc_rhs(const int64_t n[3], const orc_config *c, const vprim_t *P, double *rv)
{
    int64_t nreal = n[0] * n[1] * n[2];
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n[2]; ++k)
        for (int64_t j = 0; j < n[1]; ++j)
            for (int64_t i = 0; i < n[0]; ++i) {
                int64_t ri = i + n[0] * (j + n[1] * k);
                double acc[4] = {0.0, 0.0, 0.0, 0.0};
                for (int d = 0; d < c->dim; ++d) {
                    int64_t a[3] = {i, j, k}, am[3] = {i, j, k};
                    am[d] -= 1;
                    double Fp[4], Fm[4];
                    visc_face(P, n, c, d, a, Fp);
                    visc_face(P, n, c, d, am, Fm);
                    for (int q = 0; q < 4; ++q)
                        acc[q] = d == 0 ? (Fp[q] - Fm[q]) / c->dr[d]
                                        : acc[q] + (Fp[q] - Fm[q]) / c->dr[d];
                }
                rv[ri] = 0.0;
                for (int q = 0; q < 4; ++q) rv[(q + 1) * nreal + ri] = acc[q];
            }
}

void orc_viscous_residual(const int64_t n[3], const orc_config *c, const double *p_real,
                          double *rv)
{
    int64_t nreal = n[0] * n[1] * n[2];
    vprim_t *P = malloc(sizeof(vprim_t) * nreal);
    visc_prims(n, c, p_real, 5, P);
    visc_rhs(n, c, P, rv);
    free(P);
}

/* rhs += viscous residual of the block's current primitives */
static void viscous_residual(blk_t *b, const orc_config *c, double *rhs)
{
    int64_t nreal = b->n[0] * b->n[1] * b->n[2];
    double *preal = malloc(sizeof(double) * 5 * nreal);
    double *rv = malloc(sizeof(double) * NEQ * nreal);
    vprim_t *P = malloc(sizeof(vprim_t) * nreal);
    for (int64_t k = 0; k < b->n[2]; ++k)
        for (int64_t j = 0; j < b->n[1]; ++j)
            for (int64_t i = 0; i < b->n[0]; ++i)
                memcpy(preal + 5 * (i + b->n[0] * (j + b->n[1] * k)),
                       b->p + 5 * gidx(b, i + NG, j + NG, k + b->gz), 5 * sizeof(double));
    visc_prims(b->n, c, preal, 5, P);
    visc_rhs(b->n, c, P, rv);
    for (int e = 1; e < NEQ; ++e)
        for (int64_t v = 0; v < nreal; ++v) rhs[e * nreal + v] = rhs[e * nreal + v] + rv[e * nreal + v];
    free(preal); free(rv); free(P);
}

double orc_cfl_dt(const int64_t n[3], int dim, const double *q_real,
                  double gamma, const double dr[3], double cfl)
{
    int64_t nreal = n[0] * n[1] * n[2];
    double smax = 0.0;
    for (int64_t c = 0; c < nreal; ++c) {
        double p[5];
        q_to_p_point(gamma, q_real + c, nreal, p);
        double a = sqrt(gamma * p[4] / p[0]);
        double s = (fabs(p[1]) + a) / dr[0] + (fabs(p[2]) + a) / dr[1];
        if (dim == 3) s = s + (fabs(p[3]) + a) / dr[2];
        if (s > smax) smax = s;
    }
    return cfl / smax;
}

/* Copy the real region of the ghost-inclusive q into a dense (n0,n1,n2,5). */
static void gather_real(const blk_t *b, double *dst)
{
    int64_t cs = b->ng[0] * b->ng[1] * b->ng[2];
    int64_t nreal = b->n[0] * b->n[1] * b->n[2];
    for (int e = 0; e < NEQ; ++e)
        for (int64_t k = 0; k < b->n[2]; ++k)
            for (int64_t j = 0; j < b->n[1]; ++j)
                for (int64_t i = 0; i < b->n[0]; ++i)
                    dst[e * nreal + i + b->n[0] * (j + b->n[1] * k)] =
                        b->q[e * cs + gidx(b, i + NG, j + NG, k + b->gz)];
}

static void scatter_real(blk_t *b, const double *src)
{
    int64_t cs = b->ng[0] * b->ng[1] * b->ng[2];
    int64_t nreal = b->n[0] * b->n[1] * b->n[2];
    for (int e = 0; e < NEQ; ++e)
        for (int64_t k = 0; k < b->n[2]; ++k)
            for (int64_t j = 0; j < b->n[1]; ++j)
                for (int64_t i = 0; i < b->n[0]; ++i)
                    b->q[e * cs + gidx(b, i + NG, j + NG, k + b->gz)] =
                        src[e * nreal + i + b->n[0] * (j + b->n[1] * k)];
}

/* Residual norms of one block, class_iteration.f90:210-248 (single block,
 * accumulators starting from zero). */
static void residual_norms(int64_t nreal, const double *rhs, double *res)
{
    double res2[NEQ], res2t = 0.0;
    for (int e = 0; e < NEQ; ++e) {
        double s = 0.0;
        for (int64_t c = 0; c < nreal; ++c) s = s + rhs[e * nreal + c] * rhs[e * nreal + c];
        res2[e] = s / (double)nreal;
        res2t = res2t + res2[e];
    }
    res2t = res2t / (double)NEQ;
    for (int e = 0; e < NEQ; ++e) res[e] = sqrt(0.0 * 0.0 + res2[e]);
    res[NEQ] = sqrt(0.0 * 0.0 + res2t);
}

static int solve_core(const int64_t n[3], const orc_config *c, double dt_fix, double cfl,
                      double t_end, int nsteps, const double *p_real, const double *q_real,
                      double *q_out, double *rhs_out, double *res_out, double *dt_out,
                      int *steps_out, double *t_out);

/* No-slip walls after a stage's combine: a vertex on such a face keeps rho,
 * loses its momentum, and its total energy loses the momentum's kinetic
 * energy 0.5 |rho u|^2 / rho (pressure and temperature unchanged).  q:
 * component-major real vertices (the stage output). */
static void noslip_fix(const int64_t n[3], const orc_config *c, double *q)
{
    int64_t nreal = n[0] * n[1] * n[2];
    int any = 0;
    for (int f = 0; f < 2 * c->dim; ++f) any = any || c->bc[f] == ORC_BC_NOSLIP;
    if (!any) return;
    for (int64_t k = 0; k < n[2]; ++k)
        for (int64_t j = 0; j < n[1]; ++j)
            for (int64_t i = 0; i < n[0]; ++i) {
                int64_t x[3] = {i, j, k};
                int on = 0;
                for (int d = 0; d < c->dim; ++d)
                    on = on || (x[d] == 0 && c->bc[2 * d] == ORC_BC_NOSLIP) ||
                         (x[d] == n[d] - 1 && c->bc[2 * d + 1] == ORC_BC_NOSLIP);
                if (!on) continue;
                int64_t v = i + n[0] * (j + n[1] * k);
                double m2 = q[nreal + v] * q[nreal + v] + q[2 * nreal + v] * q[2 * nreal + v];
                if (c->dim == 3) m2 = m2 + q[3 * nreal + v] * q[3 * nreal + v];
                double ke = 0.5 * m2 / q[v];
                q[4 * nreal + v] = q[4 * nreal + v] - ke;
                for (int e = 1; e <= c->dim; ++e) q[e * nreal + v] = 0.0;
            }
}

int orc_solve(const int64_t n[3], const orc_config *c, double dt_fix, double cfl,
              int nsteps, const double *p_real, double *q_out, double *rhs_out,
              double *res_out, double *dt_out)
{
    return solve_core(n, c, dt_fix, cfl, INFINITY, nsteps, p_real, NULL, q_out, rhs_out,
                      res_out, dt_out, NULL, NULL);
}

int orc_solve_q(const int64_t n[3], const orc_config *c, double dt_fix, double cfl,
                int nsteps, const double *q_real, double *q_out, double *rhs_out,
                double *res_out, double *dt_out)
{
    return solve_core(n, c, dt_fix, cfl, INFINITY, nsteps, NULL, q_real, q_out, rhs_out,
                      res_out, dt_out, NULL, NULL);
}

int orc_solve_to(const int64_t n[3], const orc_config *c, double cfl, double t_end,
                 int max_steps, const double *p_real, double *q_out, int *steps_out,
                 double *t_out)
{
    return solve_core(n, c, 0.0, cfl, t_end, max_steps, p_real, NULL, q_out, NULL, NULL,
                      NULL, steps_out, t_out);
}

static int solve_core(const int64_t n[3], const orc_config *c, double dt_fix, double cfl,
                      double t_end, int nsteps, const double *p_real, const double *q_real,
                      double *q_out, double *rhs_out, double *res_out, double *dt_out,
                      int *steps_out, double *t_out)
{
    blk_t b;
    for (int d = 0; d < 3; ++d) b.n[d] = n[d];
    b.gz = c->dim == 3 ? NG : 0;
    b.ng[0] = n[0] + 2 * NG;
    b.ng[1] = n[1] + 2 * NG;
    b.ng[2] = n[2] + 2 * b.gz;
    int64_t cs = b.ng[0] * b.ng[1] * b.ng[2];
    int64_t nreal = n[0] * n[1] * n[2];
    b.q = calloc(NEQ * cs, sizeof(double));
    b.p = calloc(5 * cs, sizeof(double));
    double *q0 = malloc(sizeof(double) * NEQ * nreal);
    double *qm = malloc(sizeof(double) * NEQ * nreal);
    double *acc = malloc(sizeof(double) * NEQ * nreal);
    double *rhs = malloc(sizeof(double) * NEQ * nreal);
    if (!b.q || !b.p || !q0 || !qm || !acc || !rhs) return -1;

    /* initial state: p on real vertices, q = p -> q, optional p = q -> p;
     * or a conserved state, p = q -> p (what every stage ends with) */
    if (q_real) {
        scatter_real(&b, q_real);
        update_primitives(&b, c->gamma);
    } else {
        for (int64_t k = 0; k < n[2]; ++k)
            for (int64_t j = 0; j < n[1]; ++j)
                for (int64_t i = 0; i < n[0]; ++i) {
                    int64_t gi = gidx(&b, i + NG, j + NG, k + b.gz);
                    const double *src = p_real + 5 * (i + n[0] * (j + n[1] * k));
                    memcpy(b.p + 5 * gi, src, 5 * sizeof(double));
                    p_to_q_point(c->gamma, b.p + 5 * gi, b.q + gi, cs);
                }
        if (c->p_from_q) update_primitives(&b, c->gamma);
    }
    fill_ghosts(&b, c->bc, c->dim);

    int nstage = c->rk == ORC_RK4 ? 4 : 3;
    double t = 0.0;
    int step = 0;
    for (; step < nsteps; ++step) {
        /* an end time: stop once t is within 1e-14 |t_end| of it, clip the
         * last CFL step to it, t accumulated step by step */
        if (isfinite(t_end) && !(t < t_end - 1.0e-14 * fabs(t_end))) break;
        gather_real(&b, q0);
        double dt = dt_fix;
        if (dt_fix <= 0.0) dt = orc_cfl_dt(n, c->dim, q0, c->gamma, c->dr, cfl);
        if (isfinite(t_end) && dt > t_end - t) dt = t_end - t;
        t = t + dt;
        if (dt_out) dt_out[step] = dt;
        for (int s = 1; s <= nstage; ++s) {
            memset(rhs, 0, sizeof(double) * NEQ * nreal);
            inviscid_residual(&b, c, rhs);
            if (c->visc_mu_ref > 0.0) viscous_residual(&b, c, rhs);
            gather_real(&b, qm);
            if (c->rk == ORC_RK4) {
                orc_rk4_stage(s, dt, NEQ * nreal, q0, qm, acc, rhs);
            } else {
                orc_runge_kutta(s, dt, NEQ * nreal, q0, qm, rhs);
            }
            noslip_fix(n, c, qm);
            scatter_real(&b, qm);
            update_primitives(&b, c->gamma);
            fill_ghosts(&b, c->bc, c->dim);
        }
        if (res_out) residual_norms(nreal, rhs, res_out + (NEQ + 1) * step);
    }
    if (steps_out) *steps_out = step;
    if (t_out) *t_out = t;
    if (q_out) memcpy(q_out, b.q, sizeof(double) * NEQ * cs);
    if (rhs_out) memcpy(rhs_out, rhs, sizeof(double) * NEQ * nreal);
    free(b.q); free(b.p); free(q0); free(qm); free(acc); free(rhs);
    return 0;
}
