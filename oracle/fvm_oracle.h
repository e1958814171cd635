/*
 * fvm_oracle.h -- CPU restatement of the FVM-3D-F90 explicit residual + RK
 * path.  TEST INFRASTRUCTURE ONLY: the test suite, the smoke check and the
 * CPU-baseline leg of bench.py load liboracle.so to check results.  Nothing in
 * the product links or calls it.
 *
 * Scheme codes are the reference's (src/defs/type_defs_scheme.f90:11-36);
 * codes >= 1000 are additions of this library with no reference counterpart.
 */
#ifndef FVM_ORACLE_H
#define FVM_ORACLE_H
#include <stdint.h>

#define ORC_RK_SSP3       101   /* "Runge-Kutta": Shu-Osher SSP-RK3       */
#define ORC_RK4           1001  /* classical 4-stage RK (no reference)     */
#define ORC_SIMPLE        200
#define ORC_MUSCL         201
#define ORC_WENO          202
#define ORC_LIM_NONE      250
#define ORC_LIM_MINMOD    251
#define ORC_LIM_VANLEER   252
#define ORC_LIM_VANALBADA 253
#define ORC_ROE           300
#define ORC_VANLEER       301
#define ORC_AUSM          303
#define ORC_AUSMPLUS      1303  /* Liou's AUSM+ (no reference)             */

#define ORC_BC_EXTRAP     0     /* zeroth-order copy of the boundary vertex */
#define ORC_BC_PERIODIC   1     /* nodal periodic: vertex n == vertex 0     */
/* Walls (the reference's codes, type_defs_bc.f90:11-30, which it never
 * applies; this library's definition): ghost -l is the mirror vertex l */
#define ORC_BC_WALL       30    /* slip: normal velocity negated            */
#define ORC_BC_WALL_INVISCID 31 /* as ORC_BC_WALL                           */
#define ORC_BC_NOSLIP     33    /* adiabatic no-slip: every velocity component
                                   negated; after each stage the wall vertices'
                                   momentum is set to 0 and its kinetic energy
                                   taken out of the total energy             */
#define ORC_BC_SYMMETRY   50    /* as ORC_BC_WALL                           */

/* ---- per-pencil scheme routines (same argument meaning as the reference) */
void orc_recon_simple(int64_t nface, int64_t dghost, int64_t npg,
                      const double *pg, double *pr, int wind);
void orc_recon_muscl(int64_t nface, int64_t dghost, double kappa, int limiter,
                     int64_t npg, const double *pg, double *pr, int wind);
void orc_recon_weno(int64_t nface, int neq, int64_t dghost, int64_t npg,
                    const double *qg, double *qr, int wind, int textbook);
void orc_flux_roe(int64_t nface, int neq, double gamma, const double *norm,
                  const double *ql, const double *qr, double *flux, double entfix);
void orc_flux_vanleer(int64_t nface, int neq, double gamma, const double *norm,
                      const double *ql, const double *qr, double *flux);
void orc_flux_ausm(int64_t nface, int neq, double gamma, const double *norm,
                   const double *ql, const double *qr, double *flux);
void orc_flux_ausmplus(int64_t nface, int neq, double gamma, const double *norm,
                       const double *ql, const double *qr, double *flux);
void orc_runge_kutta(int m, double dt, int64_t n, double *u0, double *um,
                     const double *rm);

/* Classical RK4 stage s = 1..4 with an accumulator (no reference
 * counterpart); the combine orc_solve applies for rk = ORC_RK4. */
void orc_rk4_stage(int s, double dt, int64_t n, const double *u0, double *um,
                   double *acc, const double *rm);

/* ---- whole-block solver ------------------------------------------------ */
typedef struct orc_config {
    int dim;              /* 2: sweep i,j (k planes independent); 3: i,j,k   */
    int recon, limiter, flux, rk;
    int weno_textbook;    /* 0: reference WENO (IS_1 '+', unmirrored right)  */
    int p_from_q;         /* 1: first stage uses convert(q), 0: IC primitives */
    int bc[6];            /* i-,i+,j-,j+,k-,k+                               */
    double kappa, entfix, gamma;
    double dr[3];
    /* central viscous fluxes (no reference counterpart: viscous_flux_sub is
     * an empty stub there); visc_mu_ref <= 0 switches them off */
    double visc_mu_ref;   /* Sutherland reference viscosity at visc_t_ref      */
    double visc_t_ref;
    double visc_s;        /* Sutherland constant (same units as T)             */
    double visc_pr;       /* Prandtl number                                    */
    double visc_r;        /* gas constant: T = p / (rho R), cp = gamma R/(gamma-1) */
} orc_config;

/* Threads of the pencil loops (OpenMP; results do not depend on it);
 * n <= 0 keeps the current setting.  Returns the thread count in use. */
int orc_set_threads(int n);

/* Sutherland's law mu(T) = mu_ref (T/T_ref)^(3/2) (T_ref + S) / (T + S). */
double orc_sutherland(double t, double mu_ref, double t_ref, double s);

/* Viscous residual of a primitive field alone (for tests): p AoS
 * (5, n0, n1, n2) real vertices, rv (n0, n1, n2, 5) out (component 0 = 0). */
void orc_viscous_residual(const int64_t n[3], const orc_config *cfg, const double *p_real,
                          double *rv);

/* Runs nsteps steps of fixed dt (dt > 0) or CFL dt (dt <= 0, cfl > 0) on one
 * block of n[0] x n[1] x n[2] real vertices.
 *   p_real  AoS (5, n0, n1, n2) initial primitives
 *   q_out   (n0+6, n1+6, n2+2gz, 5) Fortran order, gz = 3 if dim == 3 else 0
 *   rhs_out (n0, n1, n2, 5) rhs of the last stage
 *   res_out (6, nsteps) absolute residual norms of the last stage
 *   dt_out  (nsteps) the dt used per step                                 */
int orc_solve(const int64_t n[3], const orc_config *cfg, double dt, double cfl,
              int nsteps, const double *p_real, double *q_out, double *rhs_out,
              double *res_out, double *dt_out);

/* orc_solve from a conserved state instead of primitives: q_real
 * component-major (5, n2, n1, n0) real vertices; primitives are derived
 * from it (the p_from_q path).  A step loop's state is its real conserved
 * vertices, so nsteps from the q_out of an earlier call continue that run
 * bit for bit (a long run in chunks). */
int orc_solve_q(const int64_t n[3], const orc_config *cfg, double dt, double cfl,
                int nsteps, const double *q_real, double *q_out, double *rhs_out,
                double *res_out, double *dt_out);

/* CFL steps to an end time, as the device time slot of a served solve takes
 * them (no reference counterpart: the reference's CFL time step is
 * unimplemented, src/solve/class_time_discret.f90:94-107): per step
 * dt = min(cfl / smax, t_end - t), t = t + dt, until t >= t_end - 1e-14
 * |t_end| or max_steps steps; steps_out and t_out report them.  Outputs as
 * orc_solve's q_out. */
int orc_solve_to(const int64_t n[3], const orc_config *cfg, double cfl, double t_end,
                 int max_steps, const double *p_real, double *q_out, int *steps_out,
                 double *t_out);

/* Global CFL time step of a state (no reference counterpart):
 *   dt = cfl / max_v sum_d (|u_d| + c) / dr_d,  c = sqrt(gamma p / rho)   */
double orc_cfl_dt(const int64_t n[3], int dim, const double *q_real,
                  double gamma, const double dr[3], double cfl);

#endif
