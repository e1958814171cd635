/*
 * fvm_gpu.h -- C ABI of the MI355X explicit residual + Runge-Kutta path.
 *
 * Drop-in for the solver's RK-stage call of FVM-3D-F90: the Fortran driver
 * (or any C/Python host) keeps case input, mesh and scheme selection, uploads
 * the conserved state once, and calls fvm_step / fvm_stage where the
 * reference calls ClassTimeDiscret%Runge_Kutta_sub
 * (src/solve/class_time_discret.f90:159-196, called from solve_sub :146 in
 * the time loop src/process/solver_sub.f90:196-206).  State stays in HBM
 * between calls.
 *
 * Conventions
 *   - Every call returns 0 on success and a negative FVM_E* code on failure;
 *     fvm_last_error() gives the text (the reference's ClassError flags plus
 *     mpi%stop(msg), src/core/global_error.f90:82-146,
 *     src/core/parallel_mpi.f90:111-133, become return codes; nothing
 *     aborts the process).
 *   - Host arrays are Fortran column-major q(i, j, k, eq) with eq = rho,
 *     rho*u, rho*v, rho*w, rho*E (the equation order of
 *     src/process/solver_sub.f90:131-136) and host_ng ghost layers on each
 *     side of every swept direction (the reference allocates 3,
 *     src/base/defs_ghost_cell.f90:42).  A 2-D block (nk == 1) has no k
 *     ghosts on the host; its rho*w must be identically zero.
 *   - Scheme codes are the reference's (src/defs/type_defs_scheme.f90:11-36);
 *     codes >= 1000 are additions with no reference counterpart.
 *   - One host thread per context.  Contexts are independent.
 */
#ifndef FVM_GPU_H
#define FVM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- scheme codes ------------------------------------------------------ */
#define FVM_RK_SSP3        101   /* "Runge-Kutta": 3-stage Shu-Osher SSP-RK3  */
#define FVM_RK4            1001  /* classical RK4 (no reference counterpart)   */
#define FVM_SIMPLE         200
#define FVM_MUSCL          201
#define FVM_WENO           202
#define FVM_LIM_NONE       250
#define FVM_LIM_MINMOD     251
#define FVM_LIM_VANLEER    252
#define FVM_LIM_VANALBADA  253
#define FVM_ROE            300
#define FVM_VANLEER        301
#define FVM_AUSM           303
#define FVM_AUSMPLUS       1303  /* Liou's AUSM+ (no reference counterpart)    */

/* ---- boundary kinds (fvm_block_set_bc) --------------------------------- */
#define FVM_BC_EXTRAP      0     /* zeroth-order ghost copy, the reference's only BC */
#define FVM_BC_PERIODIC    1     /* nodal periodic: vertex n is vertex 0       */
/* the reference's BC codes (src/defs/type_defs_bc.f90:11-30), which it
 * declares but never applies: */
#define FVM_BC_INFLOW             10  /* fixed ghost state (fvm_block_set_bc_state) */
#define FVM_BC_INFLOW_SUBSONIC    11  /* fixed ghost state                          */
#define FVM_BC_INFLOW_SUPERSONIC  12  /* fixed ghost state                          */
#define FVM_BC_OUTFLOW            20  /* extrapolation                               */
#define FVM_BC_OUTFLOW_SUBSONIC   21  /* extrapolation                               */
#define FVM_BC_OUTFLOW_SUPERSONIC 22  /* extrapolation                               */
#define FVM_BC_WALL               30  /* inviscid (slip) wall: mirror, normal momentum negated */
#define FVM_BC_WALL_INVISCID      31  /* as FVM_BC_WALL                              */
#define FVM_BC_WALL_VISCOUS_ISOTHERMAL 32  /* refused                                  */
#define FVM_BC_WALL_VISCOUS       33  /* adiabatic no-slip wall: mirror, every momentum
                                         component negated; each stage sets the wall
                                         vertices' momentum to 0 (kinetic energy taken
                                         out of the total energy)                   */
#define FVM_BC_FARFIELD           40  /* fixed free-stream ghost state (fvm_block_set_bc_state) */
#define FVM_BC_EXTRAPOLATE        41  /* extrapolation                               */
#define FVM_BC_SYMMETRY           50  /* mirror, normal momentum negated             */

/* ---- error codes ------------------------------------------------------- */
#define FVM_OK             0
#define FVM_EINVAL        -1     /* bad argument                              */
#define FVM_EHIP          -2     /* HIP runtime error                         */
#define FVM_ENOMEM        -3     /* device allocation failed                  */
#define FVM_ESTATE        -4     /* call out of order (e.g. no scheme set)     */
#define FVM_ENAN          -5     /* NaN in the CFL wave speed                 */

/* ---- snapshot variables ------------------------------------------------ */
/* the reference's output variables (src/defs/type_defs_variables.f90:25-30,
 * written by src/solve/class_output.f90:398-410) */
#define FVM_VAR_DENSITY     201
#define FVM_VAR_VELOCITY_X  202
#define FVM_VAR_VELOCITY_Y  203
#define FVM_VAR_VELOCITY_Z  204
#define FVM_VAR_PRESSURE    301
/* conserved components, copied unchanged (restart state; no reference code) */
#define FVM_VAR_RHO         1
#define FVM_VAR_RHOU        2
#define FVM_VAR_RHOV        3
#define FVM_VAR_RHOW        4
#define FVM_VAR_RHOE        5

typedef struct fvm_ctx fvm_ctx;
typedef struct fvm_block fvm_block;

/* Text of the last error on the calling thread ("" if none). */
const char *fvm_last_error(void);

/* Library version string (no reference counterpart). */
const char *fvm_version(void);

/* Context on HIP device `device`: owns one stream and every block created on
 * it.  Replaces the solver singletons of src/solve/global_class.f90. */
int fvm_ctx_create(int device, fvm_ctx **out);
int fvm_ctx_destroy(fvm_ctx *ctx);

/* Scheme selection, ClassSpaceDiscret%input / ClassTimeDiscret%input
 * (src/solve/class_space_discret.f90:72-94,
 *  src/solve/class_time_discret.f90:66-82):
 *   recon   FVM_SIMPLE | FVM_MUSCL | FVM_WENO
 *   limiter FVM_LIM_*  (MUSCL only)
 *   flux    FVM_ROE | FVM_VANLEER | FVM_AUSM | FVM_AUSMPLUS
 *   rk      FVM_RK_SSP3 | FVM_RK4
 *   kappa, entropy_fix as read from the case file (a missing key reads 0.0)
 *   ref_compat 1: WENO keeps the reference's IS_1 and unmirrored right state
 *              (src/scheme/scheme_WENO.f90:54-56,107-108); 0: textbook WENO5-JS */
int fvm_scheme_set(fvm_ctx *ctx, int recon, int limiter, int flux, int rk,
                   double kappa, double entropy_fix, double gamma, int ref_compat);

/* Numerics of the stage kernels (no reference counterpart):
 *   FVM_NUMERICS_EXACT (default): the reference's IEEE operation order, no
 *     FMA contraction; results are bit-identical to the CPU path.
 *   FVM_NUMERICS_FAST: the same formulas with FMA contraction and
 *     reciprocal multiplies; relative L-inf difference from the exact path
 *     <= 1e-12 after one step and <= 1e-9 after 1000 steps (tested). */
#define FVM_NUMERICS_EXACT 0
#define FVM_NUMERICS_FAST  1
int fvm_set_numerics(fvm_ctx *ctx, int mode);

/* One structured block of ni x nj x nk real vertices (nk == 1: 2-D, swept
 * along i and j only), vertex spacing dx, dy, dz (the reference's uniform
 * cell_size, src/base/method_gridgen.f90:120-131).  Replaces
 * ClassBlockData%allocate_memory (src/base/defs_block_data.f90:111-163). */
int fvm_block_create(fvm_ctx *ctx, int64_t ni, int64_t nj, int64_t nk,
                     double dx, double dy, double dz, fvm_block **out);
int fvm_block_destroy(fvm_block *blk);

/* Boundary kind of one block face: face 0..5 = i-, i+, j-, j+, k-, k+.
 * Default FVM_BC_EXTRAP (src/base/defs_block_data.f90:226-256).  Periodic
 * applies to both faces of the direction.  Wall and symmetry faces mirror
 * the block about the boundary vertex plane (ghost l = vertex l inside,
 * momentum normal to the face negated); the no-slip wall (FVM_BC_WALL_VISCOUS)
 * mirrors with every momentum component negated and sets the wall vertices'
 * momentum to zero after every stage, removing its kinetic energy from the
 * total energy (the isothermal no-slip kind 32 is refused); outflow and
 * extrapolate kinds copy the boundary vertex.  Inflow and farfield faces take
 * a state: fvm_block_set_bc_state.  A face's kind replaces its previous one
 * (connected faces excepted).  Wall, symmetry and state faces are refilled
 * on the device after every stage.  With viscous terms the stencil crosses a
 * wall face to the mirror vertex, its normal velocity (slip, symmetry) or
 * every velocity component (no-slip) negated; inflow and farfield faces are
 * not available with viscous terms. */
int fvm_block_set_bc(fvm_block *blk, int face, int kind);

/* Inflow / farfield face (FVM_BC_INFLOW*, FVM_BC_FARFIELD): every ghost
 * layer holds the primitive state prim = (rho, u, v, w, p) along the block's
 * axes, converted to conserved variables as the reference's initial state
 * is (src/solve/class_flow_field.f90:353-377; 2-D blocks: w must be 0). */
int fvm_block_set_bc_state(fvm_block *blk, int face, int kind, const double prim[5]);

/* Block-to-block connectivity (the reference builds the metadata,
 * ClassGeomBlock%connect src/base/defs_block_geom.f90:953-990 and donor ranges
 * src/base/defs_face_geom.f90:43-63,361-403, but never exchanges data; here
 * the ghost layers across the face are filled on the device, inside the stage
 * kernel).  Face `face` of block a abuts the opposite face of block b; the two
 * faces must conform (equal extents in the other directions, same
 * orientation).  shared = 1: both blocks hold the interface vertex plane (a
 * grid cut through a vertex plane, defs_block_geom.f90:610-611); shared = 0:
 * the blocks are one spacing apart (e.g. a periodic wrap across blocks).
 * Both blocks must belong to the same context. */
int fvm_block_connect(fvm_block *a, int face, fvm_block *b, int shared);

/* One 1-to-1 connection patch in the reference's face metadata
 * (ClassGeomFace range_min/range_max, donor_type, donor_range_min/max,
 * src/base/defs_face_geom.f90:43-63; donor ranges from the transform,
 * defs_face_geom.f90:384-403), as a CGNS 1-to-1 connection carries it
 * (src/base/interface_cgns.f90:105-112: range, donor range, transform).
 * Indices are 1-based block vertex indices, inclusive.  range_min..range_max
 * is a rectangle on a boundary face of `recv` (one vertex plane in the face
 * direction); donor_type[d] = +-(1 + e): this face's direction d runs along
 * donor direction e, forward or backward; donor_range_min[d] is the donor
 * index along that direction that corresponds to range_min[d] (so
 * donor_range_max[d] - donor_range_min[d] = sign * (range_max[d] -
 * range_min[d])).  The interface vertex plane belongs to both blocks: ghost
 * layer l of `recv` receives the donor vertex l planes inside its face.  The
 * patches on a face must tile it exactly; a face may mix donors.  The ghosts
 * are refilled on the device after every stage.  2-D blocks: donor_type[2]
 * = 3 and k ranges 1..1. */
int fvm_conn_add(fvm_block *recv, const int64_t range_min[3], const int64_t range_max[3],
                 fvm_block *donor, const int64_t donor_range_min[3],
                 const int64_t donor_range_max[3], const int donor_type[3]);

/* Cut positions of a uniform block cut along one direction
 * (ClassGeomBlock%cut_uniform, src/base/defs_block_geom.f90:757-769): for
 * vertices range_min..range_max and ncut pieces, pos[0] = range_min,
 * pos[c] = pos[c-1] + (range_max - pos[c-1]) / (ncut - c + 1) (integer
 * division), pos[ncut] = range_max; piece c spans vertices pos[c]..pos[c+1],
 * so neighbours share their interface plane (connect them with shared = 1).
 * Requires 1 <= ncut <= range_max - range_min.  Host only (no device). */
int fvm_cut_uniform(int64_t range_min, int64_t range_max, int ncut, int64_t *pos);

/* How a stage is launched over the blocks of a context (no reference
 * counterpart):
 *   FVM_LAUNCH_BATCHED (default): one launch per stage over the tile list of
 *     all blocks (blocks of uneven size and independent cases alike);
 *   FVM_LAUNCH_PERSISTENT: one launch of at most two workgroups per CU that
 *     loop over the tile list;
 *   FVM_LAUNCH_PER_BLOCK: one launch per block. */
#define FVM_LAUNCH_PER_BLOCK  0
#define FVM_LAUNCH_BATCHED    1
#define FVM_LAUNCH_PERSISTENT 2
int fvm_set_launch_mode(fvm_ctx *ctx, int mode);

/* Initial state computed on the device, with no host array: the block's
 * u^n at vertices x = origin[0] + i dx, y = origin[1] + j dy, z = origin[2]
 * + k dz (i, j, k from 0; the reference's grid is x = dx (i - 1),
 * src/base/method_gridgen.f90:120-131, i.e. origin 0), converted p -> q as
 * the reference does (src/solve/class_flow_field.f90:353-377) with the
 * context's gamma (fvm_scheme_set first).  origin may be NULL (0, 0, 0).
 *  FVM_IC_REGIONS: the reference's init_value (src/base/defs_block_data.f90:
 *    166-201): per region xmin, xmax, ymin, ymax, zmin, zmax, rho, u, v, w, p
 *    (11 values, up to 16 regions), applied in order with inclusive bounds,
 *    so a later region wins on shared lines; vertices in no region get zero
 *    (cover the block).  The reference's four-quadrant case is its regions
 *    1-4 (src/solve/class_flow_field.f90:207-243).
 *  FVM_IC_VORTEX: Yee's isentropic vortex, free stream rho = p = 1;
 *    params length (period in x and y), beta, xc, yc, u_inf, v_inf, w_inf.
 *  FVM_IC_TGV: the Taylor-Green vortex on [0, 2 pi)^3, params Mach number:
 *    u = sin x cos y cos z, v = -cos x sin y cos z, w = 0,
 *    p = R + (cos 2x + cos 2y)(cos 2z + 2)/16, rho = p / R, R = 1/(gamma Ma^2).
 * (No reference counterpart for the vortex and Taylor-Green states.) */
#define FVM_IC_REGIONS  1
#define FVM_IC_VORTEX   2
#define FVM_IC_TGV      3
#define FVM_IC_MAX_PARAMS 176   /* 16 regions of 11 values */
int fvm_block_init(fvm_block *blk, int kind, int nparams, const double *params,
                   const double origin[3]);
/* The same for a block that is one piece of a larger grid whose first vertex
 * is vertex first[] (0-based) of that grid: x = dx (first[0] + i), ... one
 * product as the reference's x = dx (i - 1) (src/base/method_gridgen.f90:
 * 120-131), so the pieces of a cut (fvm_cut_uniform) get the whole grid's
 * initial state bit for bit.  first may be NULL (0, 0, 0). */
int fvm_block_init_at(fvm_block *blk, int kind, int nparams, const double *params,
                      const int64_t first[3]);

/* Host <-> device copies of the conserved state (host layout above).  Upload
 * also fills the device ghost layers (ClassFlowField%initial's ghost update,
 * src/solve/class_flow_field.f90:249-253). */
int fvm_upload(fvm_block *blk, const double *q, int host_ng);
int fvm_download(fvm_block *blk, double *q, int host_ng);

/* The same copies for a host array in the reference's own allocation,
 * ClassBlockData%q (src/base/defs_block_data.f90:137-141):
 * q(ni + 2 host_ng, nj + 2 host_ng, host_nk + 2 host_ng, 5), host_nk the
 * reference's real k-planes.  Its 2-D grid has two identical k-planes
 * (gsize(3) = 2, src/base/method_gridgen.f90:98-99), so its q holds 2 + 6
 * planes; the device block is the 2-D block (nk == 1).  Upload reads real
 * plane k_plane (1-based among the host_nk real planes; its rho*w must be
 * zero).  Download writes the state into every real plane and the real
 * vertices of the boundary planes into the host_ng ghost planes on each
 * side (the reference's k ghost update, defs_block_data.f90:226-256; edges
 * and corners zero).  A 3-D block needs host_nk == nk and then equals
 * fvm_upload / fvm_download (host_nk 0 is accepted by download too). */
int fvm_upload_planes(fvm_block *blk, const double *q, int host_ng, int64_t host_nk,
                      int64_t k_plane);
int fvm_download_planes(fvm_block *blk, double *q, int host_ng, int64_t host_nk);

/* nsteps full time steps on every block of the context
 * (ClassTimeDiscret%solve, src/solve/class_time_discret.f90:125-156).
 *   dt > 0         fixed time step (the reference's "fixed" timestep_set)
 *   dt <= 0        global CFL step dt = cfl / max sum_d (|u_d|+c)/dr_d,
 *                  recomputed from the state at the start of every step
 *                  (no reference counterpart: "CFL" is unimplemented there)
 * dt_last (nullable) receives the last step's dt.  Residual norms of the
 * last stage of the last step are kept for fvm_norms. */
int fvm_step(fvm_ctx *ctx, int nsteps, double dt, double cfl, double *dt_last);

/* One Runge-Kutta stage s (1..3 SSP-RK3, 1..4 RK4) with a fixed dt > 0: the
 * body of the stage loop of ClassTimeDiscret%Runge_Kutta_sub
 * (src/solve/class_time_discret.f90:172-195: residual, combine, q -> p and
 * ghosts).  Stages are called in order with the same dt; stage 1 takes the
 * current state as the step's base state (reset_q0, :170), the last stage
 * writes u^(n+1), advances the solution time by dt and leaves the norms for
 * fvm_norms.  The stage values live in device work arrays, so between
 * stages fvm_download returns u^n.  Results equal fvm_step's bit for bit.
 * Out of order, or fvm_step / fvm_autotune while a step is in progress:
 * FVM_ESTATE; an upload or device initial state abandons the step. */
int fvm_stage(fvm_ctx *ctx, int s, double dt);

/* fvm_step for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The same steps and results as fvm_step (bit
 * for bit), but the step's own device memory -- block table, time slot and
 * residual-norm partials, which fvm_step allocates in the context on first
 * use -- is a static array of 2.7 MB per GPU, held until the process ends;
 * the call allocates nothing (device or host).  Limits: at most 64 blocks
 * and 65536 stage tiles per context (a 128^3 block at the default launch
 * shape has 512 tiles), no connection patches or wall / symmetry / inflow /
 * farfield faces; beyond them FVM_EINVAL.  Stage timing and hipGraph replay
 * do not apply.  The array is shared by every call on a GPU: calls that
 * overlap on one GPU (other threads, other contexts' streams) race. */
int fvm_step_noalloc(fvm_ctx *ctx, int nsteps, double dt, double cfl, double *dt_last);

/* Central second-order viscous fluxes with Sutherland's law (no reference
 * counterpart: viscous_flux_sub, src/solve/class_space_discret.f90:671-675,
 * is an empty stub and its call at src/solve/class_time_discret.f90:182 is
 * commented out).  mu = mu_ref (T/t_ref)^(3/2) (t_ref + s)/(T + s),
 * T = p/(rho r_gas), heat conductivity mu cp/prandtl, cp = gamma r_gas/(gamma-1).
 * mu_ref <= 0 switches the terms off.  On blocks joined by fvm_block_connect
 * a stencil point beyond a connected face is the neighbour's vertex (edge and
 * corner points through the neighbour's neighbour), so a cut grid gives the
 * results of the grid it was cut from.  Wall and symmetry faces mirror the
 * stencil (fvm_block_set_bc); across a connection patch (fvm_conn_add) the
 * point is the donor vertex the patch's index map gives, its velocity
 * rotated into the receiver's axes.  Not with inflow or farfield faces. */
int fvm_viscous_set(fvm_ctx *ctx, double mu_ref, double t_ref, double sutherland_s,
                    double prandtl, double r_gas);

/* Replay each time step of fvm_step from a captured hipGraph (no reference
 * counterpart): one graph per context holds the CFL dt update and the stage
 * launches of a step; it is captured on first use and again after any
 * change of scheme, numerics, launch mode, blocks or launch shapes.  Results
 * are identical to eager launches.  Steps whose stages are timed
 * (fvm_timing_enable) are launched eagerly.  Off by default. */
int fvm_set_graph(fvm_ctx *ctx, int enable);

/* Solution time and end time (the time loop's time/total_time,
 * src/process/solver_sub.f90:196-206).  With an end time, CFL steps clip
 * the last dt to t_end - t and the steps of an fvm_step call that would
 * start at or past t_end are skipped on the device (no host round trip per
 * step).  t_end <= t or infinite: no end.  fvm_time reports the time after
 * the last step and the number of steps run since the context was made. */
int fvm_set_time(fvm_ctx *ctx, double t, double t_end);
int fvm_time(fvm_ctx *ctx, double *t, int64_t *steps);

/* Residual norms of the last stage of the last fvm_step call, as
 * ClassIteration%compute_residual computes them
 * (src/solve/class_iteration.f90:210-248): out[0..4] per equation,
 * out[5] total; nan_flag = 1 if any is NaN or above 1e10
 * (converged_fun's abort tests, :183-191). */
int fvm_norms(fvm_ctx *ctx, double out[6], int *nan_flag);

/* Autotuner (no reference counterpart): times `reps` sweeps of one time
 * step's stage kinds for every compiled launch shape and k-chunk (all blocks
 * together in the batched launch modes, else per block), in two passes of
 * which each candidate keeps its faster time (the first pass also warms the
 * GPU), and keeps the fastest.  The probe writes only stage buffers; the
 * state is unchanged.  Call after fvm_upload.  fvm_block_launch_shape and
 * fvm_block_launch_tile report the choice. */
int fvm_autotune(fvm_ctx *ctx, int reps);
int fvm_block_launch_shape(fvm_block *blk, int *tile_x, int *tile_y, int *chunk);
/* The launch-shape id behind that shape (no reference counterpart; the
 * FVM_TILE / FVM_TILE2 values):
 * 3-D ids 0-3 the plane-staged stage kernel at 64x4, 32x8, 64x2, 16x16; ids
 * 4-5 the register-reconstruction kernel at 32x8, 16x16 (simple and MUSCL
 * reconstructions); 2-D id 0 the row kernel, 1-2 the tiled kernel. */
int fvm_block_launch_tile(fvm_block *blk, int *tile_id);

/* Stage-kernel timing with HIP events on the context's stream (no
 * reference counterpart; the reference only wall-clocks the whole solve,
 * src/solve/class_output.f90:266-334).  enable = n >= 1 brackets the stage
 * launches of every n-th step with one event pair (an event costs the stream
 * a few microseconds, so none sits between the stages and a sample keeps
 * them out of short stages' wall time); 0 turns it off.  Enabling resets the
 * counters; fvm_timing_read returns the summed milliseconds of the timed
 * steps' launches (stage kernels, with the viscous-residual and patch
 * launches between them where those are on), the number of stage launches
 * they hold and the cell updates they performed (steps skipped past t_end
 * excluded). */
int fvm_timing_enable(fvm_ctx *ctx, int enable);
int fvm_timing_read(fvm_ctx *ctx, double *stage_ms, int64_t *launches, double *cell_updates);

/* ---- serving: continuous batching of independent solves ---------------
 * (no reference counterpart).  A server owns the device state of many
 * independent solves.  Each fvm_server_poll iteration admits waiting
 * requests in arrival order while a slot and the HBM budget allow, advances
 * every active solve by one time step -- solves with the same scheme and
 * dimension share one batched launch per RK stage, each with its own CFL
 * time step (dt = min(cfl / max wave speed, t_end - t)) -- and retires the
 * solves that reached t_end, copying their state to host memory and freeing
 * their device memory.  A solve's result does not depend on what else is
 * active (same kernels, per-solve dt; tested bit for bit). */
typedef struct fvm_server fvm_server;
typedef struct fvm_request {
  int64_t n[3];             /* vertices; n[2] == 1: 2-D                         */
  double spacing[3];
  int recon, limiter, flux, rk;
  double kappa, entropy_fix, gamma;
  int ref_compat, numerics;
  int bc[6];                /* FVM_BC_EXTRAP (or the outflow / extrapolate
                               codes) or FVM_BC_PERIODIC per face; other kinds
                               are refused                                      */
  double t_end, cfl;
  const double *q0;         /* (5, nk, nj, ni) conserved state                  */
  int borrow;               /* 0: q0 is copied by fvm_server_submit.  1: q0 is
                               copied later by the server's staging thread; it
                               must stay valid and unchanged until the request
                               is done (fvm_server_poll counts it done)         */
  int ic_kind;              /* 0: q0 holds the initial state.  FVM_IC_*: the
                               state is computed on the device at admission
                               (fvm_block_init's kinds and parameters, gamma of
                               the request); q0 is not read                     */
  int ic_nparams;
  double ic_params[FVM_IC_MAX_PARAMS];
  double ic_origin[3];
} fvm_request;

/* hbm_budget_bytes <= 0: 80 % of the free device memory. */
int fvm_server_create(int device, double hbm_budget_bytes, int max_active, fvm_server **out);
int fvm_server_destroy(fvm_server *srv);
int fvm_server_submit(fvm_server *srv, const fvm_request *req, int64_t *id);
/* Run up to `iterations` scheduler iterations; each admits pending
 * requests, advances every active solve by `steps_per_sync` steps with one
 * host synchronisation (a solve reaching t_end halts on the device), and
 * retires finished solves.  Returns early when nothing is active. */
int fvm_server_poll(fvm_server *srv, int iterations, int steps_per_sync, int *n_pending,
                    int *n_active, int *n_done);
/* Final state (same layout as q0) and info[8] = {submit, admit, done (s since
 * server creation), steps, cell-updates, final time, status (2 done, -1
 * diverged, -2 not staged: no pinned host memory), device bytes}.  The first
 * call for a request waits for its asynchronous download. */
int fvm_server_result(fvm_server *srv, int64_t id, double *q_out, double info[8]);

/* Wait for all work of the context (no reference counterpart). */
int fvm_sync(fvm_ctx *ctx);

/* Asynchronous snapshot of the solution (the reference writes the output
 * variables of every block at output times, ClassOutput%write_flowfield,
 * src/solve/class_output.f90:204-237, from get_data :377-410, and names a
 * restart file :194-195).  begin: at the current point of the context's
 * stream a kernel writes the chosen variables of u^n of every block into a
 * device slot, and a copy stream moves the slot into pinned host memory
 * while later steps run; returns at once.  Up to 4 snapshots in flight; a
 * slot is reused after release.  vars: FVM_VAR_* codes (up to 10).
 * ready: 1 when the copy has landed (does not block).
 * read: waits for the copy, then copies block blk's part into out as
 * vdata(ni, nj, nk, nvars) in Fortran order (= C array [nvars][nk][nj][ni]),
 * and the solution time and step count of the snapshot (blk may be NULL to
 * read time and step only).
 * save: writes the snapshot to a file on a background thread once the copy
 * has landed (format below); release waits for the write and returns its
 * status.  File: "FVMSNAP1", int32 nblocks, int32 nvars, int32 codes[nvars],
 * float64 time, int64 step, per block int64 ni, nj, nk, float64 dx, dy, dz,
 * then per block the nvars arrays, i fastest, little endian.
 * Variables 201-301 convert q -> p as the reference does
 * (src/solve/class_flow_field.f90:383-396); 1-5 copy the conserved state
 * bit for bit. */
int fvm_snapshot_begin(fvm_ctx *ctx, int nvars, const int *vars, int64_t *id);
int fvm_snapshot_ready(fvm_ctx *ctx, int64_t id, int *ready);
int fvm_snapshot_read(fvm_ctx *ctx, int64_t id, fvm_block *blk, double *out, double *time,
                      int64_t *step);
int fvm_snapshot_save(fvm_ctx *ctx, int64_t id, const char *path);
int fvm_snapshot_release(fvm_ctx *ctx, int64_t id);

/* Resume from a snapshot file holding the conserved state (variables 1..5
 * in order): the context must hold the same blocks (count, order, extents
 * and spacing) with their boundaries and connections set; uploads every
 * block's state and sets the solution time and step count (the reference's
 * restart file, src/solve/class_output.f90:194-195, has no reader). */
int fvm_restart_load(fvm_ctx *ctx, const char *path);

#ifdef __cplusplus
}
#endif
#endif /* FVM_GPU_H */
