!=============================================================================80
!
!> ISO_C_BINDING interface to the MI355X residual + RK library
!> (include/fvm_gpu.h).  A Fortran solver calls fvm_step where FVM-3D-F90
!> calls ClassTimeDiscret%Runge_Kutta_sub (src/solve/class_time_discret.f90:
!> 159-196); the conserved state stays on the device between calls.
!
!  Host arrays keep the reference's layout: q(i, j, k, eq), column-major,
!  eq = rho, rho*u, rho*v, rho*w, rho*E, with ng ghost layers per swept side
!  (src/base/defs_block_data.f90:111-163).
!
!=============================================================================80
module fvm_gpu

  use iso_c_binding
  implicit none
  private

!...Scheme codes, src/defs/type_defs_scheme.f90:11-36 (>= 1000: additions)
  integer(c_int), parameter, public :: FVM_RK_SSP3 = 101, FVM_RK4 = 1001
  integer(c_int), parameter, public :: FVM_SIMPLE = 200, FVM_MUSCL = 201, FVM_WENO = 202
  integer(c_int), parameter, public :: FVM_LIM_NONE = 250, FVM_LIM_MINMOD = 251
  integer(c_int), parameter, public :: FVM_LIM_VANLEER = 252, FVM_LIM_VANALBADA = 253
  integer(c_int), parameter, public :: FVM_ROE = 300, FVM_VANLEER = 301, FVM_AUSM = 303
  integer(c_int), parameter, public :: FVM_AUSMPLUS = 1303
  integer(c_int), parameter, public :: FVM_BC_EXTRAP = 0, FVM_BC_PERIODIC = 1
!...the reference's BC codes (src/defs/type_defs_bc.f90:11-30)
  integer(c_int), parameter, public :: FVM_BC_INFLOW = 10, FVM_BC_INFLOW_SUBSONIC = 11
  integer(c_int), parameter, public :: FVM_BC_INFLOW_SUPERSONIC = 12, FVM_BC_OUTFLOW = 20
  integer(c_int), parameter, public :: FVM_BC_OUTFLOW_SUBSONIC = 21
  integer(c_int), parameter, public :: FVM_BC_OUTFLOW_SUPERSONIC = 22, FVM_BC_WALL = 30
  integer(c_int), parameter, public :: FVM_BC_WALL_INVISCID = 31, FVM_BC_FARFIELD = 40
  integer(c_int), parameter, public :: FVM_BC_EXTRAPOLATE = 41, FVM_BC_SYMMETRY = 50
  integer(c_int), parameter, public :: FVM_BC_WALL_VISCOUS_ISOTHERMAL = 32, FVM_BC_WALL_VISCOUS = 33
  integer(c_int), parameter, public :: FVM_NUMERICS_EXACT = 0, FVM_NUMERICS_FAST = 1
  integer(c_int), parameter, public :: FVM_LAUNCH_PER_BLOCK = 0, FVM_LAUNCH_BATCHED = 1
  integer(c_int), parameter, public :: FVM_LAUNCH_PERSISTENT = 2
!...Initial states computed on the device (fvm_block_init)
  integer(c_int), parameter, public :: FVM_IC_REGIONS = 1, FVM_IC_VORTEX = 2, FVM_IC_TGV = 3
!...Snapshot variables: the reference's output variables
!   (src/defs/type_defs_variables.f90:25-30) and the conserved components
  integer(c_int), parameter, public :: FVM_VAR_DENSITY = 201, FVM_VAR_VELOCITY_X = 202
  integer(c_int), parameter, public :: FVM_VAR_VELOCITY_Y = 203, FVM_VAR_VELOCITY_Z = 204
  integer(c_int), parameter, public :: FVM_VAR_PRESSURE = 301
  integer(c_int), parameter, public :: FVM_VAR_RHO = 1, FVM_VAR_RHOU = 2, FVM_VAR_RHOV = 3
  integer(c_int), parameter, public :: FVM_VAR_RHOW = 4, FVM_VAR_RHOE = 5

  public :: fvm_ctx_create, fvm_ctx_destroy, fvm_scheme_set, fvm_set_numerics
  public :: fvm_block_create, fvm_block_destroy, fvm_block_set_bc, fvm_block_set_bc_state
  public :: fvm_block_init, fvm_block_init_at, fvm_cut_uniform, fvm_block_launch_tile
  public :: fvm_upload, fvm_download, fvm_step, fvm_norms, fvm_sync
  public :: fvm_upload_planes, fvm_download_planes, fvm_stage
  public :: fvm_block_connect, fvm_conn_add, fvm_viscous_set, fvm_set_time, fvm_time, fvm_autotune
  public :: fvm_set_launch_mode, fvm_set_graph
  public :: fvm_snapshot_begin, fvm_snapshot_ready, fvm_snapshot_read, fvm_snapshot_save
  public :: fvm_snapshot_release, fvm_restart_load
  public :: fvm_server_create, fvm_server_destroy, fvm_server_submit, fvm_server_poll
  public :: fvm_server_result
  public :: fvm_check, fvm_error_text, scheme_index

!...A served solve (fvm_request of include/fvm_gpu.h, same layout)
  integer(c_int), parameter, public :: FVM_IC_MAX_PARAMS = 176
  type, bind(C), public :: fvm_request
    integer(c_int64_t) :: n(3) = 0
    real(c_double)     :: spacing(3) = 0.0_c_double
    integer(c_int)     :: recon = 0, limiter = 0, flux = 0, rk = 0
    real(c_double)     :: kappa = 0.0_c_double, entropy_fix = 0.0_c_double, gamma = 1.4_c_double
    integer(c_int)     :: ref_compat = 1, numerics = 0
    integer(c_int)     :: bc(6) = 0
    real(c_double)     :: t_end = 0.0_c_double, cfl = 0.0_c_double
    type(c_ptr)        :: q0 = c_null_ptr
    integer(c_int)     :: borrow = 0
    integer(c_int)     :: ic_kind = 0, ic_nparams = 0
    real(c_double)     :: ic_params(176) = 0.0_c_double
    real(c_double)     :: ic_origin(3) = 0.0_c_double
  end type fvm_request

  interface
    function fvm_last_error() bind(C, name="fvm_last_error") result(msg)
      import :: c_ptr
      type(c_ptr) :: msg
    end function fvm_last_error

    function fvm_ctx_create(device, ctx) bind(C, name="fvm_ctx_create") result(rc)
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: rc
    end function fvm_ctx_create

    function fvm_ctx_destroy(ctx) bind(C, name="fvm_ctx_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function fvm_ctx_destroy

    function fvm_scheme_set(ctx, recon, limiter, flux, rk, kappa, entropy_fix, gamma, &
                            ref_compat) bind(C, name="fvm_scheme_set") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: recon, limiter, flux, rk, ref_compat
      real(c_double), value :: kappa, entropy_fix, gamma
      integer(c_int) :: rc
    end function fvm_scheme_set

    function fvm_set_numerics(ctx, mode) bind(C, name="fvm_set_numerics") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
      integer(c_int) :: rc
    end function fvm_set_numerics

    function fvm_block_create(ctx, ni, nj, nk, dx, dy, dz, blk) &
        bind(C, name="fvm_block_create") result(rc)
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: ni, nj, nk
      real(c_double), value :: dx, dy, dz
      type(c_ptr), intent(out) :: blk
      integer(c_int) :: rc
    end function fvm_block_create

    function fvm_block_destroy(blk) bind(C, name="fvm_block_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: blk
      integer(c_int) :: rc
    end function fvm_block_destroy

    function fvm_block_set_bc(blk, face, kind) bind(C, name="fvm_block_set_bc") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: blk
      integer(c_int), value :: face, kind
      integer(c_int) :: rc
    end function fvm_block_set_bc

    function fvm_upload(blk, q, host_ng) bind(C, name="fvm_upload") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: blk
      type(c_ptr), value :: q
      integer(c_int), value :: host_ng
      integer(c_int) :: rc
    end function fvm_upload

    function fvm_download(blk, q, host_ng) bind(C, name="fvm_download") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: blk
      type(c_ptr), value :: q
      integer(c_int), value :: host_ng
      integer(c_int) :: rc
    end function fvm_download

    !> ClassBlockData%q as the reference allocates it (defs_block_data.f90:
    !> 137-141), q(ni+2ng, nj+2ng, host_nk+2ng, 5); a 2-D block reads real
    !> plane k_plane and writes every plane back
    function fvm_upload_planes(blk, q, host_ng, host_nk, k_plane) &
        bind(C, name="fvm_upload_planes") result(rc)
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: blk
      type(c_ptr), value :: q
      integer(c_int), value :: host_ng
      integer(c_int64_t), value :: host_nk, k_plane
      integer(c_int) :: rc
    end function fvm_upload_planes

    function fvm_download_planes(blk, q, host_ng, host_nk) &
        bind(C, name="fvm_download_planes") result(rc)
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: blk
      type(c_ptr), value :: q
      integer(c_int), value :: host_ng
      integer(c_int64_t), value :: host_nk
      integer(c_int) :: rc
    end function fvm_download_planes

    !> One Runge-Kutta stage (the stage loop of class_time_discret.f90:172-195)
    function fvm_stage(ctx, s, dt) bind(C, name="fvm_stage") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: s
      real(c_double), value :: dt
      integer(c_int) :: rc
    end function fvm_stage

    function fvm_step(ctx, nsteps, dt, cfl, dt_last) bind(C, name="fvm_step") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
      real(c_double), value :: dt, cfl
      real(c_double), intent(out) :: dt_last
      integer(c_int) :: rc
    end function fvm_step

    function fvm_norms(ctx, res, nan_flag) bind(C, name="fvm_norms") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: res(6)
      integer(c_int), intent(out) :: nan_flag
      integer(c_int) :: rc
    end function fvm_norms

    function fvm_server_create(device, hbm_budget, max_active, srv) &
        bind(C, name="fvm_server_create") result(rc)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: device, max_active
      real(c_double), value :: hbm_budget
      type(c_ptr), intent(out) :: srv
      integer(c_int) :: rc
    end function fvm_server_create

    function fvm_server_destroy(srv) bind(C, name="fvm_server_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: srv
      integer(c_int) :: rc
    end function fvm_server_destroy

    function fvm_server_submit(srv, req, id) bind(C, name="fvm_server_submit") result(rc)
      import :: c_int, c_int64_t, c_ptr, fvm_request
      type(c_ptr), value :: srv
      type(fvm_request), intent(in) :: req
      integer(c_int64_t), intent(out) :: id
      integer(c_int) :: rc
    end function fvm_server_submit

    function fvm_server_poll(srv, iterations, steps_per_sync, n_pending, n_active, n_done) &
        bind(C, name="fvm_server_poll") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: srv
      integer(c_int), value :: iterations, steps_per_sync
      integer(c_int), intent(out) :: n_pending, n_active, n_done
      integer(c_int) :: rc
    end function fvm_server_poll

    ! q_out: c_loc of q(ni, nj, nk, 5); info(8) as include/fvm_gpu.h lists it
    function fvm_server_result(srv, id, q_out, info) bind(C, name="fvm_server_result") result(rc)
      import :: c_int, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: srv, q_out
      integer(c_int64_t), value :: id
      real(c_double), intent(out) :: info(8)
      integer(c_int) :: rc
    end function fvm_server_result

    function fvm_sync(ctx) bind(C, name="fvm_sync") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function fvm_sync

    ! params: FVM_IC_REGIONS 11 values per region (xmin, xmax, ymin, ymax, zmin,
    ! zmax, rho, u, v, w, p); origin: c_loc of 3 reals or c_null_ptr
    function fvm_block_init(blk, kind, nparams, params, origin) bind(C, name="fvm_block_init") &
        result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: blk, origin
      integer(c_int), value :: kind, nparams
      real(c_double), intent(in) :: params(*)
      integer(c_int) :: rc
    end function fvm_block_init

    ! cut positions of ClassGeomBlock%cut_uniform along one direction: pos(0:ncut)
    function fvm_cut_uniform(range_min, range_max, ncut, pos) bind(C, name="fvm_cut_uniform") &
        result(rc)
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: range_min, range_max
      integer(c_int), value :: ncut
      integer(c_int64_t), intent(out) :: pos(*)
      integer(c_int) :: rc
    end function fvm_cut_uniform

    function fvm_block_launch_tile(blk, tile_id) bind(C, name="fvm_block_launch_tile") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: blk
      integer(c_int), intent(out) :: tile_id
      integer(c_int) :: rc
    end function fvm_block_launch_tile

    ! a block whose first vertex is vertex first (0-based) of a larger grid
    function fvm_block_init_at(blk, kind, nparams, params, first) bind(C, name="fvm_block_init_at") &
        result(rc)
      import :: c_int, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: blk
      integer(c_int), value :: kind, nparams
      real(c_double), intent(in) :: params(*)
      integer(c_int64_t), intent(in) :: first(3)
      integer(c_int) :: rc
    end function fvm_block_init_at

    ! prim = (rho, u, v, w, p) of an inflow / farfield face
    function fvm_block_set_bc_state(blk, face, kind, prim) bind(C, name="fvm_block_set_bc_state") &
        result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: blk
      integer(c_int), value :: face, kind
      real(c_double), intent(in) :: prim(5)
      integer(c_int) :: rc
    end function fvm_block_set_bc_state

    function fvm_block_connect(a, face, b, shared) bind(C, name="fvm_block_connect") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: a, b
      integer(c_int), value :: face, shared
      integer(c_int) :: rc
    end function fvm_block_connect

    ! the reference's face metadata (range_min/max, donor_range_min/max,
    ! donor_type), 1-based vertex indices
    function fvm_conn_add(recv, range_min, range_max, donor, donor_range_min, donor_range_max, &
                          donor_type) bind(C, name="fvm_conn_add") result(rc)
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: recv, donor
      integer(c_int64_t), intent(in) :: range_min(3), range_max(3)
      integer(c_int64_t), intent(in) :: donor_range_min(3), donor_range_max(3)
      integer(c_int), intent(in) :: donor_type(3)
      integer(c_int) :: rc
    end function fvm_conn_add

    function fvm_snapshot_begin(ctx, nvars, vars, id) bind(C, name="fvm_snapshot_begin") result(rc)
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nvars
      integer(c_int), intent(in) :: vars(*)
      integer(c_int64_t), intent(out) :: id
      integer(c_int) :: rc
    end function fvm_snapshot_begin

    function fvm_snapshot_ready(ctx, id, ready) bind(C, name="fvm_snapshot_ready") result(rc)
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: id
      integer(c_int), intent(out) :: ready
      integer(c_int) :: rc
    end function fvm_snapshot_ready

    ! out: c_loc of vdata(ni, nj, nk, nvars) (or c_null_ptr with blk = c_null_ptr)
    function fvm_snapshot_read(ctx, id, blk, out, time, step) bind(C, name="fvm_snapshot_read") &
        result(rc)
      import :: c_int, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx, blk, out
      integer(c_int64_t), value :: id
      real(c_double), intent(out) :: time
      integer(c_int64_t), intent(out) :: step
      integer(c_int) :: rc
    end function fvm_snapshot_read

    function fvm_snapshot_save(ctx, id, path) bind(C, name="fvm_snapshot_save") result(rc)
      import :: c_int, c_int64_t, c_ptr, c_char
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: id
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int) :: rc
    end function fvm_snapshot_save

    function fvm_snapshot_release(ctx, id) bind(C, name="fvm_snapshot_release") result(rc)
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: id
      integer(c_int) :: rc
    end function fvm_snapshot_release

    function fvm_restart_load(ctx, path) bind(C, name="fvm_restart_load") result(rc)
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int) :: rc
    end function fvm_restart_load

    function fvm_viscous_set(ctx, mu_ref, t_ref, sutherland_s, prandtl, r_gas) &
        bind(C, name="fvm_viscous_set") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: mu_ref, t_ref, sutherland_s, prandtl, r_gas
      integer(c_int) :: rc
    end function fvm_viscous_set

    function fvm_set_time(ctx, t, t_end) bind(C, name="fvm_set_time") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: t, t_end
      integer(c_int) :: rc
    end function fvm_set_time

    function fvm_time(ctx, t, steps) bind(C, name="fvm_time") result(rc)
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: t
      integer(c_int64_t), intent(out) :: steps
      integer(c_int) :: rc
    end function fvm_time

    function fvm_autotune(ctx, reps) bind(C, name="fvm_autotune") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: reps
      integer(c_int) :: rc
    end function fvm_autotune

    function fvm_set_launch_mode(ctx, mode) bind(C, name="fvm_set_launch_mode") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
      integer(c_int) :: rc
    end function fvm_set_launch_mode

    function fvm_set_graph(ctx, enable) bind(C, name="fvm_set_graph") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
      integer(c_int) :: rc
    end function fvm_set_graph

    function c_strlen(s) bind(C, name="strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function c_strlen
  end interface

contains

  !> Text of the library's last error on this thread.
  function fvm_error_text() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: i, n
    p = fvm_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function fvm_error_text

  !> Stop with the library's message on a nonzero return, as the reference
  !> stops on ClassError flags (src/core/parallel_mpi.f90:111-133).
  subroutine fvm_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= 0) then
      write(*, '(a, a, a, i0, a, a)') "fvm_gpu error in ", where, " (", rc, "): ", &
        fvm_error_text()
      error stop 1
    end if
  end subroutine fvm_check

  !> Case-file spelling -> scheme code (src/defs/type_defs_scheme.f90:41-84);
  !> -1 (the reference's NULL) for an unknown name.
  integer(c_int) function scheme_index(name) result(code)
    character(len=*), intent(in) :: name
    select case (trim(name))
      case ("Runge-Kutta");   code = FVM_RK_SSP3
      case ("RK4");           code = FVM_RK4
      case ("simple");        code = FVM_SIMPLE
      case ("MUSCL");         code = FVM_MUSCL
      case ("WENO");          code = FVM_WENO
      case ("none_");         code = FVM_LIM_NONE
      case ("minmod_");       code = FVM_LIM_MINMOD
      case ("VanLeer_");      code = FVM_LIM_VANLEER
      case ("vanAlbada_");    code = FVM_LIM_VANALBADA
      case ("Roe");           code = FVM_ROE
      case ("VanLeer");       code = FVM_VANLEER
      case ("AUSM");          code = FVM_AUSM
      case ("AUSM+");         code = FVM_AUSMPLUS
      case default;           code = -1
    end select
  end function scheme_index

end module fvm_gpu
