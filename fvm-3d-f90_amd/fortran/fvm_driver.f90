!=============================================================================80
!
!> Fortran driver of the MI355X residual + RK path.
!
!  Reads an FVM-3D-F90 case file (namelist groups General, Geometry,
!  Inviscid_Flux, Time_Stepping, ... as in case/Riemann-2D.nml), builds the
!  uniform nodal grid (src/base/method_gridgen.f90:84-136) and the four-
!  quadrant initial state (src/solve/class_flow_field.f90:195-256), then runs
!  the time loop of solve_process (src/process/solver_sub.f90:177-217) with
!  the RK stages on the GPU through the fvm_gpu module.  Residuals are
!  printed as the reference's terminal table (src/solve/class_output.f90:
!  414-463); NaN or |R| > 1e10 stops the run (class_iteration.f90:183-191).
!
!  Usage:  fvm_driver --input=<case.nml>
!  Extra group (optional, not in the reference):
!    &GPU  device = 0, numerics = "exact" | "fast", cfl = 0.5,
!          max_step = 0 (0: no limit), print_every = 1,
!          output_file = "" (raw dump of the final q with 3 ghost layers),
!          snapshot_prefix = "" (write the &Output variables at the output
!          times of &Time_Stepping to <prefix>_<step>.fvmsnap),
!          restart_file = "" (resume from this conserved-state snapshot: its
!          time and step count carry on), checkpoint_file = "" (write the
!          conserved state there at the end: the reference's .rst file,
!          src/solve/class_output.f90:194-195, as a snapshot),
!          ncut = 1, 1, 1 (run the grid as ncut(1) x ncut(2) x ncut(3) blocks
!          of the reference's uniform cut, joined across their shared vertex
!          planes; output_file still holds the whole grid) /
!
!=============================================================================80
program fvm_driver

  use iso_c_binding
  use fvm_gpu
  implicit none

  integer, parameter :: dp = c_double, NG = 3

!...Case file (namelist names as the reference spells them)
  character(len=64) :: name = "", dimension = "", equation = "", gasModel = ""
  logical :: dimensionless = .true.
  character(len=64) :: type = "", geoType = "", gridType = ""
  real(dp) :: geoSize = 1.0_dp
  integer :: gridSize(3) = [0, 0, 0]
  logical :: writeGrid = .false.
  character(len=64) :: reconstruct = "", limiter = "", flux_split = ""
  real(dp) :: kappa = 0.0_dp, entropy_fix = 0.0_dp
  character(len=64) :: time_scheme = "", timestep_set = ""
  real(dp) :: fix_timestep = 0.0_dp, total_time = 0.0_dp, output_time = 0.0_dp
  real(dp) :: absolute_residual = 0.0_dp, relative_residual = 0.0_dp
  character(len=64) :: format = ""
  logical :: Density = .false., VelocityX = .false., VelocityY = .false., VelocityZ = .false.
  logical :: Pressure = .false.
  integer :: device = 0, max_step = 0, print_every = 1
  character(len=16) :: numerics = "exact"
  real(dp) :: cfl = 0.5_dp
  character(len=256) :: output_file = "", snapshot_prefix = ""
  character(len=256) :: restart_file = "", checkpoint_file = ""
  integer :: ncut(3) = [1, 1, 1]

  namelist /General/ name, dimension, dimensionless, equation, gasModel
  namelist /Geometry/ type, geoType, geoSize, gridType, gridSize, writeGrid
  namelist /Inviscid_Flux/ reconstruct, kappa, limiter, flux_split, entropy_fix
  ! the empty groups of the case file (Viscous_Flux, Boundary_Condition,
  ! Initial_Condition) carry nothing and are skipped by the namelist reads
  namelist /Time_Stepping/ time_scheme, timestep_set, fix_timestep, total_time, &
                           output_time, absolute_residual, relative_residual
  namelist /Output/ format, Density, VelocityX, VelocityY, VelocityZ, Pressure
  namelist /GPU/ device, numerics, cfl, max_step, print_every, output_file, snapshot_prefix, &
                 restart_file, checkpoint_file, ncut

  character(len=512) :: arg, casefile
  integer :: unit, ios, i, j, k, ni, nj, nk, step, nan_flag, eq
  real(dp) :: dx, dy, dz, gamma, time, dt, dt_used, res(6), res0(6)
  logical :: have_res0
  real(dp), allocatable, target :: q(:,:,:,:)
  type(c_ptr) :: ctx
  type(c_ptr), allocatable :: pieces(:,:,:)
  integer(c_int64_t), allocatable :: px(:), py(:), pz(:)
  integer(c_int64_t) :: first(3)
  integer :: a, b, c
  real(dp) :: regions(11, 4), time_out
  integer(c_int) :: vars(5), nvars
  integer(c_int64_t) :: snaps(4), step0
  integer :: nsnap
  logical :: writing
  integer(c_int) :: rc, recon_c, lim_c, flux_c, rk_c, mode
  logical :: converged

!...--input=<file> (src/base/method_argparse.f90:175-197)
  casefile = ""
  do i = 1, command_argument_count()
    call get_command_argument(i, arg)
    if (arg(1:8) == "--input=") casefile = arg(9:)
  end do
  if (len_trim(casefile) == 0) then
    write(*, '(a)') "usage: fvm_driver --input=<case.nml>"
    error stop 2
  end if

  open(newunit=unit, file=trim(casefile), status="old", action="read", iostat=ios)
  if (ios /= 0) then
    write(*, '(a, a)') "cannot open case file ", trim(casefile)
    error stop 2
  end if
  read(unit, nml=General, iostat=ios);          call need(ios, "General")
  rewind(unit); read(unit, nml=Geometry, iostat=ios);      call need(ios, "Geometry")
  rewind(unit); read(unit, nml=Inviscid_Flux, iostat=ios); call need(ios, "Inviscid_Flux")
  rewind(unit); read(unit, nml=Time_Stepping, iostat=ios); call need(ios, "Time_Stepping")
  rewind(unit); read(unit, nml=Output, iostat=ios)         ! optional group
  rewind(unit); read(unit, nml=GPU, iostat=ios)            ! optional group
  close(unit)

!...Grid: N vertices per direction on [0, geoSize], dx = L / (N - 1)
  gamma = 1.4_dp
  ni = gridSize(1); nj = gridSize(2)
  if (trim(dimension) == "2D") then
    nk = 1
  else
    nk = gridSize(3)
  end if
  if (ni < 2 .or. nj < 2 .or. nk < 1) then
    write(*, '(a)') "invalid gridSize"
    error stop 2
  end if
  dx = geoSize / real(ni - 1, dp)
  dy = geoSize / real(nj - 1, dp)
  dz = 1.0_dp
  if (nk > 1) dz = geoSize / real(nk - 1, dp)

!...Initial state: the four quadrants of src/solve/class_flow_field.f90:
!   207-243, regions applied in order with inclusive bounds (later regions
!   win on the x = 0.8 / y = 0.8 lines), computed on the device
  regions(:, 1) = [0.8_dp, 1.0_dp, 0.8_dp, 1.0_dp, -huge(1.0_dp), huge(1.0_dp), &
                   1.5_dp, 0.0_dp, 0.0_dp, 0.0_dp, 1.5_dp]
  regions(:, 2) = [0.0_dp, 0.8_dp, 0.8_dp, 1.0_dp, -huge(1.0_dp), huge(1.0_dp), &
                   0.5323_dp, 1.206_dp, 0.0_dp, 0.0_dp, 0.3_dp]
  regions(:, 3) = [0.0_dp, 0.8_dp, 0.0_dp, 0.8_dp, -huge(1.0_dp), huge(1.0_dp), &
                   0.138_dp, 1.206_dp, 1.206_dp, 0.0_dp, 0.029_dp]
  regions(:, 4) = [0.8_dp, 1.0_dp, 0.0_dp, 0.8_dp, -huge(1.0_dp), huge(1.0_dp), &
                   0.5323_dp, 0.0_dp, 1.206_dp, 0.0_dp, 0.3_dp]
  if (nk == 1) then
    allocate(q(ni + 2*NG, nj + 2*NG, 1, 5))
  else
    allocate(q(ni + 2*NG, nj + 2*NG, nk + 2*NG, 5))
  end if

!...Scheme selection (codes of src/defs/type_defs_scheme.f90)
  recon_c = scheme_index(reconstruct)
  lim_c = scheme_index(limiter)
  flux_c = scheme_index(flux_split)
  rk_c = scheme_index(time_scheme)
  if (recon_c < 0 .or. flux_c < 0 .or. rk_c < 0) then
    write(*, '(a)') "invalid scheme in case file"
    error stop 2
  end if
  mode = FVM_NUMERICS_EXACT
  if (trim(numerics) == "fast") mode = FVM_NUMERICS_FAST

  call fvm_check(fvm_ctx_create(int(device, c_int), ctx), "ctx_create")
  call fvm_check(fvm_scheme_set(ctx, recon_c, lim_c, flux_c, rk_c, kappa, entropy_fix, &
                                gamma, 1_c_int), "scheme_set")
  call fvm_check(fvm_set_numerics(ctx, mode), "set_numerics")

!...The grid as one block, or cut into ncut(1) x ncut(2) x ncut(3) pieces
!   (ClassGeomBlock%cut_uniform, src/base/defs_block_geom.f90:757-769) that
!   share their interface vertex planes; each piece gets the initial state
!   at its offset in the whole grid
  if (any(ncut < 1) .or. (nk == 1 .and. ncut(3) /= 1)) then
    write(*, '(a)') "invalid ncut"
    error stop 2
  end if
  allocate(px(0:ncut(1)), py(0:ncut(2)), pz(0:ncut(3)))
  call fvm_check(fvm_cut_uniform(1_c_int64_t, int(ni, c_int64_t), int(ncut(1), c_int), px), "cut i")
  call fvm_check(fvm_cut_uniform(1_c_int64_t, int(nj, c_int64_t), int(ncut(2), c_int), py), "cut j")
  if (nk == 1) then
    pz = 1
  else
    call fvm_check(fvm_cut_uniform(1_c_int64_t, int(nk, c_int64_t), int(ncut(3), c_int), pz), "cut k")
  end if
  allocate(pieces(ncut(1), ncut(2), ncut(3)))
  do c = 1, ncut(3)
    do b = 1, ncut(2)
      do a = 1, ncut(1)
        call fvm_check(fvm_block_create(ctx, px(a) - px(a-1) + 1, py(b) - py(b-1) + 1, &
                                        pz(c) - pz(c-1) + 1, dx, dy, dz, pieces(a, b, c)), &
                       "block_create")
        first = [px(a-1) - 1, py(b-1) - 1, pz(c-1) - 1]
        call fvm_check(fvm_block_init_at(pieces(a, b, c), FVM_IC_REGIONS, 44_c_int, regions, &
                                         first), "init")
        if (a > 1) call fvm_check(fvm_block_connect(pieces(a-1, b, c), 1_c_int, &
                                                    pieces(a, b, c), 1_c_int), "connect i")
        if (b > 1) call fvm_check(fvm_block_connect(pieces(a, b-1, c), 3_c_int, &
                                                    pieces(a, b, c), 1_c_int), "connect j")
        if (c > 1) call fvm_check(fvm_block_connect(pieces(a, b, c-1), 5_c_int, &
                                                    pieces(a, b, c), 1_c_int), "connect k")
      end do
    end do
  end do

!...Output variables of &Output (src/solve/class_output.f90:337-374)
  nvars = 0
  if (Density) call add_var(FVM_VAR_DENSITY)
  if (VelocityX) call add_var(FVM_VAR_VELOCITY_X)
  if (VelocityY) call add_var(FVM_VAR_VELOCITY_Y)
  if (VelocityZ) call add_var(FVM_VAR_VELOCITY_Z)
  if (Pressure) call add_var(FVM_VAR_PRESSURE)
  writing = len_trim(snapshot_prefix) > 0 .and. nvars > 0 .and. output_time > 0.0_dp
  nsnap = 0
  time_out = 0.0_dp

!...Time loop (src/process/solver_sub.f90:196-206)
  write(*, '(a, a, a, i0, a, i0, a, i0, a, i0)') "case ", trim(name), ": ", ni, " x ", nj, " x ", nk, &
    ", blocks ", size(pieces)
  write(*, '(a8, a14, 5a14)') "Step", "Time", "R_rho", "R_momx", "R_momy", "R_energy", "R_total"
  time = 0.0_dp
  step = 0
  if (len_trim(restart_file) > 0) then
    call fvm_check(fvm_restart_load(ctx, trim(restart_file) // c_null_char), "restart")
    call fvm_check(fvm_time(ctx, time, step0), "time")
    step = int(step0)
    write(*, '(a, a, a, i0, a, es14.6)') "restarted from ", trim(restart_file), " at step ", &
      step, ", t = ", time
    ! the output times already passed (written by the run that saved the
    ! restart file), accumulated as write_output accumulates them
    if (writing) then
      do while (time_out <= time + 1.0e-9_dp)
        time_out = time_out + output_time
      end do
    end if
  end if
  have_res0 = .false.
  converged = .false.
  if (writing .and. step == 0) call write_output()          ! the initial flowfield
  do while (.not. converged)
    step = step + 1
    if (trim(timestep_set) == "CFL") then
      dt = -1.0_dp
    else
      dt = fix_timestep
    end if
    rc = fvm_step(ctx, 1_c_int, dt, cfl, dt_used)
    call fvm_check(rc, "step")
    time = time + dt_used
    ! ClassOutput%write_flowfield: at the output times, within 1e-9
    if (writing) then
      if (abs(time - time_out) <= 1.0e-9_dp) call write_output()
    end if
    call fvm_check(fvm_norms(ctx, res, nan_flag), "norms")
    if (.not. have_res0) then   ! relative residuals against this run's first step
      res0 = res
      have_res0 = .true.
    end if
    if (nan_flag /= 0) then
      write(*, '(a, es12.4)') "Error: ABSRES ", res(6)
      error stop 3
    end if
    if (mod(step, max(print_every, 1)) == 0) &
      write(*, '(i8, es14.6, 5es14.6)') step, time, (res(eq) / res0(eq), eq = 1, 3), &
        res(5) / res0(5), res(6) / res0(6)
    if (max_step > 0 .and. step >= max_step) converged = .true.
    if (time > total_time - 1.0e-10_dp) converged = .true.
    if (res(6) < absolute_residual) converged = .true.
  end do
  write(*, '(a, i0, a, es14.6)') "finished after ", step, " steps at t = ", time

  do i = 1, nsnap
    call fvm_check(fvm_snapshot_release(ctx, snaps(i)), "snapshot write")
  end do
  if (len_trim(checkpoint_file) > 0) then
    block
      integer(c_int) :: cons(5)
      integer(c_int64_t) :: sid
      cons = [FVM_VAR_RHO, FVM_VAR_RHOU, FVM_VAR_RHOV, FVM_VAR_RHOW, FVM_VAR_RHOE]
      call fvm_check(fvm_snapshot_begin(ctx, 5_c_int, cons, sid), "checkpoint")
      call fvm_check(fvm_snapshot_save(ctx, sid, trim(checkpoint_file) // c_null_char), "checkpoint")
      call fvm_check(fvm_snapshot_release(ctx, sid), "checkpoint write")
    end block
  end if
  do c = 1, ncut(3)
    do b = 1, ncut(2)
      do a = 1, ncut(1)
        call gather_piece(a, b, c)
      end do
    end do
  end do
  if (len_trim(output_file) > 0) then
    open(newunit=unit, file=trim(output_file), access="stream", form="unformatted", &
         status="replace", action="write")
    write(unit) int(ni, c_int32_t), int(nj, c_int32_t), int(nk, c_int32_t), &
                int(NG, c_int32_t), step, time
    write(unit) q
    close(unit)
  end if
  call fvm_check(fvm_ctx_destroy(ctx), "ctx_destroy")

contains

  subroutine need(ios, group)
    integer, intent(in) :: ios
    character(len=*), intent(in) :: group
    if (ios /= 0) then
      write(*, '(a, a, a, i0)') "case file: cannot read group &", group, ", iostat ", ios
      error stop 2
    end if
  end subroutine need

  subroutine add_var(code)
    integer(c_int), intent(in) :: code
    nvars = nvars + 1
    vars(nvars) = code
  end subroutine add_var

  !> ClassOutput%write_flowfield (src/solve/class_output.f90:204-237): an
  !! asynchronous snapshot of the output variables, saved on a background
  !! thread; at most four in flight
  subroutine write_output()
    integer(c_int64_t) :: sid
    character(len=300) :: fname
    integer :: m
    if (nsnap == size(snaps)) then
      call fvm_check(fvm_snapshot_release(ctx, snaps(1)), "snapshot write")
      do m = 1, nsnap - 1
        snaps(m) = snaps(m + 1)
      end do
      nsnap = nsnap - 1
    end if
    call fvm_check(fvm_snapshot_begin(ctx, nvars, vars, sid), "snapshot")
    write(fname, '(a, "_", i6.6, ".fvmsnap")') trim(snapshot_prefix), step
    call fvm_check(fvm_snapshot_save(ctx, sid, trim(fname) // c_null_char), "snapshot save")
    nsnap = nsnap + 1
    snaps(nsnap) = sid
    time_out = time_out + output_time
  end subroutine write_output

  !> Download piece (a, b, c) with its ghost layers and place it in the
  !! whole-grid array q: its vertices, and its ghost layers on the faces of
  !! the whole grid (ghosts across a cut belong to the neighbour's vertices)
  subroutine gather_piece(a, b, c)
    integer, intent(in) :: a, b, c
    real(dp), allocatable, target :: w(:,:,:,:)
    integer :: n(3), lo(3), hi(3), off(3), kg
    n = [int(px(a) - px(a-1)) + 1, int(py(b) - py(b-1)) + 1, int(pz(c) - pz(c-1)) + 1]
    kg = NG
    if (nk == 1) kg = 0
    allocate(w(n(1) + 2*NG, n(2) + 2*NG, n(3) + 2*kg, 5))
    call fvm_check(fvm_download(pieces(a, b, c), c_loc(w), int(NG, c_int)), "download")
    off = [int(px(a-1)) - 1, int(py(b-1)) - 1, int(pz(c-1)) - 1]
    lo = [NG + 1, NG + 1, kg + 1]
    hi = [NG + n(1), NG + n(2), kg + n(3)]
    if (a == 1) lo(1) = 1
    if (b == 1) lo(2) = 1
    if (c == 1) lo(3) = 1
    if (a == ncut(1)) hi(1) = n(1) + 2*NG
    if (b == ncut(2)) hi(2) = n(2) + 2*NG
    if (c == ncut(3)) hi(3) = n(3) + 2*kg
    q(lo(1) + off(1):hi(1) + off(1), lo(2) + off(2):hi(2) + off(2), &
      lo(3) + off(3):hi(3) + off(3), :) = w(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), :)
  end subroutine gather_piece

end program fvm_driver
