!=============================================================================80
!
!> Serving example: independent four-quadrant Riemann solves (the reference's
!> quadrant states, src/solve/class_flow_field.f90:207-243) of three grid
!> sizes, initial states computed on the device, run to a common end time
!> by the serving loop; each final state is written to <prefix>_<n>.bin as
!> q(n, n, 1, 5) in Fortran order.
!
!  Usage:  fvm_serve <prefix>
!
!=============================================================================80
program fvm_serve

  use iso_c_binding
  use fvm_gpu
  implicit none

  integer, parameter :: dp = c_double, NREQ = 3
  integer, parameter :: sizes(NREQ) = [41, 64, 101]
  real(dp) :: regions(11, 4), info(8)
  real(dp), allocatable, target :: q(:,:,:,:)
  type(fvm_request) :: req
  type(c_ptr) :: srv
  integer(c_int64_t) :: ids(NREQ)
  integer(c_int) :: n_pending, n_active, n_done
  integer :: r, n, unit
  character(len=512) :: prefix, fname

  call get_command_argument(1, prefix)
  if (len_trim(prefix) == 0) then
    write(*, '(a)') "usage: fvm_serve <prefix>"
    error stop 2
  end if

!...The reference's regions, in order, inclusive bounds
  regions(:, 1) = [0.8_dp, 1.0_dp, 0.8_dp, 1.0_dp, -huge(1.0_dp), huge(1.0_dp), &
                   1.5_dp, 0.0_dp, 0.0_dp, 0.0_dp, 1.5_dp]
  regions(:, 2) = [0.0_dp, 0.8_dp, 0.8_dp, 1.0_dp, -huge(1.0_dp), huge(1.0_dp), &
                   0.5323_dp, 1.206_dp, 0.0_dp, 0.0_dp, 0.3_dp]
  regions(:, 3) = [0.0_dp, 0.8_dp, 0.0_dp, 0.8_dp, -huge(1.0_dp), huge(1.0_dp), &
                   0.138_dp, 1.206_dp, 1.206_dp, 0.0_dp, 0.029_dp]
  regions(:, 4) = [0.8_dp, 1.0_dp, 0.0_dp, 0.8_dp, -huge(1.0_dp), huge(1.0_dp), &
                   0.5323_dp, 0.0_dp, 1.206_dp, 0.0_dp, 0.3_dp]

  call fvm_check(fvm_server_create(0_c_int, 0.0_c_double, 16_c_int, srv), "server_create")
  do r = 1, NREQ
    n = sizes(r)
    req = fvm_request()
    req%n = [int(n, c_int64_t), int(n, c_int64_t), 1_c_int64_t]
    req%spacing = [1.0_dp / real(n - 1, dp), 1.0_dp / real(n - 1, dp), 1.0_dp]
    req%recon = FVM_MUSCL
    req%limiter = FVM_LIM_MINMOD
    req%flux = FVM_ROE
    req%rk = FVM_RK_SSP3
    req%kappa = 0.33333_dp
    req%entropy_fix = 0.2_dp
    req%gamma = 1.4_dp
    req%ref_compat = 1
    req%numerics = FVM_NUMERICS_EXACT
    req%t_end = 0.01_dp
    req%cfl = 0.5_dp
    req%ic_kind = FVM_IC_REGIONS
    req%ic_nparams = 44
    req%ic_params(1:44) = reshape(regions, [44])
    call fvm_check(fvm_server_submit(srv, req, ids(r)), "submit")
  end do
  n_done = 0
  do while (n_done < NREQ)
    call fvm_check(fvm_server_poll(srv, 16_c_int, 1_c_int, n_pending, n_active, n_done), "poll")
  end do
  do r = 1, NREQ
    n = sizes(r)
    allocate(q(n, n, 1, 5))
    call fvm_check(fvm_server_result(srv, ids(r), c_loc(q), info), "result")
    if (nint(info(7)) /= 2) then
      write(*, '(a, i0)') "request not done: status ", nint(info(7))
      error stop 3
    end if
    write(fname, '(a, "_", i0, ".bin")') trim(prefix), n
    open(newunit=unit, file=trim(fname), access="stream", form="unformatted", &
         status="replace", action="write")
    write(unit) q
    close(unit)
    write(*, '(a, i0, a, i0, a, es14.6)') "served ", n, "^2: ", nint(info(4)), " steps to t = ", info(6)
    deallocate(q)
  end do
  call fvm_check(fvm_server_destroy(srv), "server_destroy")

end program fvm_serve
