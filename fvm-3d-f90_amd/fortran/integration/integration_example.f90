!=============================================================================80
!
!> The integration snippet (snippet_*.inc, quoted verbatim in the
!> integration notes) run on the reference's 2-D case: a uniform square of
!> n x n vertices and two k-planes (src/base/method_gridgen.f90:98-131),
!> ClassBlockData%q allocated as the reference allocates it, the case
!> file's scheme values read through ClassSpaceDiscret%input.
!>
!> usage: integration_example n nsteps dt q_in q_out
!>   q_in / q_out: the conserved state in ClassBlockData%q's layout,
!>   q(n+6, n+6, 8, 5), raw float64 (the state after flowfield%initial in;
!>   the state after nsteps steps out)
!
!=============================================================================80
program integration_example

  use iso_c_binding
  use fvm_gpu
  use ref_standin
  implicit none

  include 'snippet_decl.inc'
  type(ClassSpaceDiscret) :: spaceDiscr
  type(ClassFlowField)    :: flowfield
  type(ClassGeometry)     :: geometry
  type(ClassGeomBlock), pointer :: gb
  real(dp), pointer :: q(:,:,:,:)
  character(len=512) :: arg, fin, fout
  integer(i8) :: n
  integer :: nsteps, step, u
  real(dp) :: dt

  call get_command_argument(1, arg); read(arg, *) n
  call get_command_argument(2, arg); read(arg, *) nsteps
  call get_command_argument(3, arg); read(arg, *) dt
  call get_command_argument(4, fin)
  call get_command_argument(5, fout)

!...case file &Inviscid_Flux of case/Riemann-2D.nml
  call spaceDiscr%input("reconstruct", "MUSCL")
  call spaceDiscr%input("kappa", "0.33333")
  call spaceDiscr%input("limiter", "minmod_")
  call spaceDiscr%input("flux_split", "Roe")
  call spaceDiscr%input("entropy_fix", "0.2")
!...geometry: "2D" square, unit size, gsize(3) = 2
  geometry%idimension = 2
  call geometry%add_blocks(1)
  gb => geometry%block(1)
  call gb%set([n, n, 2_i8], [1.0_dp / real(n - 1, dp), 1.0_dp / real(n - 1, dp), 1.0_dp])
!...flow field: one block, q on the ghost mesh, state after flowfield%initial
  allocate(flowfield%block(1))
  call flowfield%block(1)%allocate_memory([n, n, 2_i8], 5)
  q => flowfield%block(1)%get_q_ghost()
  open(newunit=u, file=trim(fin), access="stream", form="unformatted", status="old")
  read(u) q
  close(u)

  include 'snippet_setup.inc'
  do step = 1, nsteps
    include 'snippet_step.inc'
  end do
  include 'snippet_output.inc'

  open(newunit=u, file=trim(fout), access="stream", form="unformatted", status="replace")
  write(u) qfull
  close(u)
  write(*, '(a, i0, a, 6es24.16)') "integration example: ", nsteps, " steps, norms", res
  call fvm_check(fvm_ctx_destroy(ctx), "ctx_destroy")

end program integration_example
