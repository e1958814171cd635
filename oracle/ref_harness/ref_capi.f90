!=============================================================================80
!
! C-callable wrappers around the reference's own per-pencil scheme routines.
!
! TEST INFRASTRUCTURE ONLY.  This module is compiled together with the
! unmodified reference scheme modules (src/scheme/*.f90 and their closure,
! compiled where they lie) into oracle/_ref/libfvmref.so.  It lets the test
! suite call the reference arithmetic on arbitrary face states, to pin the C
! restatement in oracle/fvm_oracle.c bit for bit and to write golden vectors.
!
! Primitive states cross this interface as AoS doubles (rho,u,v,w,p) per
! point, the memory image of Primitive_Variable
! (src/base/defs_fluid_props.f90:64-70).
!
!=============================================================================80
module ref_capi
  use iso_c_binding
  use kind_parameter,   only : dp, i8
  use defs_fluid_props, only : Primitive_Variable
  use numerical_scheme, only : TimeScheme, SpaceScheme
  implicit none
  private
  public :: aos_to_prim, prim_to_aos

contains

  subroutine aos_to_prim(n, a, pv)
    integer(i8), intent(in) :: n
    real(dp), intent(in) :: a(5, n)
    type(Primitive_Variable), intent(inout) :: pv(n)
    integer(i8) :: i
    do i = 1, n
      pv(i)%rho = a(1, i); pv(i)%u = a(2, i); pv(i)%v = a(3, i)
      pv(i)%w = a(4, i);   pv(i)%press = a(5, i)
    end do
  end subroutine aos_to_prim

  subroutine prim_to_aos(n, pv, a)
    integer(i8), intent(in) :: n
    type(Primitive_Variable), intent(in) :: pv(n)
    real(dp), intent(inout) :: a(5, n)
    integer(i8) :: i
    do i = 1, n
      a(1, i) = pv(i)%rho; a(2, i) = pv(i)%u; a(3, i) = pv(i)%v
      a(4, i) = pv(i)%w;   a(5, i) = pv(i)%press
    end do
  end subroutine prim_to_aos

  !> src/scheme/scheme_reconstruct_simple.f90:42-85
  subroutine ref_recon_simple(nface, dghost, npg, pg, pr, wind) &
      bind(C, name="ref_recon_simple")
    integer(c_int64_t), value :: nface, dghost, npg
    integer(c_int),     value :: wind
    real(c_double), intent(in)    :: pg(5, npg)
    real(c_double), intent(inout) :: pr(5, nface)
    type(Primitive_Variable), allocatable :: pgv(:), prv(:)
    allocate(pgv(npg), prv(nface))
    call aos_to_prim(npg, pg, pgv)
    call aos_to_prim(nface, pr, prv)
    call SpaceScheme%reconstruct_simple(nface, dghost, pgv, prv, int(wind))
    call prim_to_aos(nface, prv, pr)
  end subroutine ref_recon_simple

  !> src/scheme/scheme_reconstruct_MUSCL.f90:47-104
  subroutine ref_recon_muscl(nface, dghost, kappa, limiter, npg, pg, pr, wind) &
      bind(C, name="ref_recon_muscl")
    integer(c_int64_t), value :: nface, dghost, npg
    real(c_double),     value :: kappa
    integer(c_int),     value :: limiter, wind
    real(c_double), intent(in)    :: pg(5, npg)
    real(c_double), intent(inout) :: pr(5, nface)
    type(Primitive_Variable), allocatable :: pgv(:), prv(:)
    allocate(pgv(npg), prv(nface))
    call aos_to_prim(npg, pg, pgv)
    call aos_to_prim(nface, pr, prv)
    call SpaceScheme%reconstruct_MUSCL(nface, dghost, kappa, int(limiter), &
                                       pgv, prv, int(wind))
    call prim_to_aos(nface, prv, pr)
  end subroutine ref_recon_muscl

  !> src/scheme/scheme_WENO.f90:28-72 (conserved variables, SoA (npg, neq))
  subroutine ref_recon_weno(nface, neq, dghost, npg, qg, qr, wind) &
      bind(C, name="ref_recon_weno")
    integer(c_int64_t), value :: nface, dghost, npg
    integer(c_int),     value :: neq, wind
    real(c_double), intent(inout) :: qg(npg, neq)
    real(c_double), intent(inout) :: qr(nface, neq)
    call SpaceScheme%reconstruct_WENO(nface, int(neq), dghost, qg, qr, int(wind))
  end subroutine ref_recon_weno

  !> src/scheme/scheme_FDS_Roe.f90:25-138
  subroutine ref_flux_roe(nface, neq, gamma, norm, ql, qr, flux, entfix) &
      bind(C, name="ref_flux_roe")
    integer(c_int64_t), value :: nface
    integer(c_int),     value :: neq
    real(c_double),     value :: gamma, entfix
    real(c_double), intent(inout) :: norm(nface, 3)
    real(c_double), intent(in)    :: ql(5, nface), qr(5, nface)
    real(c_double), intent(inout) :: flux(nface, neq)
    type(Primitive_Variable), allocatable :: lv(:), rv(:)
    allocate(lv(nface), rv(nface))
    call aos_to_prim(nface, ql, lv)
    call aos_to_prim(nface, qr, rv)
    call SpaceScheme%FDS_Roe(nface, int(neq), gamma, norm, lv, rv, flux, entfix)
  end subroutine ref_flux_roe

  !> src/scheme/scheme_FVS_VanLeer.f90:25-98
  subroutine ref_flux_vanleer(nface, neq, gamma, norm, ql, qr, flux) &
      bind(C, name="ref_flux_vanleer")
    integer(c_int64_t), value :: nface
    integer(c_int),     value :: neq
    real(c_double),     value :: gamma
    real(c_double), intent(inout) :: norm(nface, 3)
    real(c_double), intent(in)    :: ql(5, nface), qr(5, nface)
    real(c_double), intent(inout) :: flux(nface, neq)
    type(Primitive_Variable), allocatable :: lv(:), rv(:)
    allocate(lv(nface), rv(nface))
    call aos_to_prim(nface, ql, lv)
    call aos_to_prim(nface, qr, rv)
    call SpaceScheme%FVS_VanLeer(nface, int(neq), gamma, norm, lv, rv, flux)
  end subroutine ref_flux_vanleer

  !> src/scheme/scheme_AUSM.f90:27-112
  subroutine ref_flux_ausm(nface, neq, gamma, norm, ql, qr, flux) &
      bind(C, name="ref_flux_ausm")
    integer(c_int64_t), value :: nface
    integer(c_int),     value :: neq
    real(c_double),     value :: gamma
    real(c_double), intent(inout) :: norm(nface, 3)
    real(c_double), intent(in)    :: ql(5, nface), qr(5, nface)
    real(c_double), intent(inout) :: flux(nface, neq)
    type(Primitive_Variable), allocatable :: lv(:), rv(:)
    allocate(lv(nface), rv(nface))
    call aos_to_prim(nface, ql, lv)
    call aos_to_prim(nface, qr, rv)
    call SpaceScheme%FVS_AUSM(nface, int(neq), gamma, norm, lv, rv, flux)
  end subroutine ref_flux_ausm

  !> src/scheme/scheme_Runge_Kutta.f90:30-63 on flat arrays of n values
  subroutine ref_runge_kutta(m, dt, n, u0, um, rm) bind(C, name="ref_runge_kutta")
    integer(c_int),     value :: m
    real(c_double),     value :: dt
    integer(c_int64_t), value :: n
    real(c_double), intent(inout) :: u0(n), um(n), rm(n)
    real(dp), pointer :: pu0(:,:,:,:), pum(:,:,:,:), prm(:,:,:,:)
    allocate(pu0(n,1,1,1), pum(n,1,1,1), prm(n,1,1,1))
    pu0(:,1,1,1) = u0; pum(:,1,1,1) = um; prm(:,1,1,1) = rm
    call TimeScheme%Runge_Kutta(int(m), dt, pu0, pum, prm)
    u0 = pu0(:,1,1,1); um = pum(:,1,1,1)
    deallocate(pu0, pum, prm)
  end subroutine ref_runge_kutta

end module ref_capi
