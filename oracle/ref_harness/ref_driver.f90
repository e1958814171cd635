!=============================================================================80
!
! Single-block explicit residual + SSP-RK3 driver around the reference's own
! scheme routines.
!
! TEST INFRASTRUCTURE ONLY.  The reference's driver modules (src/solve/*,
! src/base/defs_block_data.f90) cannot be compiled here: they need the Fortran
! stdlib, CGNS and MPI, none of which is installed.  This file restates the
! driver loop in the reference's own data layout (ghost-inclusive
! q(ni+6,nj+6,nk+6,neq) column-major, AoS primitives, per-pencil calls with
! dghost = 3, face arrays of nvert+1 faces) and calls the UNMODIFIED compiled
! scheme routines for every arithmetic kernel.  Whole-state golden fixtures
! come from this driver.
!
! Followed, step by step:
!   time step             src/solve/class_time_discret.f90:125-196
!   inviscid residual     src/solve/class_space_discret.f90:120-176
!   pencil dispatch       src/solve/class_space_discret.f90:179-668
!   flux differencing     src/solve/class_space_discret.f90:678-732
!   q <-> p conversion    src/solve/class_flow_field.f90:353-396
!   ghost extrapolation   src/base/defs_block_data.f90:226-256,
!                         regions src/base/defs_ghost_cell.f90:122-166
!   q0 / rhs reset        src/solve/class_flow_field.f90:447-474
!   residual norms        src/solve/class_iteration.f90:210-259
!
!=============================================================================80
module ref_driver
  use iso_c_binding
  use kind_parameter,   only : dp, i8
  use defs_fluid_props, only : Primitive_Variable
  use numerical_scheme, only : TimeScheme, SpaceScheme
  use global_type_defs, only : TypeDefs
  implicit none
  private

  integer, parameter :: NG = 3     !< ghost layers per side (defs_ghost_cell.f90:42)
  integer, parameter :: NEQ = 5    !< Euler keeps momentumZ in 2-D (solver_sub.f90:131-136)

contains

  !> p -> q on the real region (class_flow_field.f90:353-377)
  subroutine p_to_q(gamma, p, q)
    real(dp), intent(in) :: gamma
    type(Primitive_Variable), intent(in) :: p(:,:,:)
    real(dp), intent(inout) :: q(:,:,:,:)
    q(:,:,:,1) = p%rho
    q(:,:,:,2) = p%rho * p%u
    q(:,:,:,3) = p%rho * p%v
    q(:,:,:,4) = p%rho * p%w
    q(:,:,:,5) = p%press / (gamma - 1.0_dp) + &
                 0.5_dp * p%rho * (p%u**2 + p%v**2 + p%w**2)
  end subroutine p_to_q

  !> q -> p on the real region (class_flow_field.f90:383-396)
  subroutine q_to_p(gamma, q, p)
    real(dp), intent(in) :: gamma
    real(dp), intent(in) :: q(:,:,:,:)
    type(Primitive_Variable), intent(inout) :: p(:,:,:)
    p%rho   = q(:,:,:,1)
    p%u     = q(:,:,:,2) / p%rho
    p%v     = q(:,:,:,3) / p%rho
    p%w     = q(:,:,:,4) / p%rho
    p%press = (gamma - 1.0_dp) * (q(:,:,:,5) - &
                0.5_dp * p%rho * (p%u**2 + p%v**2 + p%w**2))
  end subroutine q_to_p

  !> q -> p along one face line (class_flow_field.f90:415-428)
  subroutine q_to_p_line(gamma, q, p)
    real(dp), intent(in) :: gamma
    real(dp), intent(in) :: q(:,:)
    type(Primitive_Variable), intent(inout) :: p(:)
    p%rho   = q(:,1)
    p%u     = q(:,2) / p%rho
    p%v     = q(:,3) / p%rho
    p%w     = q(:,4) / p%rho
    p%press = (gamma - 1.0_dp) * (q(:,5) - &
                0.5_dp * p%rho * (p%u**2 + p%v**2 + p%w**2))
  end subroutine q_to_p_line

  !> Zeroth-order ghost fill of both p and q: 6 face regions x NG layers,
  !! each ghost vertex copies the boundary vertex; edges and corners untouched.
  subroutine fill_ghosts(nv, q, p)
    integer(i8), intent(in) :: nv(3)
    real(dp), intent(inout) :: q(:,:,:,:)
    type(Primitive_Variable), intent(inout) :: p(:,:,:)
    integer(i8) :: i, j, k, lo(3), hi(3), src(3)
    integer :: region, layer, d, side
    do region = 1, 6
      d = (region + 1) / 2
      side = merge(-1, 1, mod(region, 2) == 1)
      do layer = 1, NG
        lo = NG + 1
        hi = NG + nv
        if (side < 0) then
          lo(d) = NG + 1 - layer
        else
          lo(d) = NG + nv(d) + layer
        end if
        hi(d) = lo(d)
        do k = lo(3), hi(3)
          do j = lo(2), hi(2)
            do i = lo(1), hi(1)
              src = (/ i, j, k /)
              src(d) = src(d) - side * layer
              p(i, j, k) = p(src(1), src(2), src(3))
              q(i, j, k, :) = q(src(1), src(2), src(3), :)
            end do
          end do
        end do
      end do
    end do
  end subroutine fill_ghosts

  !> Inviscid residual of one block, dir = 1..idim, accumulated into rhs
  subroutine inviscid_residual(nv, idim, dr, recon, limiter, fluxc, kappa, &
                               entfix, gamma, q, p, rhs, flux, pm, pp, qrec)
    integer(i8), intent(in) :: nv(3)
    integer,     intent(in) :: idim, recon, limiter, fluxc
    real(dp),    intent(in) :: dr(3), kappa, entfix, gamma
    real(dp),    intent(inout) :: q(:,:,:,:)
    type(Primitive_Variable), intent(inout) :: p(:,:,:)
    real(dp),    intent(inout) :: rhs(:,:,:,:), flux(:,:,:,:), qrec(:,:,:,:)
    type(Primitive_Variable), intent(inout) :: pm(:,:,:), pp(:,:,:)
    real(dp), allocatable :: norm(:,:)
    integer(i8) :: i, j, k, nface, a, b, dg
    integer :: dir
    dg = NG
    do dir = 1, idim
      ! reset_pool (defs_block_data.f90:204-211)
      qrec = 0.0_dp
      flux = 0.0_dp
      call zero_prim(pm)
      call zero_prim(pp)
      nface = nv(dir) + 1
      allocate(norm(nface, 3))
      norm = 0.0_dp
      norm(:, dir) = 1.0_dp
      ! one pencil per (a, b) in the plane normal to dir
      do b = 1, nv(merge(3, merge(3, 2, dir == 1), dir /= 3))
        do a = 1, nv(merge(2, 1, dir == 1))
          select case (dir)
          case (1)
            j = a; k = b
            call pencil(p(:, j + dg, k + dg), q(:, j + dg, k + dg, :), &
                        pm(1:nface, j, k), pp(1:nface, j, k), &
                        qrec(1:nface, j, k, :), flux(1:nface, j, k, :))
          case (2)
            i = a; k = b
            call pencil(p(i + dg, :, k + dg), q(i + dg, :, k + dg, :), &
                        pm(i, 1:nface, k), pp(i, 1:nface, k), &
                        qrec(i, 1:nface, k, :), flux(i, 1:nface, k, :))
          case (3)
            i = a; j = b
            call pencil(p(i + dg, j + dg, :), q(i + dg, j + dg, :, :), &
                        pm(i, j, 1:nface), pp(i, j, 1:nface), &
                        qrec(i, j, 1:nface, :), flux(i, j, 1:nface, :))
          end select
        end do
      end do
      deallocate(norm)
      ! update_residual_sub: rhs -= (F(f+1) - F(f)) / dr
      do k = 1, nv(3)
        do j = 1, nv(2)
          do i = 1, nv(1)
            select case (dir)
            case (1)
              rhs(i,j,k,:) = rhs(i,j,k,:) - (flux(i+1,j,k,:) - flux(i,j,k,:)) / dr(dir)
            case (2)
              rhs(i,j,k,:) = rhs(i,j,k,:) - (flux(i,j+1,k,:) - flux(i,j,k,:)) / dr(dir)
            case (3)
              rhs(i,j,k,:) = rhs(i,j,k,:) - (flux(i,j,k+1,:) - flux(i,j,k,:)) / dr(dir)
            end select
          end do
        end do
      end do
    end do
  contains
    subroutine pencil(pg, qg, pminus, pplus, qr, fl)
      type(Primitive_Variable), intent(inout) :: pg(:)
      real(dp), intent(inout) :: qg(:,:)
      type(Primitive_Variable), intent(inout) :: pminus(:), pplus(:)
      real(dp), intent(inout) :: qr(:,:), fl(:,:)
      type(Primitive_Variable), allocatable :: lm(:), lp(:)
      real(dp), allocatable :: fx(:,:)
      allocate(lm(nface), lp(nface), fx(nface, NEQ))
      lm = pminus
      lp = pplus
      if (recon == TypeDefs%Scheme%simple) then
        call SpaceScheme%reconstruct_simple(nface, dg, pg, lm, wind=-1)
        call SpaceScheme%reconstruct_simple(nface, dg, pg, lp, wind=1)
      else if (recon == TypeDefs%Scheme%MUSCL) then
        call SpaceScheme%reconstruct_MUSCL(nface, dg, kappa, limiter, pg, lm, wind=-1)
        call SpaceScheme%reconstruct_MUSCL(nface, dg, kappa, limiter, pg, lp, wind=1)
      else if (recon == TypeDefs%Scheme%WENO) then
        call SpaceScheme%reconstruct_WENO(nface, NEQ, dg, qg, qr, wind=-1)
        call q_to_p_line(gamma, qr, lm)
        call SpaceScheme%reconstruct_WENO(nface, NEQ, dg, qg, qr, wind=1)
        call q_to_p_line(gamma, qr, lp)
      end if
      pminus = lm
      pplus = lp
      fx = fl
      if (fluxc == TypeDefs%Scheme%Roe) then
        call SpaceScheme%FDS_Roe(nface, NEQ, gamma, norm, lm, lp, fx, entfix)
      else if (fluxc == TypeDefs%Scheme%VanLeer) then
        call SpaceScheme%FVS_VanLeer(nface, NEQ, gamma, norm, lm, lp, fx)
      else if (fluxc == TypeDefs%Scheme%AUSM) then
        call SpaceScheme%FVS_AUSM(nface, NEQ, gamma, norm, lm, lp, fx)
      end if
      fl = fx
    end subroutine pencil
  end subroutine inviscid_residual

  subroutine zero_prim(pv)
    type(Primitive_Variable), intent(inout) :: pv(:,:,:)
    pv%rho = 0.0_dp; pv%u = 0.0_dp; pv%v = 0.0_dp; pv%w = 0.0_dp; pv%press = 0.0_dp
  end subroutine zero_prim

  !> Run nsteps fixed-dt SSP-RK3 steps on one block.
  !!  p_real(5, ni, nj, nk)            initial primitives on real vertices (AoS)
  !!  p_from_q = 1: first stage reconstructs from convert(q) instead of the
  !!                exact initial primitives (class_flow_field.f90:244-253)
  !!  q_out(ni+6, nj+6, nk+6, 5)       final conserved state with ghosts
  !!  rhs_out(ni, nj, nk, 5)           rhs of the last stage
  !!  res_out(6, nsteps)               per-step absolute residuals
  subroutine ref_solve(ni, nj, nk, idim, dx, dy, dz, recon, limiter, fluxc, &
                       kappa, entfix, gamma, dt, nsteps, p_from_q, p_real, &
                       q_out, rhs_out, res_out) bind(C, name="ref_solve")
    integer(c_int64_t), value :: ni, nj, nk
    integer(c_int),     value :: idim, recon, limiter, fluxc, nsteps, p_from_q
    real(c_double),     value :: dx, dy, dz, kappa, entfix, gamma, dt
    real(c_double), intent(in)    :: p_real(5, ni, nj, nk)
    real(c_double), intent(inout) :: q_out(ni + 2*NG, nj + 2*NG, nk + 2*NG, NEQ)
    real(c_double), intent(inout) :: rhs_out(ni, nj, nk, NEQ)
    real(c_double), intent(inout) :: res_out(NEQ + 1, max(nsteps, 1))
    real(dp), pointer :: q(:,:,:,:), q0(:,:,:,:), rhs(:,:,:,:), qreal(:,:,:,:)
    real(dp), allocatable :: flux(:,:,:,:), qrec(:,:,:,:)
    type(Primitive_Variable), allocatable :: p(:,:,:), pm(:,:,:), pp(:,:,:)
    integer(i8) :: nv(3), i, j, k
    integer :: step, irk, ieq
    real(dp) :: dr(3), res2(NEQ), res2t, resabs(NEQ + 1)

    nv = (/ ni, nj, nk /)
    dr = (/ dx, dy, dz /)
    allocate(q(ni + 2*NG, nj + 2*NG, nk + 2*NG, NEQ))
    allocate(p(ni + 2*NG, nj + 2*NG, nk + 2*NG))
    allocate(q0(ni, nj, nk, NEQ), rhs(ni, nj, nk, NEQ))
    allocate(flux(ni + 1, nj + 1, nk + 1, NEQ), qrec(ni + 1, nj + 1, nk + 1, NEQ))
    allocate(pm(ni + 1, nj + 1, nk + 1), pp(ni + 1, nj + 1, nk + 1))
    q = 0.0_dp
    qreal => q(NG + 1:NG + ni, NG + 1:NG + nj, NG + 1:NG + nk, :)

    ! initial state (class_flow_field.f90:195-256)
    do k = 1, nk
      do j = 1, nj
        do i = 1, ni
          p(i + NG, j + NG, k + NG)%rho   = p_real(1, i, j, k)
          p(i + NG, j + NG, k + NG)%u     = p_real(2, i, j, k)
          p(i + NG, j + NG, k + NG)%v     = p_real(3, i, j, k)
          p(i + NG, j + NG, k + NG)%w     = p_real(4, i, j, k)
          p(i + NG, j + NG, k + NG)%press = p_real(5, i, j, k)
        end do
      end do
    end do
    call p_to_q(gamma, p(NG + 1:NG + ni, NG + 1:NG + nj, NG + 1:NG + nk), qreal)
    if (p_from_q /= 0) then
      call q_to_p(gamma, qreal, p(NG + 1:NG + ni, NG + 1:NG + nj, NG + 1:NG + nk))
    end if
    call fill_ghosts(nv, q, p)

    do step = 1, nsteps
      q0 = qreal                                   ! reset_q0
      do irk = 1, 3
        rhs = 0.0_dp                               ! reset_rhs
        call inviscid_residual(nv, int(idim), dr, int(recon), int(limiter), &
                               int(fluxc), kappa, entfix, gamma, q, p, rhs, &
                               flux, pm, pp, qrec)
        call TimeScheme%Runge_Kutta(irk, dt, q0, qreal, rhs)
        call q_to_p(gamma, qreal, p(NG + 1:NG + ni, NG + 1:NG + nj, NG + 1:NG + nk))
        call fill_ghosts(nv, q, p)
      end do
      ! compute_residual_sub on a single block, starting from zero
      res2 = 0.0_dp
      res2t = 0.0_dp
      do ieq = 1, NEQ
        do k = 1, nk
          do j = 1, nj
            do i = 1, ni
              res2(ieq) = res2(ieq) + rhs(i, j, k, ieq)**2
            end do
          end do
        end do
        res2(ieq) = res2(ieq) / real(ni * nj * nk, dp)
        res2t = res2t + res2(ieq)
      end do
      res2t = res2t / real(NEQ, dp)
      resabs = 0.0_dp
      do ieq = 1, NEQ
        resabs(ieq) = sqrt(resabs(ieq)**2 + res2(ieq))
      end do
      resabs(NEQ + 1) = sqrt(resabs(NEQ + 1)**2 + res2t)
      res_out(:, step) = resabs
    end do

    q_out = q
    rhs_out = rhs
    deallocate(q, p, q0, rhs, flux, qrec, pm, pp)
  end subroutine ref_solve

end module ref_driver
