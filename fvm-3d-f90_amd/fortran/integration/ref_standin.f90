!=============================================================================80
!
!> Stand-ins for the reference solver's types, with only the members the
!> integration snippet touches and the reference's visibility of each
!> (public scheme codes, private MUSCL / entropy-fix parameters and private
!> state array behind get_q_ghost).  They let the snippet files in this
!> folder compile and run exactly as a maintainer pastes them into
!> FVM-3D-F90:
!>   ClassSpaceDiscret   src/solve/class_space_discret.f90:25-94
!>   ClassFlowField      src/solve/class_flow_field.f90:36-45
!>   ClassBlockData      src/base/defs_block_data.f90:26-52,111-141,315-320
!>   ClassGeomBlock      src/base/defs_block_geom.f90:81-92,208-213,306-314
!>   ClassGeometry       src/solve/class_geometry.f90:88-117,266-286
!> The getters the snippet calls are the reference-side addition
!> (getters.inc), compiled here as the maintainer adds them.
!
!=============================================================================80
module ref_standin

  use iso_c_binding, only : c_double, c_int64_t
  use fvm_gpu,       only : scheme_index
  implicit none
  private

  integer, parameter, public :: dp = c_double, i8 = c_int64_t

  type, public :: ClassSpaceDiscret
    integer,  public  :: reconstruct = 0
    integer,  public  :: flux_split = 0
    real(dp), private :: kappa = -1.0_dp
    integer,  private :: limiter = -1
    real(dp), private :: entropy_fix_coeff = 0.05_dp
  contains
    procedure :: input => space_input_sub
    include 'getters_decl.inc'
  end type ClassSpaceDiscret

  type, public :: ClassBlockData
    private
    integer(i8) :: nreal(3) = 1
    real(dp), pointer :: q(:,:,:,:) => null()
  contains
    procedure :: allocate_memory => block_allocate_sub
    procedure :: get_q_ghost     => block_q_ghost_fun
  end type ClassBlockData

  type, public :: ClassFlowField
    real(dp), public :: gamma = 1.40_dp
    type(ClassBlockData), pointer, public :: block(:) => null()
  end type ClassFlowField

  type, public :: ClassGeomBlock
    private
    integer(i8) :: range_max(3) = 1
    real(dp)    :: cell_size(3) = 0.0_dp
  contains
    procedure :: set   => geom_block_set_sub
    procedure :: nvert => geom_block_nvert_fun
    procedure :: dr    => geom_block_dr_fun
  end type ClassGeomBlock

  type, public :: ClassGeometry
    integer, public :: idimension = 0
    type(ClassGeomBlock), pointer, private :: blocks(:) => null()
  contains
    procedure :: add_blocks => geometry_add_sub
    procedure :: block      => geometry_block_fun
  end type ClassGeometry

contains

  !> Case-file values, as ClassSpaceDiscret%input takes them
  subroutine space_input_sub(this, vname, vvalue)
    class(ClassSpaceDiscret), intent(inout) :: this
    character(len=*), intent(in) :: vname, vvalue
    select case (trim(vname))
      case ("reconstruct"); this%reconstruct = scheme_index(vvalue)
      case ("limiter");     this%limiter = scheme_index(vvalue)
      case ("flux_split");  this%flux_split = scheme_index(vvalue)
      case ("kappa");       read(vvalue, *) this%kappa
      case ("entropy_fix"); read(vvalue, *) this%entropy_fix_coeff
    end select
  end subroutine space_input_sub

  include 'getters.inc'

  !> q on the ghost mesh: 3 ghost layers on each side of every direction,
  !> also the third one of a 2-D grid
  subroutine block_allocate_sub(this, nreal, neq)
    class(ClassBlockData), intent(inout) :: this
    integer(i8), intent(in) :: nreal(3)
    integer, intent(in) :: neq
    this%nreal = nreal
    allocate(this%q(nreal(1) + 6, nreal(2) + 6, nreal(3) + 6, neq))
    this%q = 0.0_dp
  end subroutine block_allocate_sub

  function block_q_ghost_fun(this) result(q_ptr)
    class(ClassBlockData), intent(inout) :: this
    real(dp), pointer :: q_ptr(:,:,:,:)
    q_ptr => this%q(:,:,:,:)
  end function block_q_ghost_fun

  subroutine geom_block_set_sub(this, nvert, cell_size)
    class(ClassGeomBlock), intent(inout) :: this
    integer(i8), intent(in) :: nvert(3)
    real(dp), intent(in) :: cell_size(3)
    this%range_max = nvert
    this%cell_size = cell_size
  end subroutine geom_block_set_sub

  function geom_block_nvert_fun(this) result(nvert)
    class(ClassGeomBlock), intent(inout) :: this
    integer(i8) :: nvert(3)
    nvert = this%range_max
  end function geom_block_nvert_fun

  function geom_block_dr_fun(this, i, j, k, dir) result(dr)
    class(ClassGeomBlock), intent(inout) :: this
    integer(i8), intent(in) :: i, j, k
    integer, intent(in) :: dir
    real(dp) :: dr
    dr = this%cell_size(dir)
  end function geom_block_dr_fun

  subroutine geometry_add_sub(this, nblocks)
    class(ClassGeometry), intent(inout) :: this
    integer, intent(in) :: nblocks
    allocate(this%blocks(nblocks))
  end subroutine geometry_add_sub

  function geometry_block_fun(this, iblk) result(blk)
    class(ClassGeometry), intent(inout) :: this
    integer, intent(in) :: iblk
    type(ClassGeomBlock), pointer :: blk
    blk => this%blocks(iblk)
  end function geometry_block_fun

end module ref_standin
