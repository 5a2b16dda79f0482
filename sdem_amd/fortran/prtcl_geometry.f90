! prtcl_geometry.f90 -- plane walls from the &GeometryMakingNumPlane and
! &GeometryMakingParam namelists (G_MakeGeometry with GeometrySource = 1,
! src/DEM/Prtcl_Geometry.f90:542-637).  Point1..Point4 and TraVel are read as
! 3 x nplane arrays, the storage of the reference's real3 arrays.
module Prtcl_Geometry
  use, intrinsic :: iso_c_binding
  use sdem_capi
  use Prtcl_Parameters
  implicit none
  private

  type, public :: Geometry
    integer :: nPW = 0
    real(RK), allocatable :: Point1(:, :), Point2(:, :), Point3(:, :), Point4(:, :), TraVel(:, :)
    integer, allocatable :: user_id(:), wall_Type(:)
    logical, allocatable :: bothSide(:), IsInfinite(:)
  contains
    procedure :: MakeGeometry => G_MakeGeometry
    procedure :: push => G_push
  end type Geometry

  type(Geometry), public, save :: DEMGeometry

contains

  subroutine G_MakeGeometry(this, chFile)
    class(Geometry) :: this
    character(*), intent(in) :: chFile
    integer :: nplane, nUnitFile, myistat
    integer, allocatable :: user_id(:), wall_Type(:)
    logical, allocatable :: bothSide(:), IsInfinite(:)
    real(RK), allocatable :: Point1(:, :), Point2(:, :), Point3(:, :), Point4(:, :), TraVel(:, :)
    NAMELIST /GeometryMakingNumPlane/ nplane
    NAMELIST /GeometryMakingParam/ Point1, Point2, Point3, Point4, TraVel, user_id, wall_Type, bothSide, IsInfinite

    this%nPW = 0
    if (DEM_opt%GeometrySource /= 1) then
      print *, "MakeGeometry: only GeometrySource = 1 (walls from the .prm file) is supported"
      stop 1
    end if
    open(newunit=nUnitFile, file=chFile, status='old', form='formatted', iostat=myistat)
    if (myistat /= 0) then
      print *, "MakeGeometry: cannot open file: "//trim(chFile); stop 1
    end if
    nplane = 0
    read(nUnitFile, nml=GeometryMakingNumPlane)
    if (nplane < 1) then
      close(nUnitFile); return
    end if
    allocate(user_id(nplane), wall_Type(nplane), bothSide(nplane), IsInfinite(nplane))
    allocate(Point1(3, nplane), Point2(3, nplane), Point3(3, nplane), Point4(3, nplane), TraVel(3, nplane))
    TraVel = 0.0_RK; bothSide = .false.; IsInfinite = .false.; user_id = 1; wall_Type = 1
    rewind(nUnitFile)
    read(nUnitFile, nml=GeometryMakingParam)
    close(nUnitFile)
    this%nPW = nplane
    call move_alloc(Point1, this%Point1); call move_alloc(Point2, this%Point2)
    call move_alloc(Point3, this%Point3); call move_alloc(Point4, this%Point4)
    call move_alloc(TraVel, this%TraVel); call move_alloc(user_id, this%user_id)
    call move_alloc(wall_Type, this%wall_Type); call move_alloc(bothSide, this%bothSide)
    call move_alloc(IsInfinite, this%IsInfinite)
  end subroutine G_MakeGeometry

  ! add_PlaneWall for every plane, in file order (wall ids Base_wall_id+1, +2, ...)
  integer function G_push(this, ctx) result(rc)
    class(Geometry), intent(in) :: this
    type(c_ptr), intent(in) :: ctx
    integer :: i
    integer(c_int32_t) :: wid
    rc = SDEM_OK
    do i = 1, this%nPW
      rc = sdem_add_plane_wall(ctx, this%Point1(:, i), this%Point2(:, i), this%Point3(:, i), this%Point4(:, i), &
                               int(this%user_id(i), c_int32_t), int(this%wall_Type(i), c_int32_t), &
                               merge(1_c_int32_t, 0_c_int32_t, this%bothSide(i)), &
                               merge(1_c_int32_t, 0_c_int32_t, this%IsInfinite(i)), this%TraVel(:, i), wid)
      if (rc /= SDEM_OK) return
    end do
  end function G_push

end module Prtcl_Geometry
