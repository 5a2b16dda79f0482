! prtcl_geometry.f90 -- the plane walls of G_MakeGeometry
! (src/DEM/Prtcl_Geometry.f90:542-637): GeometrySource = 0, the geometry
! built into the program (an hourglass, :557-615); GeometrySource = 1, the
! &GeometryMakingNumPlane and &GeometryMakingParam namelists, whose Point1..
! Point4 and TraVel are read as 3 x nplane arrays, the storage of the
! reference's real3 arrays.  STL input (GeometrySource = 2) is not built.
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
    if (DEM_opt%GeometrySource == 0) then
      call hourglass(this)
      return
    end if
    if (DEM_opt%GeometrySource /= 1) then
      print *, "MakeGeometry: GeometrySource = 2 (walls from an STL file) is not supported"
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

  ! The reference program's own geometry (GeometrySource = 0,
  ! Prtcl_Geometry.f90:557-615): an hourglass along y, square in x and z, of
  ! six levels -- half widths 0.075 m at y = 0.30 and 0.14, 0.015 m (the
  ! neck) at y = 0.05 and -0.05, 0.075 m at y = -0.14 and -0.30 -- walled by
  ! 20 side quads, the sides z = -h, x = -h, z = +h, x = +h in that order,
  ! each from the top level down, then the top and the bottom quad; all
  ! one-sided, finite, wall type 1, user id 1, normals inwards.  Corner c of
  ! level l is (-h, y, -h), (h, y, -h), (h, y, h), (-h, y, h) for c = 1..4.
  subroutine hourglass(this)
    class(Geometry) :: this
    real(RK), parameter :: ylev(6) = [0.30_RK, 0.14_RK, 0.05_RK, -0.05_RK, -0.14_RK, -0.30_RK]
    real(RK), parameter :: hw(6) = [0.075_RK, 0.075_RK, 0.015_RK, 0.015_RK, 0.075_RK, 0.075_RK]
    real(RK), parameter :: sx(4) = [-1.0_RK, 1.0_RK, 1.0_RK, -1.0_RK], sz(4) = [-1.0_RK, -1.0_RK, 1.0_RK, 1.0_RK]
    ! per side: the corners (level offset 0 or 1, corner index) of its quads
    integer, parameter :: side(2, 4, 4) = reshape([ &
      0, 1,  1, 1,  1, 2,  0, 2, &   ! z = -h
      0, 1,  0, 4,  1, 4,  1, 1, &   ! x = -h
      0, 3,  1, 3,  1, 4,  0, 4, &   ! z = +h
      0, 2,  1, 2,  1, 3,  0, 3], [2, 4, 4])   ! x = +h
    integer :: f, l, q, k
    this%nPW = 22
    allocate(this%Point1(3, 22), this%Point2(3, 22), this%Point3(3, 22), this%Point4(3, 22), this%TraVel(3, 22))
    allocate(this%user_id(22), this%wall_Type(22), this%bothSide(22), this%IsInfinite(22))
    this%TraVel = 0.0_RK; this%user_id = 1; this%wall_Type = 1
    this%bothSide = .false.; this%IsInfinite = .false.
    k = 0
    do f = 1, 4
      do l = 1, 5
        k = k + 1
        do q = 1, 4
          call put(k, q, l + side(1, q, f), side(2, q, f))
        end do
      end do
    end do
    ! top (p01 p02 p03 p04) and bottom (p21 p24 p23 p22)
    k = 21
    do q = 1, 4
      call put(k, q, 1, q)
    end do
    k = 22
    call put(k, 1, 6, 1); call put(k, 2, 6, 4); call put(k, 3, 6, 3); call put(k, 4, 6, 2)
  contains
    subroutine put(iw, iq, lvl, c)
      integer, intent(in) :: iw, iq, lvl, c
      real(RK) :: p(3)
      p = [sx(c) * hw(lvl), ylev(lvl), sz(c) * hw(lvl)]
      select case (iq)
      case (1); this%Point1(:, iw) = p
      case (2); this%Point2(:, iw) = p
      case (3); this%Point3(:, iw) = p
      case default; this%Point4(:, iw) = p
      end select
    end subroutine put
  end subroutine hourglass

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
