! prtcl_variables.f90 -- host copies of the particle state in the reference's
! array shapes (src/DEM/Prtcl_Variables.f90:26-36): GPrtcl_PosR(4,n),
! GPrtcl_linVel(3,tsize,n), ...  GL_MakingAllPrtcl places the particles on
! the lattice of the &ParticleMakingOption namelist
! (Prtcl_Variables.f90:191-307); its random shuffles and velocity directions
! come from a seeded generator instead of the clock.
module Prtcl_Variables
  use, intrinsic :: iso_c_binding
  use Prtcl_Parameters
  use Prtcl_Property
  implicit none
  private

  integer, public, save :: nlocal = 0, tsize = 1, rsize = 1
  integer(c_int32_t), allocatable, public, save :: GPrtcl_id(:), GPrtcl_pType(:), GPrtcl_UsrMark(:)
  real(RK), allocatable, public, save :: GPrtcl_PosR(:, :), GPrtcl_linVel(:, :, :), GPrtcl_linAcc(:, :, :)
  real(RK), allocatable, public, save :: GPrtcl_theta(:, :), GPrtcl_rotVel(:, :, :), GPrtcl_rotAcc(:, :, :)

  public :: GL_AllocateAllVar, GL_MakingAllPrtcl

contains

  integer function level_size(m)
    integer, intent(in) :: m
    level_size = merge(1, merge(2, 3, m == PIM_AB2), m == PIM_FE)
  end function level_size

  subroutine GL_AllocateAllVar(n)
    integer, intent(in) :: n
    tsize = level_size(DEM_opt%PI_Method)
    rsize = level_size(DEM_opt%PRI_Method)
    if (allocated(GPrtcl_id)) deallocate(GPrtcl_id, GPrtcl_pType, GPrtcl_UsrMark, GPrtcl_PosR, GPrtcl_linVel, &
                                        GPrtcl_linAcc, GPrtcl_theta, GPrtcl_rotVel, GPrtcl_rotAcc)
    allocate(GPrtcl_id(n), GPrtcl_pType(n), GPrtcl_UsrMark(n), GPrtcl_PosR(4, n))
    allocate(GPrtcl_linVel(3, tsize, n), GPrtcl_linAcc(3, tsize, n), GPrtcl_theta(3, n))
    allocate(GPrtcl_rotVel(3, rsize, n), GPrtcl_rotAcc(3, rsize, n))
    GPrtcl_UsrMark = 1
    GPrtcl_linVel = 0.0_RK; GPrtcl_linAcc = 0.0_RK; GPrtcl_theta = 0.0_RK
    GPrtcl_rotVel = 0.0_RK; GPrtcl_rotAcc = 0.0_RK
  end subroutine GL_AllocateAllVar

  ! unit vectors of the fill order ("x","z","y": fill x first, then z, then y)
  subroutine Fill_Vectors(order, l1, l2, l3)
    character, intent(in) :: order(3)
    real(RK), intent(out) :: l1(3), l2(3), l3(3)
    call axis(order(1), l1); call axis(order(2), l2); call axis(order(3), l3)
  contains
    subroutine axis(c, v)
      character, intent(in) :: c
      real(RK), intent(out) :: v(3)
      v = 0.0_RK
      select case (c)
      case ('x', 'X'); v(1) = 1.0_RK
      case ('y', 'Y'); v(2) = 1.0_RK
      case ('z', 'Z'); v(3) = 1.0_RK
      case default
        print *, "MakingAllPrtcl: bad Fill_order"; stop 1
      end select
    end subroutine axis
  end subroutine Fill_Vectors

  subroutine GL_MakingAllPrtcl(chFile)
    character(*), intent(in) :: chFile
    real(RK), dimension(3) :: MkPrtclMinpoint, MkPrtclMaxpoint, l1, l2, l3, lmin_p, lmax_p, dx, cntr, vel_dir, p1
    real(RK) :: Distance_Ratio, VelMag, maxRad, randt, rvelt
    character, dimension(3) :: Fill_order
    NAMELIST /ParticleMakingOption/ MkPrtclMinpoint, MkPrtclMaxpoint, Fill_order, Distance_Ratio, VelMag
    integer :: i, j, k, n, nUnitFile, myistat, bin_id
    integer, allocatable :: sum_bin(:)

    open(newunit=nUnitFile, file=chFile, status='old', form='formatted', iostat=myistat)
    if (myistat /= 0) then
      print *, "MakingAllPrtcl: cannot open file: "//trim(chFile); stop 1
    end if
    read(nUnitFile, nml=ParticleMakingOption)
    close(nUnitFile)
    n = DEM_opt%numPrtcl
    call GL_AllocateAllVar(n)
    call Fill_Vectors(Fill_order, l1, l2, l3)
    maxRad = 0.5_RK * maxval(DEMProperty%Diameter)
    lmin_p = MkPrtclMinpoint + Distance_Ratio * maxRad
    lmax_p = MkPrtclMaxpoint - Distance_Ratio * maxRad
    dx = Distance_Ratio * 2.0_RK * maxRad
    allocate(sum_bin(DEMProperty%nPType))
    sum_bin(1) = DEMProperty%nPrtcl_in_Bin(1)
    do j = 2, DEMProperty%nPType
      sum_bin(j) = sum_bin(j - 1) + DEMProperty%nPrtcl_in_Bin(j)
    end do
    cntr = lmin_p
    do i = 1, n
      bin_id = DEMProperty%nPType
      do j = 1, DEMProperty%nPType
        if (i <= sum_bin(j)) then
          bin_id = j; exit
        end if
      end do
      GPrtcl_id(i) = i
      GPrtcl_pType(i) = bin_id
      GPrtcl_PosR(1:3, i) = cntr
      GPrtcl_PosR(4, i) = 0.5_RK * DEMProperty%Diameter(bin_id)
      cntr = cntr + l1 * dx
      if (dot_product(l1, cntr) >= dot_product(l1, lmax_p)) then
        cntr = lmin_p * l1 + (cntr + dx) * l2 + cntr * l3
        if (dot_product(l2, cntr) >= dot_product(l2, lmax_p)) then
          cntr = cntr * l1 + lmin_p * l2 + (cntr + dx) * l3
          if (dot_product(l3, cntr) >= dot_product(l3, lmax_p) .and. i < n) then
            print *, "MakingAllPrtcl: not enough space for positioning"; stop 1
          end if
        end if
      end if
    end do
    nlocal = n
    ! shuffle the positions (three passes) so that sizes mix
    do k = 1, 3
      do i = 1, n
        call random_number(randt)
        j = int(randt * n) + 1
        p1 = GPrtcl_PosR(1:3, i)
        GPrtcl_PosR(1:3, i) = GPrtcl_PosR(1:3, j)
        GPrtcl_PosR(1:3, j) = p1
      end do
    end do
    do i = 1, n
      call random_number(vel_dir)
      vel_dir = 2.0_RK * vel_dir - 1.0_RK
      rvelt = sqrt(vel_dir(1)**2 + vel_dir(2)**2 + vel_dir(3)**2)
      if (rvelt > 1.0E-10_RK) then
        vel_dir = vel_dir / rvelt
      else
        vel_dir = -l3
      end if
      GPrtcl_linVel(:, 1, i) = VelMag * vel_dir
    end do
  end subroutine GL_MakingAllPrtcl

end module Prtcl_Variables
