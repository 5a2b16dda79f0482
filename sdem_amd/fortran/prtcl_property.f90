! prtcl_property.f90 -- the &ParticlePhysicalProperty and &WallPhysicalProperty
! namelists (InitPrtclProperty / InitWallProperty, src/DEM/Prtcl_Property.f90:62-195)
! and the split of numPrtcl into size bins.  The pair coefficients themselves
! are computed by the library (sdem_set_particle_property).  The optional
! &LinearContactModel namelist (kn, kt_over_kn) has no counterpart in the
! reference; it is read only when CF_Type = 3.
module Prtcl_Property
  use, intrinsic :: iso_c_binding
  use sdem_capi
  use Prtcl_Parameters
  implicit none
  private

  type, public :: PhysicalProperty
    integer :: nPType = 0, nWType = 0
    real(RK), allocatable :: Bin_Divided(:), Density(:), Diameter(:), YoungsModulus_P(:), PoissonRatio_P(:)
    real(RK) :: FrictionCoe_s_PP, FrictionCoe_k_PP, FrictionCoe_Roll_PP, RestitutionCoe_n_PP, RestitutionCoe_t_PP
    real(RK), allocatable :: YoungsModulus_W(:), PoissonRatio_W(:)
    real(RK) :: FrictionCoe_s_PW, FrictionCoe_k_PW, FrictionCoe_Roll_PW, RestitutionCoe_n_PW, RestitutionCoe_t_PW
    real(RK) :: kn = 0.0_RK, kt_over_kn = 0.8_RK
    integer, allocatable :: nPrtcl_in_Bin(:)
  contains
    procedure :: InitPrtclProperty
    procedure :: InitWallProperty
    procedure :: push => PP_push
  end type PhysicalProperty

  type(PhysicalProperty), public, save :: DEMProperty

contains

  subroutine InitPrtclProperty(this, chFile, seed)
    class(PhysicalProperty) :: this
    character(*), intent(in) :: chFile
    integer, intent(in) :: seed
    real(RK), allocatable, dimension(:) :: Bin_Divided, Density, Diameter, YoungsModulus_P, PoissonRatio_P
    real(RK) :: FrictionCoe_s_PP, FrictionCoe_k_PP, FrictionCoe_Roll_PP, RestitutionCoe_n_PP, RestitutionCoe_t_PP
    real(RK) :: kn, kt_over_kn, sum_divided, rtemp
    namelist /ParticlePhysicalProperty/ Bin_Divided, Density, Diameter, YoungsModulus_P, PoissonRatio_P, &
             FrictionCoe_s_PP, FrictionCoe_k_PP, FrictionCoe_Roll_PP, RestitutionCoe_n_PP, RestitutionCoe_t_PP
    namelist /LinearContactModel/ kn, kt_over_kn
    integer :: i, n, nUnitFile, myistat, sum_prtcl, prdiff, bin_id
    integer, allocatable :: put(:)

    n = DEM_opt%numPrtcl_Type
    this%nPType = n
    allocate(Bin_Divided(n), Density(n), Diameter(n), YoungsModulus_P(n), PoissonRatio_P(n))
    Bin_Divided = 1.0_RK
    open(newunit=nUnitFile, file=chFile, status='old', form='formatted', iostat=myistat)
    if (myistat /= 0) then
      print *, "InitPrtclProperty: cannot open file: "//trim(chFile); stop 1
    end if
    read(nUnitFile, nml=ParticlePhysicalProperty)
    kn = 0.0_RK; kt_over_kn = 0.8_RK
    if (DEM_opt%CF_Type == CFT_Linear) then
      rewind(nUnitFile)
      read(nUnitFile, nml=LinearContactModel)
    end if
    close(nUnitFile)
    call move_alloc(Bin_Divided, this%Bin_Divided)
    call move_alloc(Density, this%Density)
    call move_alloc(Diameter, this%Diameter)
    call move_alloc(YoungsModulus_P, this%YoungsModulus_P)
    call move_alloc(PoissonRatio_P, this%PoissonRatio_P)
    this%FrictionCoe_s_PP = FrictionCoe_s_PP; this%FrictionCoe_k_PP = FrictionCoe_k_PP
    this%FrictionCoe_Roll_PP = FrictionCoe_Roll_PP; this%RestitutionCoe_n_PP = RestitutionCoe_n_PP
    this%RestitutionCoe_t_PP = RestitutionCoe_t_PP
    this%kn = kn; this%kt_over_kn = kt_over_kn

    ! particles per bin (Prtcl_Property.f90:100-121); the remainder goes to
    ! random bins from a seeded generator instead of the clock
    call random_seed(size=i)
    allocate(put(i))
    put = seed + 63946 * [(i - 1, i = 1, size(put))]
    call random_seed(put=put)
    allocate(this%nPrtcl_in_Bin(n))
    sum_divided = sum(this%Bin_Divided)
    sum_prtcl = 0
    do i = 1, n
      this%nPrtcl_in_Bin(i) = int(DEM_opt%numPrtcl * this%Bin_Divided(i) / sum_divided)
      sum_prtcl = sum_prtcl + this%nPrtcl_in_Bin(i)
    end do
    prdiff = DEM_opt%numPrtcl - sum_prtcl
    do i = 1, prdiff
      call random_number(rtemp)
      bin_id = int(rtemp * n) + 1
      this%nPrtcl_in_Bin(bin_id) = this%nPrtcl_in_Bin(bin_id) + 1
    end do
  end subroutine InitPrtclProperty

  subroutine InitWallProperty(this, chFile)
    class(PhysicalProperty) :: this
    character(*), intent(in) :: chFile
    real(RK), allocatable, dimension(:) :: YoungsModulus_W, PoissonRatio_W
    real(RK) :: FrictionCoe_s_PW, FrictionCoe_k_PW, FrictionCoe_Roll_PW, RestitutionCoe_n_PW, RestitutionCoe_t_PW
    namelist /WallPhysicalProperty/ YoungsModulus_W, PoissonRatio_W, FrictionCoe_s_PW, FrictionCoe_k_PW, &
             FrictionCoe_Roll_PW, RestitutionCoe_n_PW, RestitutionCoe_t_PW
    integer :: nUnitFile, myistat

    this%nWType = DEM_opt%numWall_type
    allocate(YoungsModulus_W(this%nWType), PoissonRatio_W(this%nWType))
    open(newunit=nUnitFile, file=chFile, status='old', form='formatted', iostat=myistat)
    if (myistat /= 0) then
      print *, "InitWallProperty: cannot open file: "//trim(chFile); stop 1
    end if
    read(nUnitFile, nml=WallPhysicalProperty)
    close(nUnitFile)
    call move_alloc(YoungsModulus_W, this%YoungsModulus_W)
    call move_alloc(PoissonRatio_W, this%PoissonRatio_W)
    this%FrictionCoe_s_PW = FrictionCoe_s_PW; this%FrictionCoe_k_PW = FrictionCoe_k_PW
    this%FrictionCoe_Roll_PW = FrictionCoe_Roll_PW; this%RestitutionCoe_n_PW = RestitutionCoe_n_PW
    this%RestitutionCoe_t_PW = RestitutionCoe_t_PW
  end subroutine InitWallProperty

  ! hand the properties to a library context
  integer function PP_push(this, ctx) result(rc)
    class(PhysicalProperty), intent(in) :: this
    type(c_ptr), intent(in) :: ctx
    rc = SDEM_OK
    if (DEM_opt%CF_Type == CFT_Linear) rc = sdem_set_linear_stiffness(ctx, this%kn, this%kt_over_kn)
    if (rc /= SDEM_OK) return
    rc = sdem_set_particle_property(ctx, int(this%nPType, c_int32_t), this%Diameter, this%Density, &
                                    this%YoungsModulus_P, this%PoissonRatio_P, this%FrictionCoe_s_PP, &
                                    this%FrictionCoe_k_PP, this%FrictionCoe_Roll_PP, this%RestitutionCoe_n_PP, &
                                    this%RestitutionCoe_t_PP)
    if (rc /= SDEM_OK) return
    rc = sdem_set_wall_property(ctx, int(this%nWType, c_int32_t), this%YoungsModulus_W, this%PoissonRatio_W, &
                                this%FrictionCoe_s_PW, this%FrictionCoe_k_PW, this%FrictionCoe_Roll_PW, &
                                this%RestitutionCoe_n_PW, this%RestitutionCoe_t_PW)
  end function PP_push

end module Prtcl_Property
