! prtcl_parameters.f90 -- the &DEMOptions namelist of the reference's .prm
! files (DO_ReadDEMOption, src/DEM/Prtcl_Parameters.f90:104-205).  The
! namelist group has the same variables, so the reference's input files are
! read unchanged; the values the GPU step needs are forwarded through
! struct sdem_options.
module Prtcl_Parameters
  use, intrinsic :: iso_c_binding
  use sdem_capi
  implicit none
  private

  integer, parameter, public :: RK = c_double
  integer, parameter, public :: CSM_NBS_Munjiza = 1, CSM_NBS_Munjiza_Hrchl = 2
  integer, parameter, public :: CFT_LSD = 1, CFT_nLin = 2, CFT_Linear = 3
  integer, parameter, public :: CTM_ConstantTorque = 1, CTM_VariableTorque = 2
  integer, parameter, public :: PIM_FE = 1, PIM_AB2 = 2, PIM_AB3 = 3

  ! defaults as the reference's DEMS_Options (Prtcl_Parameters.f90:45-95)
  type, public :: DEMS_Options
    logical :: RestartFlag = .false.
    integer :: numPrtcl = 8000
    integer :: numPrtclFix = 1000
    integer :: np_InDomain = 0
    integer :: ifirst = 1
    integer :: ilast = 1
    real(RK) :: dt = 1.0E-5_RK
    real(RK) :: gravity(3) = (/0.0_RK, -9.81_RK, 0.0_RK/)
    real(RK) :: SimDomain_min(3) = 0.0_RK
    real(RK) :: SimDomain_max(3) = 1.0_RK
    logical :: IsPeriodic(3) = .false.
    real(RK) :: Prtcl_cs_ratio = 1.0_RK
    integer :: CS_Method = CSM_NBS_Munjiza
    integer :: CF_Type = CFT_LSD
    integer :: CT_Model = CTM_ConstantTorque
    integer :: PI_Method = PIM_FE
    integer :: PRI_Method = PIM_FE
    integer :: numPrtcl_Type = 1
    integer :: numWall_type = 1
    integer :: CntctList_Size = 6
    integer :: CS_numlvls = 0
    integer :: Base_wall_id = 10000000
    integer :: Wall_max_update_iter = 100
    real(RK) :: Wall_neighbor_ratio = 2.0_RK
    character(64) :: RunName = "DEMRun"
    character(64) :: Res_Dir = "."
    character(64) :: RestartDir = "."
    integer :: SaveVisu = 1000
    integer :: BackupFreq = 100000
    integer :: Cmd_LFile_Freq = 500
    integer :: LF_file_lvl = 5
    integer :: LF_cmdw_lvl = 3
    integer :: GeometrySource = 0
    character(64) :: Geom_Dir = "DEMGeom.stl"
  contains
    procedure :: ReadDEMOption => DO_ReadDEMOption
    procedure :: to_c => DO_to_c
  end type DEMS_Options

  type(DEMS_Options), public, save :: DEM_opt

contains

  subroutine DO_ReadDEMOption(this, chFile)
    class(DEMS_Options) :: this
    character(*), intent(in) :: chFile
    character(64) :: RunName, Res_Dir, RestartDir, Geom_Dir
    integer :: numPrtcl, numPrtclFix, CS_Method, CF_Type, CT_Model, PI_Method, PRI_Method
    integer :: numPrtcl_Type, numWall_type, CntctList_Size, CS_numlvls, Base_wall_id
    integer :: Wall_max_update_iter, SaveVisuDEM, BackupFreqDEM, Cmd_LFile_Freq, LF_file_lvl, LF_cmdw_lvl
    integer :: GeometrySource, ifirstDEM, ilastDEM, nUnitFile, myistat
    real(RK) :: dtDEM, Wall_neighbor_ratio, Prtcl_cs_ratio
    real(RK), dimension(3) :: gravity, minpoint, maxpoint
    logical, dimension(3) :: IsPeriodic
    logical :: RestartFlag
    NAMELIST /DEMOptions/ RestartFlag, numPrtcl, numPrtclFix, dtDEM, gravity, minpoint, maxpoint, CS_Method, &
                          CF_Type, CT_Model, PI_Method, PRI_Method, numPrtcl_Type, numWall_type, CS_numlvls, &
                          CntctList_Size, Base_wall_id, Wall_max_update_iter, Wall_neighbor_ratio, RunName, &
                          Res_Dir, RestartDir, BackupFreqDEM, SaveVisuDEM, Cmd_LFile_Freq, LF_file_lvl, &
                          LF_cmdw_lvl, GeometrySource, Geom_Dir, ifirstDEM, ilastDEM, Prtcl_cs_ratio, IsPeriodic

    ! start from the defaults so that keys missing from the file keep them
    RestartFlag = this%RestartFlag; numPrtcl = this%numPrtcl; numPrtclFix = this%numPrtclFix
    dtDEM = this%dt; gravity = this%gravity; minpoint = this%SimDomain_min; maxpoint = this%SimDomain_max
    CS_Method = this%CS_Method; CF_Type = this%CF_Type; CT_Model = this%CT_Model
    PI_Method = this%PI_Method; PRI_Method = this%PRI_Method; numPrtcl_Type = this%numPrtcl_Type
    numWall_type = this%numWall_type; CS_numlvls = this%CS_numlvls; CntctList_Size = this%CntctList_Size
    Base_wall_id = this%Base_wall_id; Wall_max_update_iter = this%Wall_max_update_iter
    Wall_neighbor_ratio = this%Wall_neighbor_ratio; RunName = this%RunName; Res_Dir = this%Res_Dir
    RestartDir = this%RestartDir; BackupFreqDEM = this%BackupFreq; SaveVisuDEM = this%SaveVisu
    Cmd_LFile_Freq = this%Cmd_LFile_Freq; LF_file_lvl = this%LF_file_lvl; LF_cmdw_lvl = this%LF_cmdw_lvl
    GeometrySource = this%GeometrySource; Geom_Dir = this%Geom_Dir; ifirstDEM = this%ifirst
    ilastDEM = this%ilast; Prtcl_cs_ratio = this%Prtcl_cs_ratio; IsPeriodic = this%IsPeriodic

    open(newunit=nUnitFile, file=chFile, status='old', form='formatted', iostat=myistat)
    if (myistat /= 0) then
      print *, "Cannot open file: "//trim(adjustl(chFile)); stop 1
    end if
    read(nUnitFile, nml=DEMOptions)
    close(nUnitFile)

    this%RestartFlag = RestartFlag
    this%numPrtcl = numPrtcl
    this%numPrtclFix = numPrtclFix
    this%dt = dtDEM
    this%ifirst = ifirstDEM
    this%ilast = ilastDEM
    this%gravity = gravity
    this%SimDomain_min = minpoint
    this%SimDomain_max = maxpoint
    this%IsPeriodic = IsPeriodic
    this%Prtcl_cs_ratio = Prtcl_cs_ratio
    this%CS_Method = CS_Method
    this%CF_Type = CF_Type
    this%CT_Model = CT_Model
    this%PI_Method = PI_Method
    this%PRI_Method = PRI_Method
    this%numPrtcl_Type = numPrtcl_Type
    this%numWall_type = numWall_type
    this%CntctList_Size = CntctList_Size
    this%CS_numlvls = CS_numlvls
    this%Base_wall_id = Base_wall_id
    this%Wall_max_update_iter = Wall_max_update_iter
    this%Wall_neighbor_ratio = Wall_neighbor_ratio
    this%RunName = RunName
    this%Res_Dir = Res_Dir
    this%RestartDir = RestartDir
    this%BackupFreq = BackupFreqDEM
    this%SaveVisu = SaveVisuDEM
    this%Cmd_LFile_Freq = Cmd_LFile_Freq
    this%LF_file_lvl = LF_file_lvl
    this%LF_cmdw_lvl = LF_cmdw_lvl
    this%GeometrySource = GeometrySource
    if (this%GeometrySource == 2) this%Geom_Dir = Geom_Dir
  end subroutine DO_ReadDEMOption

  ! the fields the GPU step reads, as struct sdem_options
  function DO_to_c(this) result(o)
    class(DEMS_Options), intent(in) :: this
    type(sdem_options) :: o
    integer :: rc, d
    rc = sdem_options_default(o)
    o%dt = this%dt
    o%gravity = this%gravity
    o%domain_min = this%SimDomain_min
    o%domain_max = this%SimDomain_max
    o%prtcl_cs_ratio = this%Prtcl_cs_ratio
    o%wall_neighbor_ratio = this%Wall_neighbor_ratio
    do d = 1, 3
      o%is_periodic(d) = merge(1, 0, this%IsPeriodic(d))
    end do
    o%cs_method = this%CS_Method
    o%cf_type = this%CF_Type
    o%ct_model = this%CT_Model
    o%pi_method = this%PI_Method
    o%pri_method = this%PRI_Method
    o%cs_numlvls = this%CS_numlvls
    o%cntct_list_size = this%CntctList_Size
    o%base_wall_id = this%Base_wall_id
    o%wall_max_update_iter = this%Wall_max_update_iter
    o%num_prtcl = this%numPrtcl
    o%ifirst = this%ifirst
  end function DO_to_c

end module Prtcl_Parameters
