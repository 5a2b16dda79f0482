! prtcl_demsystem.f90 -- DEMSystem with the reference's call sequence
! (src/DEM/Prtcl_DEMSystem.f90): Initialize reads the .prm file's property,
! geometry and particle-making namelists, or the restart file, and pushes
! them into a libsdem_hip context (DEMS_Initialize :58-195); iterate(itime)
! is one DEMS_iterate (:201-317), i.e. one sdem_step, followed by the
! reference's visualisation dumps and restart backups and its terminal
! report every Cmd_LFile_Freq steps.  Errors stop the program with the
! library's message, as DEMLogInfo%CheckForError(ErrT_Abort, ...) does.
module Prtcl_DEMSystem
  use, intrinsic :: iso_c_binding
  use sdem_capi
  use Prtcl_Parameters
  use Prtcl_Property
  use Prtcl_Geometry
  use Prtcl_Variables
  use Prtcl_IOAndVisu
  implicit none
  private

  integer, parameter :: DEFAULT_SEED = 20261203

  type, public :: DEMSystem
    type(c_ptr) :: ctx = c_null_ptr
    real(RK) :: t_steps = 0.0_RK
  contains
    procedure :: Initialize => DEMS_Initialize
    procedure :: iterate => DEMS_iterate
    procedure :: get_state => DEMS_get_state
    procedure :: dump_state => DEMS_dump_state
    procedure :: Final => DEMS_Final
  end type DEMSystem

  type(DEMSystem), public, save :: DEM

contains

  subroutine check(this, rc, where)
    class(DEMSystem), intent(in) :: this
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    if (rc /= SDEM_OK) then
      print '(A,I0,A)', "DEM error in "//where//" (code ", rc, "): "//trim(sdem_last_error(this%ctx))
      stop 1
    end if
  end subroutine check

  subroutine DEMS_Initialize(this, chFile)
    class(DEMSystem) :: this
    character(*), intent(in) :: chFile
    type(sdem_options) :: o
    character(32) :: envseed
    integer :: seed, stat
    integer(c_int32_t) :: n

    seed = DEFAULT_SEED
    call get_environment_variable("SDEM_SEED", envseed, status=stat)
    if (stat == 0) read(envseed, *) seed
    call execute_command_line('mkdir -p '//trim(DEM_opt%Res_Dir)//' '//trim(DEM_opt%RestartDir)//' 2> /dev/null')
    call DEMProperty%InitPrtclProperty(chFile, seed)
    call DEMProperty%InitWallProperty(chFile)
    call DEMGeometry%MakeGeometry(chFile)
    o = DEM_opt%to_c()
    call check(this, sdem_create(o, this%ctx), "sdem_create")
    call check(this, DEMProperty%push(this%ctx), "properties")
    call check(this, DEMGeometry%push(this%ctx), "geometry")
    ! fixed spheres (PIO_ReadFixedCoord): the library takes them before the
    ! moving particles' state
    if (DEM_opt%numPrtclFix > 0) call DEM_IO%ReadFixedCoord(this%ctx)
    if (DEM_opt%RestartFlag) then
      ! PIO_Read_Restart and PIO_RestartCL: state and contact list
      call DEM_IO%Read_Restart(this%ctx)
      call check(this, sdem_get_num_particles(this%ctx, n), "sdem_get_num_particles")
      DEM_opt%np_InDomain = n
    else
      call GL_MakingAllPrtcl(chFile)
      DEM_opt%np_InDomain = nlocal
      call check(this, sdem_set_state(this%ctx, int(nlocal, c_int32_t), GPrtcl_id, GPrtcl_pType, &
                                      GPrtcl_UsrMark, GPrtcl_PosR, GPrtcl_linVel, GPrtcl_linAcc, GPrtcl_theta, &
                                      GPrtcl_rotVel, GPrtcl_rotAcc), "sdem_set_state")
    end if
    call DEM_IO%Init_visu(chFile, DEM_opt%np_InDomain)
    print '(A,I0,A,I0,A)', "DEM initialized: ", DEM_opt%np_InDomain, " particles, ", DEMGeometry%nPW, " walls"
  end subroutine DEMS_Initialize

  subroutine DEMS_iterate(this, itime)
    class(DEMSystem) :: this
    integer, intent(in) :: itime
    integer(c_int32_t) :: cnt(4), n
    integer(c_int64_t) :: t0, t1, rate
    call system_clock(t0, rate)
    call check(this, sdem_step(this%ctx, 1_c_int32_t), "sdem_step")
    call system_clock(t1)
    this%t_steps = this%t_steps + real(t1 - t0, RK) / real(rate, RK)
    ! results and backups (DEMS_iterate :250-265)
    if (mod(itime, DEM_opt%SaveVisu) == 0 .or. mod(itime, DEM_opt%BackupFreq) == 0 .or. &
        itime == DEM_opt%ilast) then
      call check(this, sdem_get_num_particles(this%ctx, n), "sdem_get_num_particles")
      DEM_opt%np_InDomain = n
      if (mod(itime, DEM_opt%SaveVisu) == 0) call DEM_IO%Dump_visu(this%ctx, itime, int(n))
      if (mod(itime, DEM_opt%BackupFreq) == 0 .or. itime == DEM_opt%ilast) then
        call DEM_IO%Write_Restart(this%ctx, itime)
        call DEM_IO%Delete_Prev_Restart(itime)
      end if
    end if
    if (mod(itime, DEM_opt%Cmd_LFile_Freq) == 0 .or. itime == DEM_opt%ilast) then
      call check(this, sdem_get_counts(this%ctx, cnt), "sdem_get_counts")
      call check(this, sdem_get_num_particles(this%ctx, n), "sdem_get_num_particles")
      print '(A,I0,A,I0,A,4(1X,I0))', "iter ", itime, ": ", n, " particles, contacts P-P P-GP P-FP P-W:", cnt
    end if
  end subroutine DEMS_iterate

  ! current state into the host arrays (ascending particle id)
  subroutine DEMS_get_state(this)
    class(DEMSystem) :: this
    integer(c_int32_t) :: n
    call check(this, sdem_get_num_particles(this%ctx, n), "sdem_get_num_particles")
    call GL_AllocateAllVar(int(n))
    call check(this, sdem_get_state(this%ctx, n, n, GPrtcl_id, GPrtcl_pType, GPrtcl_UsrMark, GPrtcl_PosR, &
                                    GPrtcl_linVel, GPrtcl_linAcc, GPrtcl_theta, GPrtcl_rotVel, GPrtcl_rotAcc), &
               "sdem_get_state")
    nlocal = n
  end subroutine DEMS_get_state

  ! raw state dump (stream file): n, tsize, rsize (int32), id, type, mark
  ! (int32 x n), PosR (4 x n), linVel, linAcc (3 x tsize x n), theta (3 x n),
  ! rotVel, rotAcc (3 x rsize x n) as doubles
  subroutine DEMS_dump_state(this, chFile)
    class(DEMSystem) :: this
    character(*), intent(in) :: chFile
    integer :: u
    call this%get_state()
    open(newunit=u, file=chFile, access='stream', form='unformatted', status='replace')
    write(u) int(nlocal, c_int32_t), int(tsize, c_int32_t), int(rsize, c_int32_t)
    write(u) GPrtcl_id(1:nlocal), GPrtcl_pType(1:nlocal), GPrtcl_UsrMark(1:nlocal)
    write(u) GPrtcl_PosR(:, 1:nlocal), GPrtcl_linVel(:, :, 1:nlocal), GPrtcl_linAcc(:, :, 1:nlocal)
    write(u) GPrtcl_theta(:, 1:nlocal), GPrtcl_rotVel(:, :, 1:nlocal), GPrtcl_rotAcc(:, :, 1:nlocal)
    close(u)
  end subroutine DEMS_dump_state

  subroutine DEMS_Final(this)
    class(DEMSystem) :: this
    if (c_associated(this%ctx)) call sdem_destroy(this%ctx)
    this%ctx = c_null_ptr
  end subroutine DEMS_Final

end module Prtcl_DEMSystem
