! prtcl_demsystem.f90 -- DEMSystem with the reference's call sequence
! (src/DEM/Prtcl_DEMSystem.f90): Initialize reads the .prm file's property,
! geometry and particle-making namelists, or the restart file, and pushes
! them into a libsdem_hip context (DEMS_Initialize :58-195); iterate(itime)
! is one DEMS_iterate (:201-317), i.e. one sdem_step, followed by the
! reference's visualisation dumps and restart backups and its terminal
! report every Cmd_LFile_Freq steps.  With SDEM_PHASES=1 in the environment
! the step runs as the reference's calls one at a time (sdem_pre_iteration
! ... sdem_remv_released, the same bits); with SDEM_TIMERS=1 the report
! includes the reference's step timers (:280-305).  Errors stop the program
! with the library's message, as DEMLogInfo%CheckForError(ErrT_Abort, ...)
! does.
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
    logical :: phases = .false., timers = .false.
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
    call get_environment_variable("SDEM_PHASES", envseed, status=stat)
    this%phases = stat == 0 .and. trim(envseed) == "1"
    call get_environment_variable("SDEM_TIMERS", envseed, status=stat)
    this%timers = stat == 0 .and. trim(envseed) == "1"
    if (this%timers) call check(this, sdem_set_profiling(this%ctx, 1_c_int32_t), "sdem_set_profiling")
    call DEM_IO%Init_visu(chFile, DEM_opt%np_InDomain)
    print '(A,I0,A,I0,A)', "DEM initialized: ", DEM_opt%np_InDomain, " particles, ", DEMGeometry%nPW, " walls"
  end subroutine DEMS_Initialize

  subroutine DEMS_iterate(this, itime)
    class(DEMSystem) :: this
    integer, intent(in) :: itime
    integer(c_int32_t) :: cnt(4), n
    integer(c_int64_t) :: t0, t1, rate
    call system_clock(t0, rate)
    if (this%phases) then
      call phase_by_phase(this)
    else
      call check(this, sdem_step(this%ctx, 1_c_int32_t), "sdem_step")
    end if
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
      if (this%timers) call report_timers(this)
    end if
  end subroutine DEMS_iterate

  ! one DEMS_iterate as the reference's calls (Prtcl_DEMSystem.f90:213-248);
  ! a step that ran out of capacity is run again from its start
  subroutine phase_by_phase(this)
    class(DEMSystem) :: this
    integer :: attempt
    integer(c_int) :: rc
    do attempt = 1, 13
      call check(this, sdem_pre_iteration(this%ctx), "sdem_pre_iteration")
      call check(this, sdem_comm_for_cntct(this%ctx), "sdem_comm_for_cntct")
      rc = sdem_find_contacts(this%ctx)
      if (rc == SDEM_E_CAPACITY) cycle
      call check(this, rc, "sdem_find_contacts")
      rc = sdem_find_contacts_pw(this%ctx)
      if (rc == SDEM_E_CAPACITY) cycle
      call check(this, rc, "sdem_find_contacts_pw")
      call check(this, sdem_clc_acceleration(this%ctx), "sdem_clc_acceleration")
      call check(this, sdem_integrate(this%ctx), "sdem_integrate")
      call check(this, sdem_comm_for_exchange(this%ctx), "sdem_comm_for_exchange")
      call check(this, sdem_remv_released(this%ctx), "sdem_remv_released")
      return
    end do
    call check(this, SDEM_E_CAPACITY, "phase_by_phase")
  end subroutine phase_by_phase

  ! the reference's timer lines (Prtcl_DEMSystem.f90:283-305), in seconds
  subroutine report_timers(this)
    class(DEMSystem) :: this
    real(c_double) :: ms(9)
    integer(c_int64_t) :: nc(9)
    character(len=36), parameter :: label(8) = [character(len=36) :: &
      "PreItertion time [tot, ave]        : ", "Comm_For_Contact [tot, ave]        : ", &
      "CS and CF P-P time [tot, ave]      : ", "CS and CF P-W time [tot, ave]      : ", &
      "Acceleration time [tot, ave]       : ", "Integration time [tot, ave]        : ", &
      "Comm_For_Exchange [tot, ave]       : ", "Write to file time [tot, ave]      : "]
    integer :: k
    call check(this, sdem_get_timers(this%ctx, ms, nc), "sdem_get_timers")
    print '(A,ES12.5,A,ES12.5)', "Execution time [tot, ave] [sec]: ", ms(1) / 1000.0_RK, ", ", &
          ms(1) / 1000.0_RK / real(max(nc(1), 1_c_int64_t), RK)
    do k = 2, 9
      print '(A,ES12.5,A,ES12.5)', label(k - 1), ms(k) / 1000.0_RK, ", ", &
            ms(k) / 1000.0_RK / real(max(nc(k), 1_c_int64_t), RK)
    end do
  end subroutine report_timers

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
