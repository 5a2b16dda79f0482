! prtcl_ioandvisu.f90 -- DEM_IO with the reference's procedures
! (src/DEM/Prtcl_IOAndVisu.f90) for one rank: the restart files
! RestartFor<RunName><itime>, the fixed spheres' FixedSpheresCoord.dat and
! the visualisation dumps
! PartVisuFor<RunName><itime> with their XDMF index, all written and read by
! libsdem_hip.so (include/sdem_io.h) in the reference's binary layouts.
! &PrtclVisuOption is read as PIO_Init_visu does (:64-132); a .prm file
! without it writes no visualisation files.
module Prtcl_IOAndVisu
  use, intrinsic :: iso_c_binding
  use sdem_capi
  use Prtcl_Parameters
  implicit none
  private

  type, public :: Prtcl_IO_Visu
    logical :: visu_on = .false.
    logical :: saveXDMFOnce = .true.
    integer(c_int32_t) :: fields = 0
    integer :: Prev_BackUp_itime = 1234567
  contains
    procedure :: Init_visu => PIO_Init_visu
    procedure :: Dump_visu => PIO_Dump_visu
    procedure :: Final_visu => PIO_Final_visu
    procedure :: Read_Restart => PIO_Read_Restart
    procedure :: ReadFixedCoord => PIO_ReadFixedCoord
    procedure :: Write_Restart => PIO_Write_Restart
    procedure :: Delete_Prev_Restart => PIO_Delete_Prev_Restart
  end type Prtcl_IO_Visu

  type(Prtcl_IO_Visu), public, save :: DEM_IO

contains

  subroutine io_check(ctx, rc, where)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    if (rc /= SDEM_OK) then
      if (c_associated(ctx)) then
        print '(A,I0,A)', "DEM error in "//where//" (code ", rc, "): "//trim(sdem_last_error(ctx))
      else
        print '(A,I0,A)', "DEM error in "//where//" (code ", rc, ")"
      end if
      stop 1
    end if
  end subroutine io_check

  function xdmf_file() result(f)
    character(len=256) :: f
    f = trim(DEM_opt%Res_Dir)//"PartVisuFor"//trim(DEM_opt%RunName)//".xmf"
  end function xdmf_file

  function visu_base() result(f)
    character(len=256) :: f
    f = "PartVisuFor"//trim(DEM_opt%RunName)
  end function visu_base

  function restart_file(itime) result(f)
    integer, intent(in) :: itime
    character(len=256) :: f
    write(f, "(A,I10.10)") trim(DEM_opt%RestartDir)//"RestartFor"//trim(DEM_opt%RunName), itime
  end function restart_file

  ! &PrtclVisuOption, the XDMF header and, with saveXDMFOnce, every grid
  ! and the tail at once (PIO_Init_visu :64-132)
  subroutine PIO_Init_visu(this, chFile, np)
    class(Prtcl_IO_Visu) :: this
    character(*), intent(in) :: chFile
    integer, intent(in) :: np
    logical :: saveXDMFOnce, save_ID, save_Diameter, save_Type, save_UsrMark, save_LinVel
    logical :: save_LinAcc, save_Theta, save_RotVel, save_RotAcc, save_CntctForce, save_Torque
    NAMELIST /PrtclVisuOption/ saveXDMFOnce, save_ID, save_Diameter, save_Type, save_UsrMark, save_LinVel, &
                               save_LinAcc, save_Theta, save_RotVel, save_RotAcc, save_CntctForce, save_Torque
    integer :: u, ios, ifld, nflds
    saveXDMFOnce = .true.
    save_ID = .false.; save_Diameter = .false.; save_Type = .false.; save_UsrMark = .false.
    save_LinVel = .false.; save_LinAcc = .false.; save_Theta = .false.; save_RotVel = .false.
    save_RotAcc = .false.; save_CntctForce = .false.; save_Torque = .false.
    open(newunit=u, file=chFile, status='old', form='formatted', iostat=ios)
    if (ios /= 0) then
      print *, "PIO_Init_visu: cannot open file: "//trim(chFile); stop 1
    end if
    read(u, nml=PrtclVisuOption, iostat=ios)
    close(u)
    this%visu_on = (ios == 0)
    if (.not. this%visu_on) return
    this%saveXDMFOnce = saveXDMFOnce
    this%fields = 0
    if (save_ID) this%fields = ior(this%fields, SDEM_VISU_ID)
    if (save_Diameter) this%fields = ior(this%fields, SDEM_VISU_DIAMETER)
    if (save_Type) this%fields = ior(this%fields, SDEM_VISU_TYPE)
    if (save_UsrMark) this%fields = ior(this%fields, SDEM_VISU_USRMARK)
    if (save_LinVel) this%fields = ior(this%fields, SDEM_VISU_LINVEL)
    if (save_LinAcc) this%fields = ior(this%fields, SDEM_VISU_LINACC)
    if (save_Theta) this%fields = ior(this%fields, SDEM_VISU_THETA)
    if (save_RotVel) this%fields = ior(this%fields, SDEM_VISU_ROTVEL)
    if (save_RotAcc) this%fields = ior(this%fields, SDEM_VISU_ROTACC)
    if (save_CntctForce) this%fields = ior(this%fields, SDEM_VISU_CNTCTFORCE)
    if (save_Torque) this%fields = ior(this%fields, SDEM_VISU_TORQUE)
    call io_check(c_null_ptr, sdem_xdmf_begin(c_str(xdmf_file()), int(DEM_opt%ifirst, c_int32_t), &
                  int(DEM_opt%ilast, c_int32_t), int(DEM_opt%SaveVisu, c_int32_t)), "sdem_xdmf_begin")
    if (.not. this%saveXDMFOnce) return
    nflds = (DEM_opt%ilast - DEM_opt%ifirst + 1) / DEM_opt%SaveVisu + 1
    do ifld = 1, nflds
      call io_check(c_null_ptr, sdem_xdmf_grid(c_str(xdmf_file()), c_str(visu_base()), &
                    int((ifld - 1) * DEM_opt%SaveVisu + DEM_opt%ifirst - 1, c_int32_t), int(np, c_int32_t), &
                    this%fields), "sdem_xdmf_grid")
    end do
    call io_check(c_null_ptr, sdem_xdmf_end(c_str(xdmf_file())), "sdem_xdmf_end")
  end subroutine PIO_Init_visu

  ! PIO_Dump_visu (:1193-1267): the XDMF grid (unless written at once) and
  ! the binary dump
  subroutine PIO_Dump_visu(this, ctx, itime, np)
    class(Prtcl_IO_Visu) :: this
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: itime, np
    character(len=256) :: f
    if (.not. this%visu_on) return
    if (.not. this%saveXDMFOnce) &
      call io_check(ctx, sdem_xdmf_grid(c_str(xdmf_file()), c_str(visu_base()), int(itime, c_int32_t), &
                    int(np, c_int32_t), this%fields), "sdem_xdmf_grid")
    if (np <= 0) return
    write(f, "(A,I10.10)") trim(DEM_opt%Res_Dir)//trim(visu_base()), itime
    call io_check(ctx, sdem_write_visu(ctx, c_str(f), this%fields), "sdem_write_visu")
  end subroutine PIO_Dump_visu

  ! PIO_Final_visu (:1167-1188)
  subroutine PIO_Final_visu(this)
    class(Prtcl_IO_Visu) :: this
    if (.not. this%visu_on .or. this%saveXDMFOnce) return
    call io_check(c_null_ptr, sdem_xdmf_end(c_str(xdmf_file())), "sdem_xdmf_end")
  end subroutine PIO_Final_visu

  ! PIO_Read_Restart + PIO_RestartCL (:731-844, :296-377): the file of
  ! itime = ifirst - 1
  subroutine PIO_Read_Restart(this, ctx)
    class(Prtcl_IO_Visu) :: this
    type(c_ptr), intent(in) :: ctx
    call io_check(ctx, sdem_read_restart(ctx, c_str(restart_file(DEM_opt%ifirst - 1))), "sdem_read_restart")
  end subroutine PIO_Read_Restart

  ! PIO_ReadFixedCoord (:164-291): RestartDir/FixedSpheresCoord.dat
  subroutine PIO_ReadFixedCoord(this, ctx)
    class(Prtcl_IO_Visu) :: this
    type(c_ptr), intent(in) :: ctx
    call io_check(ctx, sdem_read_fixed_coord(ctx, c_str(trim(DEM_opt%RestartDir)//"FixedSpheresCoord.dat"), &
                  int(DEM_opt%numPrtclFix, c_int32_t)), "sdem_read_fixed_coord")
  end subroutine PIO_ReadFixedCoord

  ! PIO_Write_Restart (:849-1031); the marks become the link counts, as
  ! the reference's writer leaves them
  subroutine PIO_Write_Restart(this, ctx, itime)
    class(Prtcl_IO_Visu) :: this
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: itime
    call io_check(ctx, sdem_write_restart(ctx, c_str(restart_file(itime)), 1_c_int32_t), "sdem_write_restart")
  end subroutine PIO_Write_Restart

  ! PIO_Delete_Prev_Restart (:137-159): only the newest restart file stays
  subroutine PIO_Delete_Prev_Restart(this, itime)
    class(Prtcl_IO_Visu) :: this
    integer, intent(in) :: itime
    integer :: u, ios
    open(newunit=u, file=restart_file(this%Prev_BackUp_itime), status='old', iostat=ios)
    if (ios == 0) close(u, status='delete')
    this%Prev_BackUp_itime = itime
  end subroutine PIO_Delete_Prev_Restart

end module Prtcl_IOAndVisu
