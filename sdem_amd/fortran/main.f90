! main.f90 -- the reference's driver (src/DEM/main.f90:16-38) on one GPU:
! read the .prm file, initialise (from a restart file with RestartFlag = T),
! write the visualisation dump of ifirst-1, iterate ifirst..ilast with the
! reference's dumps and restart backups, close the XDMF index.  No MPI.
!
!   sdem_main <file.prm>          run the case
!   sdem_main <file.prm> <dir>    also write raw state dumps <dir>/state_<iter>.bin
!                                 before the first and after the last step
!   sdem_main --abi-check         print the size of struct sdem_options and exit
!   sdem_main --dry-run <file.prm>  read the input and make the particles, print
!                                 a summary and the walls' corners, touch no GPU
program main_DEM
  use, intrinsic :: iso_c_binding
  use sdem_capi
  use Prtcl_Parameters
  use Prtcl_Property
  use Prtcl_Geometry
  use Prtcl_Variables
  use Prtcl_DEMSystem
  use Prtcl_IOAndVisu
  implicit none
  character(len=256) :: chDEMPrm, chDump, fname
  integer :: i, nargs
  type(sdem_options) :: o

  nargs = command_argument_count()
  if (nargs < 1) then
    print *, "usage: sdem_main <file.prm> [dump_dir] | --abi-check"; stop 1
  end if
  call get_command_argument(1, chDEMPrm)
  if (trim(chDEMPrm) == "--abi-check") then
    print '(A,I0)', "sizeof(sdem_options) = ", c_sizeof(o)
    stop
  end if
  if (trim(chDEMPrm) == "--dry-run") then
    call get_command_argument(2, chDEMPrm)
    call DEM_opt%ReadDEMOption(chDEMPrm)
    call DEMProperty%InitPrtclProperty(chDEMPrm, 20261203)
    call DEMProperty%InitWallProperty(chDEMPrm)
    call DEMGeometry%MakeGeometry(chDEMPrm)
    call GL_MakingAllPrtcl(chDEMPrm)
    o = DEM_opt%to_c()
    print '(A,I0,A,I0,A,I0,A,I0)', "numPrtcl ", DEM_opt%numPrtcl, " types ", DEMProperty%nPType, &
          " walls ", DEMGeometry%nPW, " steps ", DEM_opt%ilast - DEM_opt%ifirst + 1
    print '(A,ES14.6,A,3I2,A,5I2)', "dt ", o%dt, " periodic", o%is_periodic, " methods", o%cs_method, &
          o%cf_type, o%ct_model, o%pi_method, o%pri_method
    print '(A,3ES14.6)', "max point ", o%domain_max
    print '(A,*(1X,I0))', "bins", DEMProperty%nPrtcl_in_Bin
    print '(A,4ES14.6)', "first particle ", GPrtcl_PosR(:, 1)
    do i = 1, DEMGeometry%nPW
      print '(A,I0,12ES24.16)', "wall ", i, DEMGeometry%Point1(:, i), DEMGeometry%Point2(:, i), &
            DEMGeometry%Point3(:, i), DEMGeometry%Point4(:, i)
    end do
    stop
  end if
  chDump = ""
  if (nargs >= 2) call get_command_argument(2, chDump)

  call DEM_opt%ReadDEMOption(chDEMPrm)
  call DEM%Initialize(chDEMPrm)
  if (len_trim(chDump) > 0) then
    write(fname, '(A,"/state_",I10.10,".bin")') trim(chDump), DEM_opt%ifirst - 1
    call DEM%dump_state(fname)
  end if
  call DEM_IO%Dump_visu(DEM%ctx, DEM_opt%ifirst - 1, DEM_opt%np_InDomain)
  do i = DEM_opt%ifirst, DEM_opt%ilast
    call DEM%iterate(i)
  end do
  call DEM_IO%Final_visu()
  if (len_trim(chDump) > 0) then
    write(fname, '(A,"/state_",I10.10,".bin")') trim(chDump), DEM_opt%ilast
    call DEM%dump_state(fname)
  end if
  print '(A,ES12.5,A)', "stepping time ", DEM%t_steps, " s"
  call DEM%Final()
  print *, "Good job! DEM finished successfully!"
end program main_DEM
