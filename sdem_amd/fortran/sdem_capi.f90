! sdem_capi.f90 -- ISO_C_BINDING interfaces to libsdem_hip.so (include/sdem.h).
!
! One interface per C entry point, argument for argument; arrays are passed
! as assumed-size dummies so the host's Fortran arrays go through unchanged
! (GPrtcl_PosR(4,n) = n x 4 doubles, GPrtcl_linVel(3,tsize,n) = n x tsize x 3).
module sdem_capi
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: SDEM_OK = 0, SDEM_E_ARG = -1, SDEM_E_HIP = -2, &
                               SDEM_E_CAPACITY = -3, SDEM_E_STATE = -4, SDEM_E_NODEV = -5
  integer(c_int32_t), parameter :: SDEM_CFT_LSD = 1, SDEM_CFT_NLIN = 2, SDEM_CFT_LINEAR = 3

  ! DEMS_Options fields the step reads (Prtcl_Parameters.f90:45-95 of the
  ! reference); doubles first, then 32-bit integers, as struct sdem_options
  type, bind(C) :: sdem_options
    real(c_double) :: dt
    real(c_double) :: gravity(3)
    real(c_double) :: domain_min(3)
    real(c_double) :: domain_max(3)
    real(c_double) :: prtcl_cs_ratio
    real(c_double) :: wall_neighbor_ratio
    integer(c_int32_t) :: is_periodic(3)
    integer(c_int32_t) :: cs_method, cf_type, ct_model, pi_method, pri_method
    integer(c_int32_t) :: cs_numlvls, cntct_list_size, base_wall_id, wall_max_update_iter
    integer(c_int32_t) :: num_prtcl, ifirst
  end type sdem_options

  interface
    integer(c_int) function sdem_options_default(o) bind(C, name="sdem_options_default")
      import :: c_int, sdem_options
      type(sdem_options), intent(out) :: o
    end function
    integer(c_int) function sdem_create(o, ctx) bind(C, name="sdem_create")
      import :: c_int, c_ptr, sdem_options
      type(sdem_options), intent(in) :: o
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function sdem_set_linear_stiffness(ctx, kn, kt_over_kn) &
        bind(C, name="sdem_set_linear_stiffness")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: kn, kt_over_kn
    end function
    integer(c_int) function sdem_set_particle_property(ctx, ntype, diameter, density, youngs, poisson, &
        fs, fk, froll, en, et) bind(C, name="sdem_set_particle_property")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ntype
      real(c_double), intent(in) :: diameter(*), density(*), youngs(*), poisson(*)
      real(c_double), value :: fs, fk, froll, en, et
    end function
    integer(c_int) function sdem_set_wall_property(ctx, nwtype, youngs, poisson, fs, fk, froll, en, et) &
        bind(C, name="sdem_set_wall_property")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nwtype
      real(c_double), intent(in) :: youngs(*), poisson(*)
      real(c_double), value :: fs, fk, froll, en, et
    end function
    integer(c_int) function sdem_add_plane_wall(ctx, p1, p2, p3, p4, user_id, wall_type, both_side, &
        is_infinite, trans_vel, wall_id) bind(C, name="sdem_add_plane_wall")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: p1(3), p2(3), p3(3), p4(3), trans_vel(3)
      integer(c_int32_t), value :: user_id, wall_type, both_side, is_infinite
      integer(c_int32_t), intent(out) :: wall_id
    end function
    integer(c_int) function sdem_set_fixed_particles(ctx, nfix, id, ptype, posr) &
        bind(C, name="sdem_set_fixed_particles")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nfix
      integer(c_int32_t), intent(in) :: id(*), ptype(*)
      real(c_double), intent(in) :: posr(*)
    end function
    integer(c_int) function sdem_set_state(ctx, n, id, ptype, usr_mark, posr, lin_vel, lin_acc, theta, &
        rot_vel, rot_acc) bind(C, name="sdem_set_state")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n
      integer(c_int32_t), intent(in) :: id(*), ptype(*), usr_mark(*)
      real(c_double), intent(in) :: posr(*), lin_vel(*), lin_acc(*), theta(*), rot_vel(*), rot_acc(*)
    end function
    integer(c_int) function sdem_set_contacts(ctx, m, kind, owner_id, partner_id, tan_delta, vrel_init) &
        bind(C, name="sdem_set_contacts")
      import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: m
      integer(c_int32_t), intent(in) :: kind(*), owner_id(*), partner_id(*)
      real(c_double), intent(in) :: tan_delta(*), vrel_init(*)
    end function
    integer(c_int) function sdem_step(ctx, nsteps) bind(C, name="sdem_step")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
    end function
    ! the calls of DEMS_iterate one at a time (Prtcl_DEMSystem.f90:213-248)
    integer(c_int) function sdem_pre_iteration(ctx) bind(C, name="sdem_pre_iteration")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function sdem_comm_for_cntct(ctx) bind(C, name="sdem_comm_for_cntct")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function sdem_find_contacts(ctx) bind(C, name="sdem_find_contacts")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function sdem_find_contacts_pw(ctx) bind(C, name="sdem_find_contacts_pw")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function sdem_clc_acceleration(ctx) bind(C, name="sdem_clc_acceleration")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function sdem_integrate(ctx) bind(C, name="sdem_integrate")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function sdem_comm_for_exchange(ctx) bind(C, name="sdem_comm_for_exchange")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function sdem_remv_released(ctx) bind(C, name="sdem_remv_released")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function sdem_set_profiling(ctx, enable) bind(C, name="sdem_set_profiling")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: enable
    end function
    integer(c_int) function sdem_get_timers(ctx, tot_ms, calls) bind(C, name="sdem_get_timers")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: tot_ms(9)
      integer(c_int64_t), intent(out) :: calls(9)
    end function
    integer(c_int) function sdem_get_num_particles(ctx, n) bind(C, name="sdem_get_num_particles")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: n
    end function
    integer(c_int) function sdem_get_state(ctx, cap, n, id, ptype, usr_mark, posr, lin_vel, lin_acc, theta, &
        rot_vel, rot_acc) bind(C, name="sdem_get_state")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: cap
      integer(c_int32_t), intent(out) :: n
      integer(c_int32_t), intent(out) :: id(*), ptype(*), usr_mark(*)
      real(c_double), intent(out) :: posr(*), lin_vel(*), lin_acc(*), theta(*), rot_vel(*), rot_acc(*)
    end function
    integer(c_int) function sdem_get_forces(ctx, cap, id, force, torque) bind(C, name="sdem_get_forces")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: cap
      integer(c_int32_t), intent(out) :: id(*)
      real(c_double), intent(out) :: force(*), torque(*)
    end function
    integer(c_int) function sdem_get_num_contacts(ctx, m) bind(C, name="sdem_get_num_contacts")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: m
    end function
    integer(c_int) function sdem_get_contacts(ctx, cap, m, kind, owner_id, partner_id, tan_delta, vrel_init) &
        bind(C, name="sdem_get_contacts")
      import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: cap
      integer(c_int64_t), intent(out) :: m
      integer(c_int32_t), intent(out) :: kind(*), owner_id(*), partner_id(*)
      real(c_double), intent(out) :: tan_delta(*), vrel_init(*)
    end function
    integer(c_int) function sdem_get_counts(ctx, num_cntcts) bind(C, name="sdem_get_counts")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: num_cntcts(4)
    end function
    integer(c_int) function sdem_get_iter(ctx, iter_number, next_wall_update) bind(C, name="sdem_get_iter")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: iter_number, next_wall_update
    end function
    integer(c_int) function sdem_util_random_pack(n, radius, lo, hi, periodic, seed, max_attempts, pos) &
        bind(C, name="sdem_util_random_pack")
      import :: c_int, c_double, c_int32_t, c_int64_t
      integer(c_int64_t), value :: n, seed
      real(c_double), intent(in) :: radius(*), lo(3), hi(3)
      integer(c_int32_t), intent(in) :: periodic(3)
      integer(c_int32_t), value :: max_attempts
      real(c_double), intent(out) :: pos(*)
    end function
    type(c_ptr) function sdem_last_error_c(ctx) bind(C, name="sdem_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    subroutine sdem_destroy(ctx) bind(C, name="sdem_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    ! include/sdem_io.h: the reference's restart and visualisation files
    integer(c_int) function sdem_write_restart(ctx, path, mark_links) bind(C, name="sdem_write_restart")
      import :: c_int, c_ptr, c_char, c_int32_t
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: mark_links
    end function
    integer(c_int) function sdem_read_restart(ctx, path) bind(C, name="sdem_read_restart")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
    end function
    integer(c_int) function sdem_read_fixed_coord(ctx, path, num_fix) bind(C, name="sdem_read_fixed_coord")
      import :: c_int, c_ptr, c_char, c_int32_t
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: num_fix
    end function
    integer(c_int) function sdem_write_visu(ctx, path, fields) bind(C, name="sdem_write_visu")
      import :: c_int, c_ptr, c_char, c_int32_t
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: fields
    end function
    integer(c_int) function sdem_xdmf_begin(path, ifirst, ilast, save_visu) bind(C, name="sdem_xdmf_begin")
      import :: c_int, c_char, c_int32_t
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: ifirst, ilast, save_visu
    end function
    integer(c_int) function sdem_xdmf_grid(path, data_base, itime, np, fields) bind(C, name="sdem_xdmf_grid")
      import :: c_int, c_char, c_int32_t
      character(kind=c_char), intent(in) :: path(*), data_base(*)
      integer(c_int32_t), value :: itime, np, fields
    end function
    integer(c_int) function sdem_xdmf_end(path) bind(C, name="sdem_xdmf_end")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: path(*)
    end function
  end interface

  ! fields of the visualisation dump (SDEM_VISU_* of include/sdem_io.h)
  integer(c_int32_t), parameter :: SDEM_VISU_ID = 1, SDEM_VISU_DIAMETER = 2, SDEM_VISU_TYPE = 4, &
                                   SDEM_VISU_USRMARK = 8, SDEM_VISU_LINVEL = 16, SDEM_VISU_LINACC = 32, &
                                   SDEM_VISU_THETA = 64, SDEM_VISU_ROTVEL = 128, SDEM_VISU_ROTACC = 256, &
                                   SDEM_VISU_CNTCTFORCE = 512, SDEM_VISU_TORQUE = 1024

contains

  ! the context's last error message as a Fortran string
  function sdem_last_error(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: i
    msg = ' '
    p = sdem_last_error_c(ctx)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [256])
    do i = 1, 256
      if (s(i) == c_null_char) exit
      msg(i:i) = s(i)
    end do
  end function sdem_last_error

  ! a Fortran string as a NUL-terminated C string
  function c_str(s) result(c)
    character(*), intent(in) :: s
    character(kind=c_char, len=len_trim(s) + 1) :: c
    c = trim(s) // c_null_char
  end function c_str

end module sdem_capi
