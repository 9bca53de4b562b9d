! fcmd_capi.f90 -- ISO_C_BINDING interfaces to the engine's C ABI
! (include/fcmd.h).  Used by the drop-in plugin subroutines in this directory;
! any Fortran program can use it directly.
module fcmd_capi
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: FCMD_MAXSORTS = 10
  integer(c_int), parameter :: FCMD_FP64 = 0, FCMD_MIXED = 1, FCMD_MIXED_KAHAN = 2
  integer(c_int), parameter :: FCMD_INT_VV = 0, FCMD_INT_MAV = 1, FCMD_INT_NH = 2
  integer(c_int), parameter :: FCMD_NB_PBC = 0, FCMD_NB_MAV = 1
  integer(c_int), parameter :: FCMD_ACC_1 = 0, FCMD_ACC_M = 1, FCMD_ACC_TABLE = 2
  integer(c_int), parameter :: FCMD_ST_N = 0, FCMD_ST_NM = 1, FCMD_ST_FRE = 2

  type, bind(C) :: fcmd_params
     integer(c_int) :: natom, nsorts
     real(c_double) :: box(3)
     real(c_double) :: mas(FCMD_MAXSORTS), sig(FCMD_MAXSORTS), eps(FCMD_MAXSORTS)
     integer(c_int) :: num(FCMD_MAXSORTS)
     real(c_double) :: atrcut, atrskin, atskin, dt, Q, att
     integer(c_int) :: nbtim, btim, atnenum, ethst_nh
     integer(c_int) :: integrator, neighbor, accatom, stoddard
     real(c_double) :: eps_ab(FCMD_MAXSORTS*FCMD_MAXSORTS)
     real(c_double) :: sig_ab(FCMD_MAXSORTS*FCMD_MAXSORTS)
     integer(c_int) :: nh_mask, pad
  end type fcmd_params

  type, bind(C) :: fcmd_scalars
     real(c_double) :: atrskin, psi, s, ekin, epot, sumv2
     real(c_double) :: dmmh(256)
     real(c_double) :: dmms, dmmsu, dmmax, dmmav, dmmsr
     integer(c_int) :: dmmst, dmmi, step, npairs, nrebuild, nh_iters, err, nlist
     real(c_double) :: virial, pressure
  end type fcmd_scalars

  interface
     integer(c_int) function fcmd_device_count() bind(C, name='fcmd_device_count')
       import :: c_int
     end function fcmd_device_count

     type(c_ptr) function fcmd_error_text(code) bind(C, name='fcmd_error_text')
       import :: c_int, c_ptr
       integer(c_int), value :: code
     end function fcmd_error_text

     integer(c_int) function fcmd_create(sys, ntraj, params, precision, list_width) &
          bind(C, name='fcmd_create')
       import :: c_int, c_ptr, fcmd_params
       type(c_ptr) :: sys
       integer(c_int), value :: ntraj, precision, list_width
       type(fcmd_params) :: params
     end function fcmd_create

     subroutine fcmd_destroy(sys) bind(C, name='fcmd_destroy')
       import :: c_ptr
       type(c_ptr), value :: sys
     end subroutine fcmd_destroy

     integer(c_int) function fcmd_set_params(sys, traj, params) bind(C, name='fcmd_set_params')
       import :: c_int, c_ptr, fcmd_params
       type(c_ptr), value :: sys
       integer(c_int), value :: traj
       type(fcmd_params) :: params
     end function fcmd_set_params

     integer(c_int) function fcmd_upload(sys, traj, x, y, z, vx, vy, vz, ax, ay, az, api) &
          bind(C, name='fcmd_upload')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: sys
       integer(c_int), value :: traj
       real(c_double) :: x(*), y(*), z(*), vx(*), vy(*), vz(*), ax(*), ay(*), az(*)
       integer(c_int) :: api(*)
     end function fcmd_upload

     integer(c_int) function fcmd_download(sys, traj, x, y, z, vx, vy, vz, ax, ay, az) &
          bind(C, name='fcmd_download')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: sys
       integer(c_int), value :: traj
       real(c_double) :: x(*), y(*), z(*), vx(*), vy(*), vz(*), ax(*), ay(*), az(*)
     end function fcmd_download

     integer(c_int) function fcmd_set_scalars(sys, traj, sc) bind(C, name='fcmd_set_scalars')
       import :: c_int, c_ptr, fcmd_scalars
       type(c_ptr), value :: sys
       integer(c_int), value :: traj
       type(fcmd_scalars) :: sc
     end function fcmd_set_scalars

     integer(c_int) function fcmd_get_scalars(sys, traj, sc) bind(C, name='fcmd_get_scalars')
       import :: c_int, c_ptr, fcmd_scalars
       type(c_ptr), value :: sys
       integer(c_int), value :: traj
       type(fcmd_scalars) :: sc
     end function fcmd_get_scalars

     integer(c_int) function fcmd_neighbor(sys, traj, step) bind(C, name='fcmd_neighbor')
       import :: c_int, c_ptr
       type(c_ptr), value :: sys
       integer(c_int), value :: traj
       integer(c_int) :: step
     end function fcmd_neighbor

     integer(c_int) function fcmd_accatom(sys) bind(C, name='fcmd_accatom')
       import :: c_int, c_ptr
       type(c_ptr), value :: sys
     end function fcmd_accatom

     integer(c_int) function fcmd_integrator(sys, step) bind(C, name='fcmd_integrator')
       import :: c_int, c_ptr
       type(c_ptr), value :: sys
       integer(c_int) :: step
     end function fcmd_integrator

     integer(c_int) function fcmd_stoddard(sys, traj, step, cl, cs, cc, csi, cek, ceks, &
          clf, csf, ccf, csif, cekf, ceksf, cgs, cgsf, cgsfa, cfpt, cfptf, counters) &
          bind(C, name='fcmd_stoddard')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: sys
       integer(c_int), value :: traj, step
       integer(c_int) :: cl(*), cs(*), cc(*), csi(*), clf(*), csf(*), ccf(*), csif(*)
       integer(c_int) :: cgs(*), cgsf(*), cgsfa(*), cfpt(*), cfptf(*), counters(6)
       real(c_double) :: cek(*), ceks(*), cekf(*), ceksf(*)
     end function fcmd_stoddard

     integer(c_int) function fcmd_init_random(seed, natom, box, rmin, att, mas, api, &
          x, y, z, vx, vy, vz) bind(C, name='fcmd_init_random')
       import :: c_int, c_double, c_long_long
       integer(c_long_long), value :: seed
       integer(c_int), value :: natom
       real(c_double) :: box(3)
       real(c_double), value :: rmin, att
       real(c_double) :: mas(*)
       integer(c_int) :: api(*)
       real(c_double) :: x(*), y(*), z(*), vx(*), vy(*), vz(*)
     end function fcmd_init_random

     integer(c_int) function fcmd_halflist(sys, traj, atnlist, cap, atnidx, length) &
          bind(C, name='fcmd_halflist')
       import :: c_int, c_ptr, c_long_long
       type(c_ptr), value :: sys
       integer(c_int), value :: traj
       integer(c_int) :: atnlist(*), atnidx(*)
       integer(c_long_long), value :: cap
       integer(c_long_long) :: length
     end function fcmd_halflist
  end interface

contains

  ! the reference's STOP message for an engine error code
  subroutine fcmd_stop_on_error(rc)
    integer(c_int), intent(in) :: rc
    character(kind=c_char), pointer :: msg(:)
    character(len=200) :: text
    integer :: k
    if (rc == 0) return
    call c_f_pointer(fcmd_error_text(rc), msg, [200])
    text = ' '
    do k = 1, 200
       if (msg(k) == c_null_char) exit
       text(k:k) = msg(k)
    end do
    write (*,*) trim(text)
    stop
  end subroutine fcmd_stop_on_error

end module fcmd_capi
