! fcmd_dropin_state.f90 -- the device handle shared by the drop-in plugin
! subroutines.  Names carry an fcmd_ prefix so they cannot collide with the
! host program's COMMON variables (ATOMP has s, Q, psi ...).
module fcmd_dropin_state
  use fcmd_capi
  implicit none
  type(c_ptr), save :: fcmd_sys = c_null_ptr
  ! ATOMC stays on the device between the driver's output steps (the
  ! main.f:117-168 schedule); FCMD_RESIDENT=0 copies it both ways on every
  ! call instead
  logical, save :: fcmd_resident = .true.
  ! initialisation: until the driver's stoddard(-1) (main.f:112, called once
  ! the initial state is complete, right before the step loop) the host may
  ! read ATOMC after any integrator call (init-cube1.f:119-125 runs warm-up
  ! steps and then stopdead), so every integrator call hands it back
  logical, save :: fcmd_initphase = .true.
  ! the linked neighbor plugin prints the per-rebuild line (neighbor-pbc*.f)
  logical, save :: fcmd_verbose = .true.
  integer(c_int), save :: fcmd_nrebuild = 0
  ! the host received ATOMC since the last upload and may have changed it
  logical, save :: fcmd_hostdirty = .true.
end module fcmd_dropin_state
