! fcmd_dropin_state.f90 -- the device handle shared by the drop-in plugin
! subroutines.  Names carry an fcmd_ prefix so they cannot collide with the
! host program's COMMON variables (ATOMP has s, Q, psi ...).
module fcmd_dropin_state
  use fcmd_capi
  implicit none
  type(c_ptr), save :: fcmd_sys = c_null_ptr
  ! FCMD_RESIDENT=1: keep ATOMC on the device between the driver's output
  ! steps (main.f:117-168 schedule); otherwise copy both ways on every call
  logical, save :: fcmd_resident = .false.
  ! the linked neighbor plugin prints the per-rebuild line (neighbor-pbc*.f)
  logical, save :: fcmd_verbose = .true.
  integer(c_int), save :: fcmd_nrebuild = 0
  ! the host received ATOMC since the last upload and may have changed it
  logical, save :: fcmd_hostdirty = .true.
end module fcmd_dropin_state
