*===============================================================================
*	NEIGHBOR LIST - PBC, ADAPTIVE SKIN, ON THE GPU
*===============================================================================
* file: neighbor-pbc-mav-hip.f
*-------------------------------------------------------------------------------
* Drop-in for neighbor-pbc-mav.f:13-109: same subroutine name and
* arguments, computed on the GPU through the C ABI (include/fcmd.h).
*-------------------------------------------------------------------------------

	subroutine neighbor(step)

	use fcmd_dropin_state
	implicit none

	integer step
	integer(c_int) :: st, rc
	type(fcmd_scalars) :: sc

	call fcmdensure()
	call fcmdpushdirty()
	st = step
	rc = fcmd_neighbor(fcmd_sys, 0, st)
	call fcmd_stop_on_error(rc)
	step = st
*** wrapped positions are the host's positions from here on, as in the
*** reference (positions are wrapped in place at every rebuild)
	call fcmdpull()
	call fcmdpullscal(sc)
	if (fcmd_verbose) call fcmdlog(step, sc)
	fcmd_nrebuild = sc%nrebuild
	return
	end

	subroutine fcmdvarnb(v, verb)
	use fcmd_capi
	implicit none
	integer(c_int) :: v
	logical verb
	v = FCMD_NB_MAV
	verb = .true.
	return
	end
*-------------------------------------------------------------------------------
