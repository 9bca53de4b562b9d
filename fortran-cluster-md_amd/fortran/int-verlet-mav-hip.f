*===============================================================================
*	INTEGRATOR VERLET + MOVING-AVERAGE LIST TRIGGER ON THE GPU
*===============================================================================
* file: int-verlet-mav-hip.f
*-------------------------------------------------------------------------------
* Drop-in for int-verlet-mav.f:18-95: same subroutine name and
* arguments, computed on the GPU through the C ABI (include/fcmd.h).
*-------------------------------------------------------------------------------

	subroutine integrator(step)

	use fcmd_dropin_state
	implicit none

	integer step
	integer(c_int) :: st, rc

	call fcmdensure()
	call fcmdpushdirty()
	st = step
	rc = fcmd_integrator(fcmd_sys, st)
	call fcmd_stop_on_error(rc)
	call fcmdafter(step)
	return
	end

	subroutine fcmdvarint(v)
	use fcmd_capi
	implicit none
	integer(c_int) :: v
	v = FCMD_INT_MAV
	return
	end
*-------------------------------------------------------------------------------
