*===============================================================================
*	ACCELERATION ATOM-ATOM (LJ, two species) ON THE GPU
*===============================================================================
* file: for-a-lj-n-pbc-m-hip.f
*-------------------------------------------------------------------------------
* Drop-in for for-a-lj-n-pbc-m.f:18-154: same subroutine name and
* arguments, computed on the GPU through the C ABI (include/fcmd.h).
*-------------------------------------------------------------------------------

	subroutine AccAtom()

	use fcmd_dropin_state
	implicit none

	integer(c_int) :: rc

	call fcmdensure()
	call fcmdpushdirty()
	rc = fcmd_accatom(fcmd_sys)
	call fcmd_stop_on_error(rc)
	call fcmdpull()
	return
	end

	subroutine fcmdvaracc(v)
	use fcmd_capi
	implicit none
	integer(c_int) :: v
	v = FCMD_ACC_M
	return
	end
*-------------------------------------------------------------------------------
