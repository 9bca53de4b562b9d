*===============================================================================
*	CLUSTER SIZES + TEN WOLDE-FRENKEL ON THE GPU
*===============================================================================
* file: stoddard-pbc-n-m-fre-hip.f
*-------------------------------------------------------------------------------
* Drop-in for stoddard-pbc-n-m-fre.f:11-261: same subroutine name and
* arguments, computed on the GPU through the C ABI (include/fcmd.h).
*-------------------------------------------------------------------------------

	subroutine stoddard(step)

	use fcmd_dropin_state
	implicit none

	include 'const.inc'
	include 'atomp.inc'
	include 'cvtpp.inc'

	integer step, i
	integer(c_int) :: rc, cnt(6)

	call fcmdensure()
	call fcmdpushdirty()
*** main.f:112 calls stoddard(-1) once initialisation is over
	if (step.lt.0) fcmd_initphase = .false.
	cnt(1) = cnum
	cnt(2) = cnumf
	cnt(3) = cbiggest
	cnt(4) = cbiggestf
	cnt(5) = clookfor
	cnt(6) = clookforf
	rc = fcmd_stoddard(fcmd_sys, 0, step, cl, cs, cc, csi, cek, ceks,
     $	       clf, csf, ccf, csif, cekf, ceksf, cgs, cgsf, cgsfa,
     $	       cfpt, cfptf, cnt)
	call fcmd_stop_on_error(rc)
	cstep = step
	cnum = cnt(1)
	cnumf = cnt(2)
	cbiggest = cnt(3)
	cbiggestf = cnt(4)
	clookfor = cnt(5)
	clookforf = cnt(6)
	do i=1,natom
	  cgsa(i) = 0
	enddo
	return
	end

	subroutine fcmdvarst(v)
	use fcmd_capi
	implicit none
	integer(c_int) :: v
	v = FCMD_ST_FRE
	return
	end
*-------------------------------------------------------------------------------
