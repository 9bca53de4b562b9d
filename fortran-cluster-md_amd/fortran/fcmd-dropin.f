*===============================================================================
*	DEVICE RESIDENCY FOR THE DROP-IN PLUGIN SUBROUTINES
*===============================================================================
* file: fcmd-dropin.f
*-------------------------------------------------------------------------------
* Compiled against the host program's own include files, so the COMMON layout
* (and mxa, mxnlist) is whatever the host program was built with.
*
* fcmdensure   creates the device system on first use from the COMMON
*              values and the plugin variants linked in (fcmdvar* below are
*              provided by the plugin files), then uploads the state
* fcmdpushdirty uploads ATOMC and the scalars if the host may have changed
*              them since the last upload (i.e. after any download)
* fcmdpull     downloads ATOMC and the scalars
* fcmdafter    after a device call: scalars always; ATOMC when the driver
*              will read it before the next device call
*-------------------------------------------------------------------------------

	subroutine fcmdensure()

	use fcmd_dropin_state
	implicit none

	include 'const.inc'
	include 'atomc.inc'
	include 'atomp.inc'
	include 'boxpp.inc'
	include 'timep.inc'
	include 'energ.inc'

	type(fcmd_params) :: prm
	integer(c_int) :: rc, iv, nv, av, sv, prec
	logical verb
	integer k
	character*16 env

	if (c_associated(fcmd_sys)) return

	call fcmdvarint(iv)
	call fcmdvarnb(nv, verb)
	call fcmdvaracc(av)
	call fcmdvarst(sv)
	fcmd_verbose = verb

	prm%natom = natom
	prm%nsorts = atsorts
	prm%box(1) = boxx
	prm%box(2) = boxy
	prm%box(3) = boxz
	do k=1,FCMD_MAXSORTS
	  prm%mas(k) = 0.0d0
	  prm%sig(k) = 0.0d0
	  prm%eps(k) = 0.0d0
	  prm%num(k) = 0
	enddo
	do k=1,atsorts
	  prm%mas(k) = apmas(k)
	  prm%sig(k) = apsig(k)
	  prm%eps(k) = apeps(k)
	  prm%num(k) = apnum(k)
	enddo
	prm%atrcut = atrcut
	prm%atrskin = atrskin
	prm%atskin = atskin
	prm%dt = dt
	prm%Q = Q
	prm%att = att
	prm%nbtim = nbtim
	prm%btim = btim
	prm%atnenum = atnenum
	prm%ethst_nh = 0
	if (ethst.eq.'N') prm%ethst_nh = 1
	prm%integrator = iv
	prm%neighbor = nv
	prm%accatom = av
	prm%stoddard = sv
	do k=1,FCMD_MAXSORTS*FCMD_MAXSORTS
	  prm%eps_ab(k) = 0.0d0
	  prm%sig_ab(k) = 0.0d0
	enddo
	prm%nh_mask = 0
	prm%pad = 0

	prec = FCMD_FP64
	env = ' '
	call get_environment_variable('FCMD_PRECISION', env)
	if (env(1:5).eq.'mixed') prec = FCMD_MIXED
	if (env(1:5).eq.'kahan') prec = FCMD_MIXED_KAHAN
	env = ' '
	call get_environment_variable('FCMD_RESIDENT', env)
	fcmd_resident = env(1:1).ne.'0'

	rc = fcmd_create(fcmd_sys, 1, prm, prec, 0)
	call fcmd_stop_on_error(rc)
	fcmd_hostdirty = .true.
	call fcmdpushdirty()
	return
	end

*-------------------------------------------------------------------------------
	subroutine fcmdpushdirty()

	use fcmd_dropin_state
	implicit none

	include 'const.inc'
	include 'atomc.inc'
	include 'atomp.inc'
	include 'energ.inc'
	include 'dmmpp.inc'

	type(fcmd_scalars) :: sc
	integer(c_int) :: rc
	integer k

	if (.not.fcmd_hostdirty) return
	rc = fcmd_upload(fcmd_sys, 0, x, y, z, vx, vy, vz, ax, ay, az, api)
	call fcmd_stop_on_error(rc)
	rc = fcmd_get_scalars(fcmd_sys, 0, sc)
	call fcmd_stop_on_error(rc)
	sc%atrskin = atrskin
	sc%psi = psi
	sc%s = s
	sc%ekin = ekin
	sc%epot = epot
	sc%sumv2 = sumv2
	do k=0,255
	  sc%dmmh(k+1) = dmmh(k)
	enddo
	sc%dmms = dmms
	sc%dmmsu = dmmsu
	sc%dmmax = dmmax
	sc%dmmav = dmmav
	sc%dmmst = dmmst
	sc%dmmi = dmmi
	rc = fcmd_set_scalars(fcmd_sys, 0, sc)
	call fcmd_stop_on_error(rc)
	fcmd_hostdirty = .false.
	return
	end

*-------------------------------------------------------------------------------
	subroutine fcmdpullscal(sc)

	use fcmd_dropin_state
	implicit none

	include 'const.inc'
	include 'atomp.inc'
	include 'energ.inc'
	include 'dmmpp.inc'

	type(fcmd_scalars) :: sc
	integer(c_int) :: rc
	integer k

	rc = fcmd_get_scalars(fcmd_sys, 0, sc)
	call fcmd_stop_on_error(rc)
	atrskin = sc%atrskin
	psi = sc%psi
	s = sc%s
	ekin = sc%ekin
	epot = sc%epot
	sumv2 = sc%sumv2
	do k=0,255
	  dmmh(k) = sc%dmmh(k+1)
	enddo
	dmms = sc%dmms
	dmmsu = sc%dmmsu
	dmmax = sc%dmmax
	dmmav = sc%dmmav
	dmmst = sc%dmmst
	dmmi = sc%dmmi
	return
	end

*-------------------------------------------------------------------------------
	subroutine fcmdpull()

	use fcmd_dropin_state
	implicit none

	include 'const.inc'
	include 'atomc.inc'

	type(fcmd_scalars) :: sc
	integer(c_int) :: rc

	rc = fcmd_download(fcmd_sys, 0, x, y, z, vx, vy, vz, ax, ay, az)
	call fcmd_stop_on_error(rc)
	call fcmdpullscal(sc)
*** the host now holds a copy it may change: upload before the next call
	fcmd_hostdirty = .true.
	return
	end

*-------------------------------------------------------------------------------
* After integrator(step): the per-rebuild line the reference's neighbor
* prints (neighbor-pbc-mav.f:102-106), then the ATOMC download if the driver
* reads ATOMC before the next integrator call (main.f:119-162 runs its
* writers, thermostat, resize and particle changes at step+1), or while
* the driver is still initialising (fcmd_initphase: warm-up steps such as
* init-cube1.f:119-125 are followed by host code that reads ATOMC).
	subroutine fcmdafter(step)

	use fcmd_dropin_state
	implicit none

	include 'const.inc'
	include 'atomp.inc'
	include 'filep.inc'
	include 'timep.inc'
	include 'energ.inc'
	include 'boxpp.inc'

	integer step, nxt
	type(fcmd_scalars) :: sc
	logical need

	need = .not.fcmd_resident
	nxt = step + 1
	if (ftxyz.ne.0) then
	  if (mod(nxt,ftxyz).eq.0) need = .true.
	endif
	if (ftvel.ne.0) then
	  if (mod(nxt,ftvel).eq.0) need = .true.
	endif
	if (ftacc.ne.0) then
	  if (mod(nxt,ftacc).eq.0) need = .true.
	endif
	if (ftrst.ne.0) then
	  if (mod(nxt,ftrst).eq.0) need = .true.
	endif
	if (ftrat.ne.0) need = .true.
	if (ethst.eq.'X') need = .true.
	if (bshr.eq.'X') need = .true.
	if (chgNctl.eq.'X') need = .true.
	if (fcmd_initphase) need = .true.

	if (need) then
	  call fcmdpull()
	endif
	call fcmdpullscal(sc)
	if (sc%nrebuild.gt.fcmd_nrebuild) then
	  if (fcmd_verbose) call fcmdlog(step, sc)
	  fcmd_nrebuild = sc%nrebuild
	endif
	return
	end

*-------------------------------------------------------------------------------
* the rebuild line of neighbor-pbc.f:97-101 / neighbor-pbc-mav.f:102-106
	subroutine fcmdlog(step, sc)

	use fcmd_dropin_state
	implicit none

	include 'const.inc'
	include 'atomp.inc'

	integer step
	type(fcmd_scalars) :: sc

	write (*,10) step,sc%npairs,sc%dmmsr,sc%dmmav,sc%dmmax,
     $	             sc%atrskin-atrcut, sc%atrskin/atr
10	format (2(I10,2X),5(F8.4,2X))
	return
	end
*-------------------------------------------------------------------------------
