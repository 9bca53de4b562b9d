*===============================================================================
*	INITIAL CONFIGURATION: SEEDED RANDOM VAPOUR
*===============================================================================
* file: init-random-hip.f
*-------------------------------------------------------------------------------
* Drop-in alongside init-cube1.f:33-128 (same subroutine, no arguments, the
* same COMMON results): natom atoms placed uniformly at random in the
* periodic box with no pair closer than 0.9 sigma of species 1,
* Maxwell-Boltzmann velocities at att for each species' mass, zero total
* momentum and kinetic energy exactly 3/2 N k att (fcmd_init_random of the
* C ABI: SplitMix64-seeded xoshiro256**, Box-Muller); the species stay as
* readdata assigned them (api, species 1 first).  Seed: the environment
* variable FCMD_SEED if set, else 65539 with rndctrl = '-' (the reference's
* fixed seed, kps.f:30) or the time of day with rndctrl = 'X'.  Then, as
* init-cube1.f:113-116 does: neighbor(0), AccAtom(), AccWall().  No warm-up
* steps: the seeded vapour is the initial state.
*-------------------------------------------------------------------------------

	subroutine initconf()

	use fcmd_capi
	implicit none

	include 'const.inc'
	include 'atomc.inc'
	include 'atomp.inc'
	include 'boxpp.inc'
	include 'energ.inc'

	integer(c_int) :: rc
	integer(c_long_long) :: seed
	real(c_double) :: box(3), rmin
	character*32 env
	integer i, istat, tv(8)

	seed = 65539_c_long_long
	if (rndctrl.eq.'X') then
	  call date_and_time(values=tv)
	  seed = int(tv(8), c_long_long) + 1000_c_long_long*(tv(7) +
     $	         60*(tv(6) + 60*(tv(5) + 24*tv(3))))
	endif
	env = ' '
	call get_environment_variable('FCMD_SEED', env, status=istat)
	if (istat.eq.0) read (env,*) seed
	write (*,'(A,I20)') ' seeded random vapour, seed ', seed

	box(1) = boxx
	box(2) = boxy
	box(3) = boxz
	rmin = 0.9d0*apsig(1)
	rc = fcmd_init_random(seed, natom, box, rmin, att, apmas, api,
     $	                      x, y, z, vx, vy, vz)
	call fcmd_stop_on_error(rc)
	do i=1,natom
	  ax(i) = 0.0d0
	  ay(i) = 0.0d0
	  az(i) = 0.0d0
	enddo
	ekin = 0.0d0
	do i=1,natom
	  ekin = ekin+0.5d0*apmas(api(i))*(vx(i)*vx(i)+vy(i)*vy(i)+
     $	         vz(i)*vz(i))
	enddo
	epot = 0.0d0

	call neighbor(0)
	call AccAtom()
	call AccWall()

	return
	end
*-------------------------------------------------------------------------------
