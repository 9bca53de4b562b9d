*===============================================================================
* ref_harness.f -- drives the reference's own plugin subroutines on a state
* read from a binary file and dumps what they produce.
*
* Test infrastructure only: it is linked against the reference sources under
* /root/reference/cluster (one plugin set per executable, see oracle/Makefile)
* and writes golden vectors for the C restatement and the HIP kernels.  It is
* never part of the product.
*
* Input (stream, little endian, written by oracle/pyoracle.py _write_input):
*   i4 natom, atsorts, nact
*   i4 api(natom)
*   f8 boxx, boxy, boxz
*   per sort s: f8 apmas(s), apsig(s), apeps(s); i4 apnum(s)
*   f8 atrcut, atrskin, atskin, dt, Q, att
*   i4 nbtim, btim, atnenum, ethst code (1 = 'N', else '-')
*   f8 x,y,z,vx,vy,vz,ax,ay,az (natom each)
*   f8 psi, s
*   f8 dmmh(0:255), dmms, dmmsu, dmmax, dmmav ; i4 dmmst, dmmi
*   nact x (i4 code, i4 arg1, i4 arg2)
*     code 1: neighbor(arg1)            (neighbor-pbc*.f)
*     code 2: a=0, epot=0, AccAtom()    (for-a-lj-n-pbc*.f)
*     code 3: integrator(i), i=arg1..arg2 (int-verlet*.f)
*     code 4: stoddard(arg1)            (stoddard-pbc-n*.f)
* Output: one record per action (see dumpstate below); on stdout, after
* each action, a line 'HARNESS-TIME code seconds' with its wall time.
*===============================================================================

	program refharness

	implicit none

	include 'const.inc'
	include 'atomc.inc'
	include 'atomp.inc'
	include 'boxpp.inc'
	include 'timep.inc'
	include 'energ.inc'
	include 'dmmpp.inc'
	include 'cvtpp.inc'

	character*256 fin, fout
	integer nact, ia, code, arg1, arg2, i, ieth, istep
	integer iacts(3,1000)
	integer*8 ct0, ct1, crate

	call getarg(1, fin)
	call getarg(2, fout)

	open (11, file=fin, access='stream', form='unformatted',
     $        status='old')
	read (11) natom, atsorts, nact
	if (natom.gt.mxa) then
	  write (*,*) 'HARNESS: natom exceeds mxa', natom, mxa
	  stop 2
	endif
	read (11) (api(i), i=1,natom)
	read (11) boxx, boxy, boxz
	do i=1,atsorts
	  read (11) apmas(i), apsig(i), apeps(i), apnum(i)
	enddo
	read (11) atrcut, atrskin, atskin, dt, Q, att
	read (11) nbtim, btim, atnenum, ieth
	ethst = '-'
	if (ieth.eq.1) ethst = 'N'
	read (11) (x(i), i=1,natom), (y(i), i=1,natom), (z(i), i=1,natom)
	read (11) (vx(i), i=1,natom), (vy(i), i=1,natom),
     $            (vz(i), i=1,natom)
	read (11) (ax(i), i=1,natom), (ay(i), i=1,natom),
     $            (az(i), i=1,natom)
	read (11) psi, s
	read (11) (dmmh(i), i=0,255), dmms, dmmsu, dmmax, dmmav
	read (11) dmmst, dmmi
	do ia=1,nact
	  read (11) iacts(1,ia), iacts(2,ia), iacts(3,ia)
	enddo
	close (11)

*** cluster bookkeeping as read-3d.f leaves it
	clookfor = 1
	clookforf = 1
	cbiggest = 0
	cbiggestf = 0
	cstep = -1
	do i=1,natom
	  cfpt(i) = -1
	  cfptf(i) = -1
	enddo
	do i=0,natom
	  csi(i) = 0
	  ceks(i) = 0.0d0
	enddo
	ekin = 0.0d0
	epot = 0.0d0
	sumv2 = 0.0d0

	open (12, file=fout, access='stream', form='unformatted',
     $        status='replace')
	do ia=1,nact
	  code = iacts(1,ia)
	  arg1 = iacts(2,ia)
	  arg2 = iacts(3,ia)
	  istep = arg1
	  call system_clock(ct0, crate)
	  if (code.eq.1) then
	    call neighbor(istep)
	  elseif (code.eq.2) then
	    epot = 0.0d0
	    do i=1,natom
	      ax(i) = 0.0d0
	      ay(i) = 0.0d0
	      az(i) = 0.0d0
	    enddo
	    call AccAtom()
	  elseif (code.eq.3) then
	    do i=arg1,arg2
	      istep = i
	      call integrator(istep)
	    enddo
	  elseif (code.eq.4) then
	    call stoddard(istep)
	  endif
	  call system_clock(ct1)
	  write (*,'(A,I3,1X,F16.6)') 'HARNESS-TIME', code,
     $	    dble(ct1-ct0)/dble(crate)
	  call dumpstate(12, code, istep)
	enddo
	close (12)
	end

*-------------------------------------------------------------------------------
	subroutine dumpstate(u, code, istep)

	implicit none

	include 'const.inc'
	include 'atomc.inc'
	include 'atomp.inc'
	include 'timep.inc'
	include 'energ.inc'
	include 'dmmpp.inc'
	include 'cvtpp.inc'

	integer u, code, istep, i, nlen

	nlen = atnidx(natom)
	write (u) code, istep, natom, nlen
	write (u) (x(i), i=1,natom), (y(i), i=1,natom), (z(i), i=1,natom)
	write (u) (vx(i), i=1,natom), (vy(i), i=1,natom),
     $            (vz(i), i=1,natom)
	write (u) (ax(i), i=1,natom), (ay(i), i=1,natom),
     $            (az(i), i=1,natom)
	write (u) epot, ekin, sumv2, psi, s, atrskin
	write (u) (dmmh(i), i=0,255), dmms, dmmsu, dmmax, dmmav
	write (u) dmmst, dmmi
	write (u) (atnlist(i), i=1,nlen), (atnidx(i), i=1,natom)
	write (u) cnum, cnumf, cbiggest, cbiggestf, clookfor, clookforf
	write (u) (cl(i), i=1,natom), (cs(i), i=1,natom),
     $            (cc(i), i=1,natom), (csi(i), i=0,natom)
	write (u) (cek(i), i=1,natom), (ceks(i), i=0,natom)
	write (u) (clf(i), i=1,natom), (csf(i), i=0,natom),
     $            (ccf(i), i=1,natom), (csif(i), i=0,natom)
	write (u) (cekf(i), i=0,natom), (ceksf(i), i=0,natom)
	write (u) (cgs(i), i=1,natom), (cgsf(i), i=0,natom),
     $            (cgsfa(i), i=0,natom)
	write (u) (cfpt(i), i=1,natom), (cfptf(i), i=1,natom)
	return
	end
