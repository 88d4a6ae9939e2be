c     moil_dump -- test-infrastructure driver for the MOIL reference.
c
c     This file is NOT part of the product.  It is linked against MOIL's
c     own, unmodified Fortran objects (compiled from the reference tree
c     by oracle/ref/Makefile into oracle/_ref/) and drives the energy
c     path the same way program pote does (pot/energy.f:375-437):
c       init_ef -> rconn -> coordinates -> cutoffs squared
c       (pot/energy.f:382-400) -> ewald_init -> [squeeze] -> nbondm ->
c       syminit -> eforce
c     It then writes every hot-path output in full precision (pote's
c     own report, pot/wener.f, prints f12.3 only).
c
c     Coordinates are read as raw float64 (3,npt) instead of a CHARMM
c     CRD file so that no f10.5 quantisation is applied (a path ending in
c     .crd is read by MOIL's getcrd instead); the topology is read
c     through MOIL's own rconn from a .con text file.
c
c     Usage: moil_dump <param-file>
c     param-file (list-directed, one item group per line):
c       con file path
c       coordinate file path (raw float64, 3*npt values, atom-major)
c       a b c                          (box edges, Angstrom)
c       rvmx relx rvbg rebg cutm       (Angstrom, squared here)
c       dtol nfft1 nfft2 nfft3 iord
c       eps
c       dosqz neval nlist sigma seed dodump
c       output file path
c     neval>1 runs a timing loop: before every evaluation after the
c     first the coordinates get N(0,sigma) noise, and squeeze/nbondm/
c     syminit run every nlist evaluations (list build included in the
c     wall time).  The dump is written after the last evaluation.
c
      program moil_dump
      implicit none
      include 'COMMON/LENGTH.BLOCK'
      include 'COMMON/CONNECT.BLOCK'
      include 'COMMON/UNITS.BLOCK'
      include 'COMMON/DEBUG.BLOCK'
      include 'COMMON/LINE.BLOCK'
      include 'COMMON/NBLIST.BLOCK'
      include 'COMMON/SYMM.BLOCK'
      include 'COMMON/COORD.BLOCK'
      include 'COMMON/ENERGY.BLOCK'
      include 'COMMON/FREEZ.BLOCK'
      include 'COMMON/PARALLEL.BLOCK'
      include 'COMMON/CONVERT.BLOCK'
      include 'COMMON/EWALD.BLOCK'
      include 'COMMON/PRESS.BLOCK'

      character*512 pfile,confile,crdfile,outfile
      double precision rvmx,relx,rvbg,rebg,cutm,sigma
      integer iord,dosqz,neval,nlist,iseed,dodump,ihvdw,imode
      integer i,k,ieval,ucon,ucrd,upar
      integer*8 c0,c1,crate,tlist,tforce
      double precision g1,g2
      integer*8 rstate
      double precision rnd_normal
      external rnd_normal

      stdi = 5
      stdo = 6
      stderr = 0
      debug = .false.
      totmon = 0
      npt = 0
      nb = 0
      nmb = 0
      nangl = 0
      ntors = 0
      nimp = 0
      lestyp = 0
      my_pe = 0
      num_pes = 1
      prll_on_off = .false.
      jnkf = 25
      open (unit=jnkf,status='scratch')

      call get_command_argument(1,pfile)
      upar = 31
      open(unit=upar,file=pfile,status='old')
      read(upar,'(a)') confile
      read(upar,'(a)') crdfile
      call init_ef()
      call init_var()
      read(upar,*) a,b,c
      read(upar,*) rvmx,relx,rvbg,rebg,cutm
      read(upar,*) dtol,nfft1,nfft2,nfft3,iord
      read(upar,*) eps
      read(upar,*) dosqz,neval,nlist,sigma,iseed,dodump
      read(upar,'(a)') outfile
c     optional last line: 0 switches the hydrogen rule off, as the 'hvdw'
c     keyword of pot/energy.f:219 does (init_ef sets hvdw0 = .true.)
      ihvdw = 1
      imode = 0
      read(upar,*,end=15) ihvdw
c     optional line after it: a MOIL mode to switch on (for the test that
c     the GPU shim refuses modes it does not evaluate): 1 distance-
c     dependent dielectric, 2 pressure, 3 one frozen particle, 4 arith;
c     5 is the pressure calculation as MOIL's pressure program runs it
c     (analysis/pressure.f:77-196): molecules from init_pressure, their
c     centres of mass before every eforce, the virial terms dumped
      read(upar,*,end=15) imode
 15   if (ihvdw.eq.0) hvdw0 = .false.
      close(upar)

      esymyes = .true.
      ewaldyes = .true.
      intrpord = iord
      sgridx = 1.d0
      sgridy = 1.d0
      sgridz = 1.d0

      ucon = 32
      open(unit=ucon,file=confile,status='old')
      call rconn(ucon)
      close(ucon)
      inofrz = npt
      npt_par = npt
      do 10 i=1,npt
         zerofrz(i) = 1
         prtc_pointer(i) = i
 10   continue
      if (imode.eq.1) ctrue = .false.
      if (imode.eq.2) pressON = .true.
      if (imode.eq.3) then
         inofrz = npt - 1
         zerofrz(npt) = 0
      end if
      if (imode.eq.4) arith = .true.
      if (imode.eq.5) then
         pressON = .true.
         call init_pressure()
      end if

      ucrd = 33
      k = len_trim(crdfile)
      if (crdfile(max(k-3,1):k) .eq. '.crd') then
c        a CHARMM-style coordinate file, read by MOIL's own getcrd
         open(unit=ucrd,file=crdfile,status='old')
         call getcrd(ucrd,'CHARM')
      else
         open(unit=ucrd,file=crdfile,status='old',access='stream',
     1        form='unformatted')
         read(ucrd) ((coor(k,i),k=1,3),i=1,npt)
      end if
      close(ucrd)

c     cutoffs exactly as pot/energy.f:382-400 (cutm always explicit)
      cutvdw2  = rvmx*rvmx
      cutvbig2 = rvbg*rvbg
      cutele2  = relx*relx
      cutebig2 = rebg*rebg
      cutmono2 = cutm*cutm

      call ewald_init()

      rstate = iseed
      tlist = 0
      tforce = 0
      call system_clock(count_rate=crate)
      do 100 ieval=1,max(neval,1)
         if (ieval.gt.1 .and. sigma.gt.0.d0) then
            do 20 i=1,npt
               do 20 k=1,3
                  coor(k,i) = coor(k,i) + sigma*rnd_normal(rstate)
 20         continue
         end if
         call system_clock(c0)
         if (mod(ieval-1,max(nlist,1)).eq.0) then
            if (dosqz.ne.0) call squeeze()
            call nbondm()
            call syminit()
         end if
         call system_clock(c1)
         tlist = tlist + (c1-c0)
         if (imode.eq.5) call mol_coords()
         call eforce()
         call system_clock(c0)
         tforce = tforce + (c0-c1)
 100  continue
      g1 = dble(tlist)/dble(crate)
      g2 = dble(tforce)/dble(crate)
      write(stdo,'(a,i8,a,es16.8,a,es16.8)') 'MOILTIME neval=',
     1     max(neval,1),' tlist_s=',g1,' tforce_s=',g2

      call dump_all(outfile,dodump)
      stop
      end

c     Molecular centres of mass and atom positions relative to them, as
c     the coordinate part of atom2molecule (analysis/pressure.f:419-505):
c     per molecule and component, the mass-weighted sum over its atoms in
c     ordLISTatm order, divided by the molecule's mass.
      subroutine mol_coords()
      implicit none
      include 'COMMON/LENGTH.BLOCK'
      include 'COMMON/CONNECT.BLOCK'
      include 'COMMON/COORD.BLOCK'
      include 'COMMON/PRESS.BLOCK'
      integer imol,k,k0,i,kd
      double precision sm
      k0 = 0
      do 20 imol=1,nmol
         do 10 kd=1,3
            sm = 0.d0
            do 5 k=k0+1,patom(imol)
               sm = sm + coor(kd,ordLISTatm(k))*ptms(ordLISTatm(k))
 5          continue
            sm = sm/massmol(imol)
            if (kd.eq.1) xmol(imol) = sm
            if (kd.eq.2) ymol(imol) = sm
            if (kd.eq.3) zmol(imol) = sm
 10      continue
         k0 = patom(imol)
 20   continue
      do 30 i=1,npt
         xrel(i) = coor(1,i) - xmol(pmol(i))
         yrel(i) = coor(2,i) - ymol(pmol(i))
         zrel(i) = coor(3,i) - zmol(pmol(i))
 30   continue
      return
      end

c     Marsaglia polar normal deviate from a 64-bit LCG; only used to
c     perturb coordinates in the timing loop.
      double precision function rnd_normal(s)
      implicit none
      integer*8 s
      double precision u,v,q
      integer*8 m
 1    continue
      s = s*6364136223846793005_8 + 1442695040888963407_8
      m = ishft(s,-11)
      u = 2.d0*dble(m)*1.1102230246251565d-16 - 1.d0
      s = s*6364136223846793005_8 + 1442695040888963407_8
      m = ishft(s,-11)
      v = 2.d0*dble(m)*1.1102230246251565d-16 - 1.d0
      q = u*u + v*v
      if (q.ge.1.d0 .or. q.eq.0.d0) goto 1
      rnd_normal = u*sqrt(-2.d0*log(q)/q)
      return
      end

      subroutine wrd(u,nm,n,arr)
      implicit none
      integer u,n,i
      character*8 nm
      double precision arr(*)
      write(u) nm,1,n,(arr(i),i=1,n)
      return
      end

      subroutine wri(u,nm,n,arr)
      implicit none
      integer u,n,i
      character*8 nm
      integer arr(*)
      write(u) nm,2,n,(arr(i),i=1,n)
      return
      end

      subroutine dump_all(outfile,dodump)
      implicit none
      include 'COMMON/LENGTH.BLOCK'
      include 'COMMON/CONNECT.BLOCK'
      include 'COMMON/NBLIST.BLOCK'
      include 'COMMON/SYMM.BLOCK'
      include 'COMMON/COORD.BLOCK'
      include 'COMMON/ENERGY.BLOCK'
      include 'COMMON/EWALD.BLOCK'
      include 'COMMON/PRESS.BLOCK'
      character*(*) outfile
      integer dodump
      integer u,n,ntab
      double precision en(20),vr(16)
      integer iv(40)

      u = 41
      open(unit=u,file=outfile,status='replace',access='stream',
     1     form='unformatted')
      en(1) = ewaldcof
      en(2) = e_self
      en(3) = e_corr
      en(4) = e_ew_receip
      en(5) = e_vdw
      en(6) = e_el
      en(7) = e_vsym
      en(8) = e_lsym
      en(9) = e_bond
      en(10) = e_theta
      en(11) = e_tors
      en(12) = e_imp
      en(13) = e_el14
      en(14) = e_vdw14
      en(15) = e_total
      en(16) = e_dir
      en(17) = erftbdns
      en(18) = fpme
      call wrd(u,'ENERGIES',18,en)
      call wrd(u,'DPOT    ',3*npt,dpot)
      call wrd(u,'COOR    ',3*npt,coor)
      ntab = int(ewaldcof*erftbdns*sqrt(cutele2)*1.5d0)
      call wrd(u,'ERFTAB  ',4*ntab,erf_arr)
      call wrd(u,'BSPMOD1 ',nfft1,bsp_mod1)
      call wrd(u,'BSPMOD2 ',nfft2,bsp_mod2)
      call wrd(u,'BSPMOD3 ',nfft3,bsp_mod3)
      iv(1) = npt
      iv(2) = nwaters
      iv(3) = totdmon
      iv(4) = point1(npt)
      iv(5) = point2(npt)
      iv(6) = point3(npt)
      iv(7) = nlist1
      iv(8) = nlist2
      iv(9) = nlist3
      iv(10) = indxsym1
      iv(11) = indxsym2
      iv(12) = indxsym3
      iv(13) = poinwat1(nwaters)
      iv(14) = poinwat2(nwaters)
      iv(15) = nwlist1
      iv(16) = nwlist2
      iv(17) = ipwat1
      iv(18) = ipwat2
      iv(19) = symnum
      iv(20) = ntab
      call wri(u,'COUNTS  ',20,iv)
      call wrd(u,'MONOCENT',3*totdmon,mono_cent)
      if (pressON .and. nmol.gt.0) then
c        the virial terms of eforce (pot/eforce.f:70-82 zeroes them) and
c        the molecule tables of init_pressure
         vr(1) = virial
         vr(2) = virXX
         vr(3) = virYY
         vr(4) = virZZ
         vr(5) = V_PIdirXX
         vr(6) = V_PIdirYY
         vr(7) = V_PIdirZZ
         vr(8) = V_PIrecXX
         vr(9) = V_PIrecYY
         vr(10) = V_PIrecZZ
         vr(11) = V_PIrelXX
         vr(12) = V_PIrelYY
         vr(13) = V_PIrelZZ
         call wrd(u,'VIRIAL  ',13,vr)
         iv(1) = nmol
         call wri(u,'NMOL    ',1,iv)
         call wri(u,'PATOM   ',nmol,patom)
         call wri(u,'ORDLIST ',npt,ordLISTatm)
         call wri(u,'PMOL    ',npt,pmol)
         call wrd(u,'MASSMOL ',nmol,massmol)
         call wrd(u,'PTMS    ',npt,ptms)
      end if
      if (dodump.ne.0) then
         call wri(u,'POINT1  ',npt+1,point1)
         call wri(u,'POINT2  ',npt+1,point2)
         call wri(u,'POINT3  ',npt+1,point3)
         if (nlist1.gt.0) call wri(u,'LIST1   ',nlist1,list1)
         if (nlist2.gt.0) call wri(u,'LIST2   ',nlist2,list2)
         if (nlist3.gt.0) call wri(u,'LIST3   ',nlist3,list3)
         call wri(u,'IBLOCK1 ',symnum,iblock1)
         call wri(u,'IBLOCK2 ',symnum,iblock2)
         call wri(u,'IBLOCK3 ',symnum,iblock3)
         if (indxsym1.gt.0) then
            call wri(u,'SYMREAL1',indxsym1,symreal1)
            call wri(u,'PSYM1   ',indxsym1,psym1)
         end if
         if (indxsym2.gt.0) then
            call wri(u,'SYMREAL2',indxsym2,symreal2)
            call wri(u,'PSYM2   ',indxsym2,psym2)
         end if
         if (indxsym3.gt.0) then
            call wri(u,'SYMREAL3',indxsym3,symreal3)
            call wri(u,'PSYM3   ',indxsym3,psym3)
         end if
         if (nwaters.gt.0) then
            call wri(u,'POINWAT1',nwaters+1,poinwat1)
            call wri(u,'POINWAT2',nwaters+1,poinwat2)
            call wri(u,'INDXWAT ',nwaters,indxwat)
         end if
         if (nwlist1.gt.0) call wri(u,'LISTWT1 ',nwlist1,listwt1)
         if (nwlist2.gt.0) call wri(u,'LISTWT2 ',nwlist2,listwt2)
         n = symnum + 1
         call wri(u,'IBLCKWT1',n,iblckwt1)
         call wri(u,'IBLCKWT2',n,iblckwt2)
         if (ipwat1.gt.0) then
            call wri(u,'SYMMWAT1',ipwat1,symmwat1)
            call wri(u,'PSYMWT1 ',ipwat1,psymwt1(1))
         end if
         if (ipwat2.gt.0) then
            call wri(u,'SYMMWAT2',ipwat2,symmwat2)
            call wri(u,'PSYMWT2 ',ipwat2,psymwt2(1))
         end if
         call wri(u,'DPOIPT  ',totdmon+1,dpoipt)
      end if
      close(u)
      return
      end
