c     moil_dyna.f -- test driver: MOIL's molecular dynamics step loop on
c     MOIL's own routines, for the parity test of the library's resident
c     integrator.  Not part of the product.
c
c     The loop follows dynamics/dyna.f:105-301 with the options the
c     library implements: every nlist steps squeeze, nbondm, syminit and
c     one eforce; then per step the position step
c     velo <- velo dt - factor1 dpot, matrix SHAKE of the waters on it
c     (mshakpt), coor <- coor + velo, velo <- velo / dt, eforce,
c     velo <- velo - factor2 dpot and the velocity SHAKE (mshakvl), with
c     factor1 = dt^2/2 / m and factor2 = dt/2 / m (dynamics/init_dyna.f:
c     279-288) and dt in MOIL's time unit (ps / tconv).  No temperature
c     control, no frozen particles, no bond SHAKE.  The matrix SHAKE
c     constants come from MOIL's mshakinit (dynamics/mshakinit.f).
c
c     Parameter file (argument 1), free format:
c       con file / coordinate file (binary coor(3,npt)) / a b c /
c       rvmx relx rvbg rebg cutm / dtol nfft1 nfft2 nfft3 order / eps /
c       nstep nlist dt_ps mshake tolcons / velocity file (binary
c       velo(3,npt), A per MOIL time unit) / dump file
c     The dump holds the final COOR, VELO, DPOT, ENERGIES (as moil_dump),
c     ETOT (e_total after each step's eforce) and MSHAKE (cmat(3,3),
c     tolcons, reqoh2, reqhh2, invmh, invmo).
      program moil_dyna
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
      include 'COMMON/VELOC.BLOCK'
      include 'COMMON/MSHAKE.BLOCK'

      character*512 pfile,confile,crdfile,velfile,outfile
      double precision rvmx,relx,rvbg,rebg,cutm,dtps,dt,dt2,twodt,tmp
      integer iord,nstep,nlist,ishm,istep,j,i,k,ucon,ucrd,upar,uout
      logical shkm
      double precision, allocatable :: f1(:),f2(:),etot(:),msh(:)

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
      read(upar,*) nstep,nlist,dtps,ishm,tolcons
      read(upar,'(a)') velfile
      read(upar,'(a)') outfile
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
         nofreez(i) = i
         prtc_pointer(i) = i
 10   continue

      ucrd = 33
      open(unit=ucrd,file=crdfile,status='old',access='stream',
     1     form='unformatted')
      read(ucrd) ((coor(k,i),k=1,3),i=1,npt)
      close(ucrd)
      open(unit=ucrd,file=velfile,status='old',access='stream',
     1     form='unformatted')
      read(ucrd) ((velo(k,i),k=1,3),i=1,npt)
      close(ucrd)

      cutvdw2  = rvmx*rvmx
      cutvbig2 = rvbg*rvbg
      cutele2  = relx*relx
      cutebig2 = rebg*rebg
      cutmono2 = cutm*cutm

      call ewald_init()
      shkm = ishm.ne.0
      if (shkm) call mshakinit(debug)

c     multiplicative constants in MOIL's time unit
      dt    = dtps/tconv
      twodt = dt*0.5d0
      dt2   = dt*dt*0.5d0
      allocate(f1(npt),f2(npt),etot(max(nstep,1)),msh(14))
      do 20 i=1,npt
         f1(i) = dt2*(1.d0/ptms(i))
         f2(i) = twodt*(1.d0/ptms(i))
 20   continue

      do 60 istep=1,nstep,nlist
         call squeeze()
         call nbondm()
         call syminit()
         call eforce()
         do 50 j=istep,min(nstep,istep+nlist-1)
            do 30 i=1,npt
               velo(1,i) = velo(1,i)*dt - f1(i)*dpot(1,i)
               velo(2,i) = velo(2,i)*dt - f1(i)*dpot(2,i)
               velo(3,i) = velo(3,i)*dt - f1(i)*dpot(3,i)
 30         continue
            if (shkm) call mshakpt(dt,dt2)
            tmp = 1.d0/dt
            do 35 i=1,npt
               coor(1,i) = coor(1,i) + velo(1,i)
               coor(2,i) = coor(2,i) + velo(2,i)
               coor(3,i) = coor(3,i) + velo(3,i)
               velo(1,i) = velo(1,i)*tmp
               velo(2,i) = velo(2,i)*tmp
               velo(3,i) = velo(3,i)*tmp
 35         continue
            call eforce()
            do 40 i=1,npt
               velo(1,i) = velo(1,i) - f2(i)*dpot(1,i)
               velo(2,i) = velo(2,i) - f2(i)*dpot(2,i)
               velo(3,i) = velo(3,i) - f2(i)*dpot(3,i)
 40         continue
            if (shkm) call mshakvl(dt)
            etot(j) = e_total
 50      continue
 60   continue

      do 70 j=1,3
         do 70 i=1,3
            msh(3*(j-1)+i) = cmat(i,j)
 70   continue
      msh(10) = tolcons
      msh(11) = reqoh2
      msh(12) = reqhh2
      msh(13) = invmh
      msh(14) = invmo

      uout = 34
      open(unit=uout,file=outfile,status='replace',access='stream',
     1     form='unformatted')
      call dwr(uout,'COOR    ',3*npt,coor)
      call dwr(uout,'VELO    ',3*npt,velo)
      call dwr(uout,'DPOT    ',3*npt,dpot)
      call dwr(uout,'ETOT    ',nstep,etot)
      call dwr(uout,'MSHAKE  ',14,msh)
      call dwr(uout,'ENERGIES',1,e_total)
      close(uout)
      stop
      end

      subroutine dwr(u,nm,n,arr)
      implicit none
      integer u,n,i
      character*8 nm
      double precision arr(*)
      write(u) nm,1,n,(arr(i),i=1,n)
      return
      end
