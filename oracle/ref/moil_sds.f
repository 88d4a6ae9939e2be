c     moil_sds -- test-infrastructure driver for MOIL's chain functional.
c
c     This file is NOT part of the product.  It is linked against MOIL's
c     own, unmodified Fortran objects (oracle/ref/Makefile, target sds)
c     including chain/sds.f, and evaluates the energy of a chain of
c     igrid replicas of one system the way MOIL's chain minimiser and
c     annealer do (chain/chmin.f:134-145 defaults, chain/chnneal.f:85):
c       init_ef -> rconn -> cutoffs squared -> ewald_init ->
c       sds(...)  (squeeze/nbondm/syminit/eforce per interior replica,
c                  then the nearest-neighbour spring and next-nearest
c                  repulsion of the chain, chain/sds.f:71-257)
c     It then recomputes every interior replica's own energy and
c     gradient (the same calls sds makes) so the chain terms can be
c     checked apart from eforce, and writes all outputs in full
c     precision.
c
c     Usage: moil_sds <param-file>
c     param-file (list-directed, one item group per line):
c       con file path
c       replica coordinate file (raw float64, 3*npt*igrid values,
c         replica-major, atom-major within a replica)
c       selection file (raw int32, npt values: MOIL's ipick)
c       a b c                          (box edges, Angstrom)
c       rvmx relx rvbg rebg cutm       (Angstrom, squared here)
c       dtol nfft1 nfft2 nfft3 iord
c       eps
c       igrid gamma rho lambda
c       output file path
c       [ihvdw]                        (optional, 0 = hydrogen rule off)
c
      program moil_sds
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

      character*512 pfile,confile,rfile,pkfile,outfile
      double precision rvmx,relx,rvbg,rebg,cutm
      double precision gamma,rho,lambda,sener
      integer iord,igrid,ihvdw,nselec,ntot
      integer i,j,k,m,ucon,ur,upar,u
      integer*8 c0,c1,crate
      double precision, allocatable :: r(:,:),dv(:,:),d0(:),e0(:)
      double precision, allocatable :: e1(:),dps(:,:),ets(:)
      integer, allocatable :: ipick(:),pointr(:)

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
      read(upar,'(a)') rfile
      read(upar,'(a)') pkfile
      call init_ef()
      call init_var()
      read(upar,*) a,b,c
      read(upar,*) rvmx,relx,rvbg,rebg,cutm
      read(upar,*) dtol,nfft1,nfft2,nfft3,iord
      read(upar,*) eps
      read(upar,*) igrid,gamma,rho,lambda
      read(upar,'(a)') outfile
      ihvdw = 1
      read(upar,*,end=15) ihvdw
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

      ntot = igrid*npt
      allocate(r(3,ntot),dv(3,ntot),d0(2*igrid),e0(igrid))
      allocate(e1(2*igrid),dps(3,ntot),ets(igrid))
      allocate(ipick(npt),pointr(npt))
      ur = 33
      open(unit=ur,file=rfile,status='old',access='stream',
     1     form='unformatted')
      read(ur) ((r(k,i),k=1,3),i=1,ntot)
      close(ur)
      open(unit=ur,file=pkfile,status='old',access='stream',
     1     form='unformatted')
      read(ur) (ipick(i),i=1,npt)
      close(ur)
      nselec = 0
      do 20 i=1,npt
         pointr(i) = i
         if (ipick(i).gt.0) nselec = nselec + 1
 20   continue

c     cutoffs exactly as pot/energy.f:382-400 (cutm always explicit)
      cutvdw2  = rvmx*rvmx
      cutvbig2 = rvbg*rvbg
      cutele2  = relx*relx
      cutebig2 = rebg*rebg
      cutmono2 = cutm*cutm

      call ewald_init()

      do 25 j=1,igrid
         e0(j) = 0.d0
 25   continue
      call system_clock(count_rate=crate)
      call system_clock(c0)
      call sds(sener,dv,r,d0,e0,e1,nselec,pointr,ipick,npt,3*ntot,
     1         gamma,rho,lambda,.true.,.false.)
      call system_clock(c1)
      write(stdo,'(a,es16.8)') 'MOILTIME sds_s=',
     1     dble(c1-c0)/dble(crate)

c     each interior replica's own energy and gradient, as sds obtains
c     them (chain/sds.f:71-110); the end replicas get none
      do 30 i=1,ntot
         dps(1,i) = 0.d0
         dps(2,i) = 0.d0
         dps(3,i) = 0.d0
 30   continue
      do 50 j=1,igrid
         ets(j) = 0.d0
 50   continue
      do 60 j=2,igrid-1
         m = (j-1)*npt
         do 40 i=1,npt
            coor(1,i) = r(1,m+i)
            coor(2,i) = r(2,m+i)
            coor(3,i) = r(3,m+i)
 40      continue
         call squeeze()
         call nbondm()
         call syminit()
         call eforce()
         ets(j) = e_total
         do 45 i=1,npt
            dps(1,m+i) = dpot(1,i)
            dps(2,m+i) = dpot(2,i)
            dps(3,m+i) = dpot(3,i)
 45      continue
 60   continue

      u = 41
      open(unit=u,file=outfile,status='replace',access='stream',
     1     form='unformatted')
      call wrd(u,'SENER   ',1,sener)
      call wrd(u,'DV      ',3*ntot,dv)
      call wrd(u,'D0      ',2*igrid-3,d0)
      call wrd(u,'E0      ',igrid,e0)
      call wrd(u,'E1      ',2*igrid-3,e1)
      call wrd(u,'DPOTS   ',3*ntot,dps)
      call wrd(u,'ETOTS   ',igrid,ets)
      call wrd(u,'BETA    ',1,ewaldcof)
      close(u)
      stop
      end

      subroutine wrd(u,nm,n,arr)
      implicit none
      integer u,n,i
      character*8 nm
      double precision arr(*)
      write(u) nm,1,n,(arr(i),i=1,n)
      return
      end
