c     moil_traj -- test-infrastructure driver: writes a binary trajectory
c     with MOIL's own wdyncrd (dynamics/wdyncrd.f) so the library's writer
c     can be compared with it byte for byte.  Not part of the product.
c
c     Usage: moil_traj <param-file>
c     param-file (list-directed):
c       npt nfrm
c       frame file path (raw float64, nfrm*3*npt values, frame-major,
c         atom-major within a frame)
c       output file path
      program moil_traj
      implicit none
      include 'COMMON/LENGTH.BLOCK'
      include 'COMMON/CONNECT.BLOCK'
      include 'COMMON/COORD.BLOCK'
      character*512 pfile,ffile,outfile
      integer nfrm,f,i,k,upar,uf,uw
      integer nofreez(1)
      call get_command_argument(1,pfile)
      upar = 31
      open(unit=upar,file=pfile,status='old')
      read(upar,*) npt,nfrm
      read(upar,'(a)') ffile
      read(upar,'(a)') outfile
      close(upar)
      uf = 32
      open(unit=uf,file=ffile,status='old',access='stream',
     1     form='unformatted')
      uw = 33
      open(unit=uw,file=outfile,status='replace',form='unformatted',
     1     access='sequential')
      nofreez(1) = 0
      do 10 f=1,nfrm
         read(uf) ((coor(k,i),k=1,3),i=1,npt)
         call wdyncrd(uw,nfrm,f,npt,nofreez,1)
 10   continue
      close(uw)
      close(uf)
      stop
      end
