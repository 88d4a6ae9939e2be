c     moil_gpu_shim_chain.f -- ISO_C_BINDING shim that routes MOIL's
c     chain functional sds (chain/sds.f:1-257) to moil_gpu_chain_sds
c     (include/moil_gpu.h).  Linked together with moil_gpu_shim.f in
c     place of the object of chain/sds.f: MOIL's chain minimiser,
c     annealer and chain dynamics (chain/chnneal.f:85,
c     chain/chndyn.f:112, chain/pwl_chn.f:200) keep calling sds with
c     the same arguments, and every call evaluates all interior
c     replicas as one batch on the GPU.
c
c     On the first call one library context per interior replica is
c     initialised from the COMMON data, the way the shim's ewald_init
c     initialises the single-system context; the chain is created over
c     them.  A later call with another number of replicas rebuilds them.
c
c     Build (from the MOIL root, with both shim files copied to pot/):
c       flang -O2 -mcmodel=medium -c pot/moil_gpu_shim.f
c       flang -O2 -mcmodel=medium -c pot/moil_gpu_shim_chain.f
c       link: as moil_gpu_shim.f says, and replace chain/sds.o by
c       pot/moil_gpu_shim_chain.o
c
      module moil_gpu_chain_c
      use iso_c_binding
      use moil_gpu_c
      implicit none
      interface
        integer(c_int32_t) function moil_gpu_chain_create(ctxs, igrid,
     &      ch) bind(C, name='moil_gpu_chain_create')
          import :: c_ptr, c_int32_t
          type(c_ptr), intent(in) :: ctxs(*)
          integer(c_int32_t), value :: igrid
          type(c_ptr), intent(out) :: ch
        end function
        integer(c_int32_t) function moil_gpu_chain_sds(ch, r, ipick,
     &      gamma, rho, lambda, sener, dv, d0, e0, e1)
     &      bind(C, name='moil_gpu_chain_sds')
          import :: c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ch
          real(c_double), intent(in) :: r(*)
          integer(c_int32_t), intent(in) :: ipick(*)
          real(c_double), value :: gamma, rho, lambda
          real(c_double), intent(out) :: sener
          real(c_double), intent(out) :: dv(*), d0(*), e1(*)
          real(c_double), intent(inout) :: e0(*)
        end function
        integer(c_int32_t) function moil_gpu_chain_destroy(ch)
     &      bind(C, name='moil_gpu_chain_destroy')
          import :: c_ptr, c_int32_t
          type(c_ptr), value :: ch
        end function
        integer(c_int32_t) function moil_gpu_finalize(ctx)
     &      bind(C, name='moil_gpu_finalize')
          import :: c_ptr, c_int32_t
          type(c_ptr), value :: ctx
        end function
      end interface
      type(c_ptr), save :: mg_chain = c_null_ptr
      integer, save :: mg_chain_igrid = 0
      type(c_ptr), allocatable, save :: mg_chain_ctx(:)
      end module moil_gpu_chain_c


c     sds (chain/sds.f:1-257): same arguments and results; nselec,
c     pointr, fixend and debug are not used by sds's arithmetic either
      subroutine sds(sener,dv,r,d0,e0,e1,nselec,pointr,ipick,npt,
     &               ndegf,gamma,rho,lambda,fixend,debug)
      use moil_gpu_chain_c
      implicit none
      double precision sener,gamma,rho,lambda
      integer nselec,npt,ndegf
      double precision dv(3,*),r(3,*),d0(*),e0(*),e1(*)
      integer pointr(*),ipick(*)
      logical fixend,debug
      integer igrid,j,rc
      type(c_ptr) :: keep
      igrid = ndegf/(3*npt)
c     sds squeezes each replica only when esymyes (chain/sds.f:84-86);
c     the library always does, and the mode check stops a run without
c     periodic images (as it stops every other mode it does not evaluate)
      call mg_modes('sds')
      if (mg_chain_igrid .ne. igrid) then
        if (c_associated(mg_chain)) then
          rc = moil_gpu_chain_destroy(mg_chain)
          do j = 1, size(mg_chain_ctx)
            rc = moil_gpu_finalize(mg_chain_ctx(j))
          end do
          deallocate(mg_chain_ctx)
          mg_chain = c_null_ptr
        end if
c       one context per interior replica, from the COMMON data
        keep = mg_ctx
        allocate(mg_chain_ctx(max(igrid-2,1)))
        do j = 1, igrid-2
          mg_ctx = c_null_ptr
          call ewald_init()
          mg_chain_ctx(j) = mg_ctx
        end do
        mg_ctx = keep
        call mg_check(moil_gpu_chain_create(mg_chain_ctx, igrid,
     &    mg_chain), 'sds')
        mg_chain_igrid = igrid
      end if
      call mg_check(moil_gpu_chain_sds(mg_chain, r, ipick, gamma, rho,
     &  lambda, sener, dv, d0, e0, e1), 'sds')
      end subroutine
