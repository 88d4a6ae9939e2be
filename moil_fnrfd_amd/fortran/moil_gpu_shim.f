c     moil_gpu_shim.f -- ISO_C_BINDING shim that routes MOIL's
c     energy/force path to libmoil_gpu.so (include/moil_gpu.h).
c
c     It is compiled inside the MOIL source tree (it includes MOIL's
c     COMMON blocks) and linked in place of the objects of
c       symm/ewald_init.f  (ewald_init)   -> moil_gpu_init
c       symm/squeeze.f     (squeeze)      -> moil_gpu_squeeze
c       pot/nbondm.f       (nbondm)       -> moil_gpu_build_lists
c       symm/syminit.f     (syminit)      -> no-op (lists built above)
c       pot/eforce.f       (eforce)       -> moil_gpu_eforce (with
c                                            pressON: eforce_virial)
c     so every MOIL driver (dyna, chain, milestoning, ...) keeps
c     calling the same argument-less subroutines on the same COMMON
c     data.  A non-zero status from the library is reported through
c     MOIL's alert(...,1), which stops the program as MOIL does.
c
c     Build (from the MOIL root, with this file copied to pot/):
c       flang -O2 -mcmodel=medium -c pot/moil_gpu_shim.f -o pot/moil_gpu_shim.o
c       link: replace the five objects above by pot/moil_gpu_shim.o and
c       add  -L<repo>/moil_fnrfd_amd -lmoil_gpu -Wl,-rpath,<repo>/moil_fnrfd_amd
c
      module moil_gpu_c
      use iso_c_binding
      implicit none
c     struct moil_topology / moil_params / moil_energies (moil_gpu.h)
      type, bind(C) :: mg_topology
        integer(c_int32_t) :: npt, totmon
        type(c_ptr) :: poipt, monoflag, ptchg, epsgm6, epsgm12
        type(c_ptr) :: flagchr, exc1, exc2
        integer(c_int32_t) :: nb
        type(c_ptr) :: ib1, ib2, kbond, req
        integer(c_int32_t) :: nangl
        type(c_ptr) :: iangl1, iangl2, iangl3, kangl, angleq
        integer(c_int32_t) :: ntors
        type(c_ptr) :: itor1, itor2, itor3, itor4, period
        type(c_ptr) :: ktors1, ktors2, ktors3
        type(c_ptr) :: phase1, phase2, phase3
        integer(c_int32_t) :: nimp
        type(c_ptr) :: iimp1, iimp2, iimp3, iimp4, kimp, impeq
        integer(c_int32_t) :: totspe
        type(c_ptr) :: spec1, spec2, p14
        integer(c_int32_t) :: nwaters
        type(c_ptr) :: idxtip3
      end type
      type, bind(C) :: mg_params
        real(c_double) :: a, b, c
        real(c_double) :: cutvdw2, cutele2, cutvbig2, cutebig2
        real(c_double) :: cutmono2, eps, dtol
        integer(c_int32_t) :: nfft1, nfft2, nfft3, order
        integer(c_int32_t) :: precision, terms
      end type
      type, bind(C) :: mg_energies
        real(c_double) :: e_total, e_bond, e_theta, e_tors, e_imp
        real(c_double) :: e_vdw, e_el, e_vsym, e_lsym, e_el14
        real(c_double) :: e_vdw14, e_corr, e_self, e_ew_receip
        real(c_double) :: e_dir
      end type
c     struct moil_virial
      type, bind(C) :: mg_virial
        real(c_double) :: virial, virxx, viryy, virzz
        real(c_double) :: dirxx, diryy, dirzz
        real(c_double) :: recxx, recyy, reczz
        real(c_double) :: relxx, relyy, relzz
      end type
      interface
        integer(c_int32_t) function moil_gpu_init(topo, par, ctx)
     &      bind(C, name='moil_gpu_init')
          import :: mg_topology, mg_params, c_ptr, c_int32_t
          type(mg_topology), intent(in) :: topo
          type(mg_params), intent(in) :: par
          type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int32_t) function moil_gpu_ewald_info(ctx, beta,
     &      eself, ntab, dens) bind(C, name='moil_gpu_ewald_info')
          import :: c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(out) :: beta, eself, dens
          integer(c_int32_t), intent(out) :: ntab
        end function
        integer(c_int32_t) function moil_gpu_squeeze(ctx, coor)
     &      bind(C, name='moil_gpu_squeeze')
          import :: c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(inout) :: coor(*)
        end function
        integer(c_int32_t) function moil_gpu_build_lists(ctx, coor)
     &      bind(C, name='moil_gpu_build_lists')
          import :: c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(in) :: coor(*)
        end function
        integer(c_int32_t) function moil_gpu_eforce(ctx, coor, dpot,
     &      e) bind(C, name='moil_gpu_eforce')
          import :: c_ptr, c_int32_t, c_double, mg_energies
          type(c_ptr), value :: ctx
          real(c_double), intent(in) :: coor(*)
          real(c_double), intent(out) :: dpot(*)
          type(mg_energies), intent(out) :: e
        end function
        integer(c_int32_t) function moil_gpu_eforce_virial(ctx, coor,
     &      dpot, e, v) bind(C, name='moil_gpu_eforce_virial')
          import :: c_ptr, c_int32_t, c_double, mg_energies, mg_virial
          type(c_ptr), value :: ctx
          real(c_double), intent(in) :: coor(*)
          real(c_double), intent(out) :: dpot(*)
          type(mg_energies), intent(out) :: e
          type(mg_virial), intent(out) :: v
        end function
        integer(c_int32_t) function moil_gpu_set_molecules(ctx, nmol,
     &      patom, ordlist, pmol, ptms)
     &      bind(C, name='moil_gpu_set_molecules')
          import :: c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: nmol
          integer(c_int32_t), intent(in) :: patom(*), ordlist(*)
          integer(c_int32_t), intent(in) :: pmol(*)
          real(c_double), intent(in) :: ptms(*)
        end function
        type(c_ptr) function moil_gpu_last_error()
     &      bind(C, name='moil_gpu_last_error')
          import :: c_ptr
        end function
      end interface
      type(c_ptr), save :: mg_ctx = c_null_ptr
c     molecules handed to the library (init_pressure's, PRESS.BLOCK)
      integer, save :: mg_nmol = 0
      integer(c_int32_t), allocatable, target, save :: mg_flags(:)
      contains
c     status check: MOIL's alert(...,1) prints and stops
      subroutine mg_check(rc, where)
      integer(c_int32_t), intent(in) :: rc
      character(len=*), intent(in) :: where
      character(kind=c_char), pointer :: msg(:)
      character(len=200) :: txt
      integer :: k
      if (rc .eq. 0) return
      call c_f_pointer(moil_gpu_last_error(), msg, [200])
      txt = ' '
      do k = 1, 200
        if (msg(k) .eq. c_null_char) exit
        txt(k:k) = msg(k)
      end do
      call alert(where, len(where), txt, len_trim(txt), 1)
      end subroutine

c     the topology arrays reach C through TARGET dummies; the library
c     copies them to the GPU during the call
      subroutine mg_init_topo(par, npt, totmon, poipt, ptchg, epsgm6,
     &  epsgm12, flagchr, exc1, exc2, nb, ib1, ib2, kbond, req, nangl,
     &  iangl1, iangl2, iangl3, kangl, angleq, ntors, itor1, itor2,
     &  itor3, itor4, period, ktors1, ktors2, ktors3, phase1, phase2,
     &  phase3, nimp, iimp1, iimp2, iimp3, iimp4, kimp, impeq, totspe,
     &  spec1, spec2, p14, nwaters, idxtip3)
      type(mg_params), intent(in) :: par
      integer, intent(in) :: npt, totmon, nb, nangl, ntors, nimp
      integer, intent(in) :: totspe, nwaters
      integer, target :: poipt(0:*), exc1(0:*), exc2(*)
      logical, target :: flagchr(*)
      double precision, target :: ptchg(*), epsgm6(*), epsgm12(*)
      integer, target :: ib1(*), ib2(*), iangl1(*), iangl2(*)
      integer, target :: iangl3(*), itor1(*), itor2(*), itor3(*)
      integer, target :: itor4(*), period(*), iimp1(*), iimp2(*)
      integer, target :: iimp3(*), iimp4(*), spec1(*), spec2(*)
      integer, target :: idxtip3(*)
      double precision, target :: kbond(*), req(*), kangl(*)
      double precision, target :: angleq(*), ktors1(*), ktors2(*)
      double precision, target :: ktors3(*), phase1(*), phase2(*)
      double precision, target :: phase3(*), kimp(*), impeq(*)
      double precision, target :: p14(3,*)
      type(mg_topology) :: t
      t%npt = npt
      t%totmon = totmon
      t%poipt = c_loc(poipt(0))
      t%monoflag = c_loc(mg_flags(1))
      t%ptchg = c_loc(ptchg(1))
      t%epsgm6 = c_loc(epsgm6(1))
      t%epsgm12 = c_loc(epsgm12(1))
      t%flagchr = c_loc(flagchr(1))
      t%exc1 = c_loc(exc1(0))
      t%exc2 = c_loc(exc2(1))
      t%nb = nb
      t%ib1 = c_loc(ib1(1))
      t%ib2 = c_loc(ib2(1))
      t%kbond = c_loc(kbond(1))
      t%req = c_loc(req(1))
      t%nangl = nangl
      t%iangl1 = c_loc(iangl1(1))
      t%iangl2 = c_loc(iangl2(1))
      t%iangl3 = c_loc(iangl3(1))
      t%kangl = c_loc(kangl(1))
      t%angleq = c_loc(angleq(1))
      t%ntors = ntors
      t%itor1 = c_loc(itor1(1))
      t%itor2 = c_loc(itor2(1))
      t%itor3 = c_loc(itor3(1))
      t%itor4 = c_loc(itor4(1))
      t%period = c_loc(period(1))
      t%ktors1 = c_loc(ktors1(1))
      t%ktors2 = c_loc(ktors2(1))
      t%ktors3 = c_loc(ktors3(1))
      t%phase1 = c_loc(phase1(1))
      t%phase2 = c_loc(phase2(1))
      t%phase3 = c_loc(phase3(1))
      t%nimp = nimp
      t%iimp1 = c_loc(iimp1(1))
      t%iimp2 = c_loc(iimp2(1))
      t%iimp3 = c_loc(iimp3(1))
      t%iimp4 = c_loc(iimp4(1))
      t%kimp = c_loc(kimp(1))
      t%impeq = c_loc(impeq(1))
      t%totspe = totspe
      t%spec1 = c_loc(spec1(1))
      t%spec2 = c_loc(spec2(1))
      t%p14 = c_loc(p14(1,1))
      t%nwaters = nwaters
      t%idxtip3 = c_loc(idxtip3(1))
      call mg_check(moil_gpu_init(t, par, mg_ctx), 'ewald_init')
      end subroutine

      end module moil_gpu_c


c     The energy terms and modes of pot/eforce.f:86-311 that the library
c     does not evaluate: with any of them switched on MOIL would compute a
c     different energy, so the shim stops through alert(...,1) instead of
c     returning the library's.  The library evaluates the constant
c     dielectric (ctrue) Ewald path with periodic images (ewaldyes,
c     esymyes), MOIL's geometric or arith combination rule, no LES
c     copies, no frozen particles, one process, and the terms bond, angle,
c     torsion, improper, LJ, Coulomb and the scaled 1-4 pairs.
      subroutine mg_modes(where)
      use moil_gpu_c
      character(len=*), intent(in) :: where
      include 'COMMON/LENGTH.BLOCK'
      include 'COMMON/CONNECT.BLOCK'
      include 'COMMON/ENERGY.BLOCK'
      include 'COMMON/CONSTRAN.BLOCK'
      include 'COMMON/EWALD.BLOCK'
      include 'COMMON/PARALLEL.BLOCK'
      include 'COMMON/SGB.BLOCK'
      include 'COMMON/SSBP.BLOCK'
      include 'COMMON/REPWALL.BLOCK'
      include 'COMMON/EFIELD.BLOCK'
      include 'COMMON/PRESS.BLOCK'
      include 'COMMON/FREEZ.BLOCK'
      character(len=60) :: why
      why = ' '
      if (.not. ctrue) why = 'distance-dependent dielectric (rdie)'
      if (shift) why = 'shifted nonbonded potential (shift)'
      if (efyes) why = 'finite nonbonded interactions (efyes)'
      if (.not. ewaldyes) why = 'nonbonded terms without Ewald'
      if (.not. esymyes) why = 'no periodic images (esymyes off)'
      if (metalyes) why = 'metal walls (metalyes)'
      if (gbsabool) why = 'generalized Born solvation (gbsa)'
      if (vp_flag) why = 'virtual particles (vp_flag)'
      if (ethyes .and. nurey .gt. 0) why = 'Urey-Bradley terms'
      if (ecnyes .and. ncnst .gt. 0) why = 'internal constraints'
      if (lcent .and. icenter .gt. 0) why = 'centre-of-mass constraint'
      if (eballyes) why = 'water sphere constraint (eball)'
      if (nmb .gt. 0) why = 'Morse or repulsion bonds'
      if (eteth_yes) why = 'tether energy'
      if (ehyes) why = 'hydrophobic energy'
      if (eenmyes) why = 'elastic network energy'
      if (eCGyes) why = 'coarse-grained energy terms'
      if (e_Kumbyes) why = 'umbrella constraint energy'
      if (qssbp) why = 'solvent boundary potential (ssbp)'
      if (torscstr) why = 'torsion constraints'
      if (alabar) why = 'alanine barrier'
      if (rgcst) why = 'radius of gyration constraint'
      if (rwall) why = 'repulsive wall'
      if (efield_yes) why = 'external electric field'
      if (esteeryes) why = 'steering energy'
      if (lesflag) why = 'LES copies'
      if (pressON .and. nmol .le. 0)
     &  why = 'pressure / virial without molecules (init_pressure)'
      if (inofrz .ne. npt) why = 'frozen particles'
      if (prll_on_off .and. num_pes .gt. 1) why = 'several processes'
      if (why .ne. ' ') call alert(where, len(where),
     &  'not supported by the GPU library: '//trim(why),
     &  len_trim('not supported by the GPU library: '//trim(why)), 1)
      end subroutine

c     ewald_init (symm/ewald_init.f:8-242): called once after the
c     cutoffs are squared (pot/energy.f:392-410)
      subroutine ewald_init
      use moil_gpu_c
      include 'COMMON/LENGTH.BLOCK'
      include 'COMMON/CONNECT.BLOCK'
      include 'COMMON/NBLIST.BLOCK'
      include 'COMMON/SYMM.BLOCK'
      include 'COMMON/ENERGY.BLOCK'
      include 'COMMON/EWALD.BLOCK'
      include 'COMMON/UNITS.BLOCK'
      type(mg_params) :: p
      integer :: i, ntab
      double precision :: dens
      character*4 nm
      call mg_modes('ewald_init')
c     monomer flags from the names MOIL's list and squeeze code test
c     (pot/nbondm.f:81-82, symm/squeeze.f:64-79)
      if (allocated(mg_flags)) deallocate(mg_flags)
      allocate(mg_flags(max(totmon,1)))
      do i = 1, totmon
        nm = moname(i)
        mg_flags(i) = 0
        if (nm .eq. 'TIP3' .or. nm(1:3) .eq. 'SPC') then
          mg_flags(i) = ior(mg_flags(i), 1)
        end if
        if (nm .eq. 'TIP3' .or. nm .eq. 'NAAT' .or. nm .eq. 'NAO'
     &      .or. nm .eq. 'CL' .or. nm .eq. 'MG' .or. nm .eq. 'CMO'
     &      .or. nm .eq. 'CO' .or. nm .eq. 'SUL' .or. nm(1:3) .eq.
     &      'SPC' .or. nm .eq. 'K' .or. nm .eq. 'AR') then
          mg_flags(i) = ior(mg_flags(i), 2)
        end if
        if (nm .eq. 'DMPC' .or. nm .eq. 'DOPC' .or. nm .eq. 'NATA')
     &    mg_flags(i) = ior(mg_flags(i), 4)
      end do
      p%a = a
      p%b = b
      p%c = c
      p%cutvdw2 = cutvdw2
      p%cutele2 = cutele2
      p%cutvbig2 = cutvbig2
      p%cutebig2 = cutebig2
      p%cutmono2 = cutmono2
      p%eps = eps
      p%dtol = dtol
      p%nfft1 = nfft1
      p%nfft2 = nfft2
      p%nfft3 = nfft3
      p%order = intrpord
c     FP32 pair arithmetic with fixed-point forces unless MOIL_GPU_FP64=1
c     asks for the FP64 parity mode
      p%precision = 1
      call get_environment_variable('MOIL_GPU_FP64', nm)
      if (nm(1:1) .eq. '1') p%precision = 0
      if (p%precision .eq. 1) write(stdo,'(1x,a)')
     &  'GPU library: FP32 pair arithmetic with fixed-point forces'//
     &  ' (MOIL_GPU_FP64=1 for the FP64 mode)'
      p%terms = 0
      if (ebyes) p%terms = ior(p%terms, 1)
      if (ethyes) p%terms = ior(p%terms, 2)
      if (etoyes) p%terms = ior(p%terms, 4)
      if (eimyes) p%terms = ior(p%terms, 8)
      if (evdyes) p%terms = ior(p%terms, 16)
      if (eelyes) p%terms = ior(p%terms, 32)
      if (e14el_yes) p%terms = ior(p%terms, 64)
      if (e14v_yes) p%terms = ior(p%terms, 128)
      if (hvdw0) p%terms = ior(p%terms, 256)
      if (arith) p%terms = ior(p%terms, 512)
      call mg_init_topo(p, npt, totmon, poipt, ptchg, epsgm6, epsgm12,
     &  flagchr, exc1, exc2, nb, ib1, ib2, kbond, req, nangl, iangl1,
     &  iangl2, iangl3, kangl, angleq, ntors, itor1, itor2, itor3,
     &  itor4, period, ktors1, ktors2, ktors3, phase1, phase2, phase3,
     &  nimp, iimp1, iimp2, iimp3, iimp4, kimp, impeq, totspe, spec1,
     &  spec2, p14, nwaters, idxtip3)
c     MOIL code elsewhere reads ewaldcof, e_self and erftbdns
      call mg_check(moil_gpu_ewald_info(mg_ctx, ewaldcof, e_self,
     &  ntab, dens), 'ewald_init')
      erftbdns = dens
      end subroutine

c     squeeze (symm/squeeze.f:1-147): coor is wrapped in place
      subroutine squeeze
      use moil_gpu_c
      include 'COMMON/LENGTH.BLOCK'
      include 'COMMON/COORD.BLOCK'
      call mg_check(moil_gpu_squeeze(mg_ctx, coor), 'squeeze')
      end subroutine

c     nbondm (pot/nbondm.f) + nbond + nbmsym + nbsym: all lists at once
      subroutine nbondm
      use moil_gpu_c
      include 'COMMON/LENGTH.BLOCK'
      include 'COMMON/COORD.BLOCK'
      call mg_check(moil_gpu_build_lists(mg_ctx, coor), 'nbondm')
      end subroutine

c     syminit (symm/syminit.f): the image lists come with nbondm
      subroutine syminit
      end subroutine

c     eforce (pot/eforce.f:1-325): energies into /ENERGY/ and /EWALD3/,
c     gradient into dpot
      subroutine eforce
      use moil_gpu_c
      include 'COMMON/LENGTH.BLOCK'
      include 'COMMON/CONNECT.BLOCK'
      include 'COMMON/COORD.BLOCK'
      include 'COMMON/ENERGY.BLOCK'
      include 'COMMON/EWALD.BLOCK'
      include 'COMMON/PRESS.BLOCK'
      type(mg_energies) :: e
      type(mg_virial) :: v
      call mg_modes('eforce')
      if (.not. pressON) then
        call mg_check(moil_gpu_eforce(mg_ctx, coor, dpot, e), 'eforce')
      else
c       the molecules of init_pressure, handed over once; the virial
c       terms into PRESS.BLOCK (pot/eforce.f:70-85 zeroes them, the
c       correction terms stay zero as in MOIL)
        if (mg_nmol .ne. nmol) then
          call mg_check(moil_gpu_set_molecules(mg_ctx, nmol, patom,
     &      ordLISTatm, pmol, ptms), 'eforce')
          mg_nmol = nmol
        end if
        call mg_check(moil_gpu_eforce_virial(mg_ctx, coor, dpot, e, v),
     &    'eforce')
        virial = v%virial
        virXX = v%virxx
        virYY = v%viryy
        virZZ = v%virzz
        V_PIdirXX = v%dirxx
        V_PIdirYY = v%diryy
        V_PIdirZZ = v%dirzz
        V_PIrecXX = v%recxx
        V_PIrecYY = v%recyy
        V_PIrecZZ = v%reczz
        V_PIrelXX = v%relxx
        V_PIrelYY = v%relyy
        V_PIrelZZ = v%relzz
        V_PIcorXX = 0.d0
        V_PIcorYY = 0.d0
        V_PIcorZZ = 0.d0
      end if
      e_total = e%e_total
      e_bond = e%e_bond
      e_theta = e%e_theta
      e_tors = e%e_tors
      e_imp = e%e_imp
      e_vdw = e%e_vdw
      e_el = e%e_el
      e_vsym = e%e_vsym
      e_lsym = e%e_lsym
      e_el14 = e%e_el14
      e_vdw14 = e%e_vdw14
      e_corr = e%e_corr
      e_self = e%e_self
      e_ew_receip = e%e_ew_receip
      e_dir = e%e_dir
      end subroutine
