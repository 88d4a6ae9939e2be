"""ctypes mirror of include/moil_gpu.h (structs, constants).

Shared by the product binding (moil_fnrfd_amd.eforce) and by the test
infrastructure; it only describes memory layouts and holds no compute.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

ABI_VERSION = 1

OK, EINVAL, ECHARGE, ECAPACITY, EHIP, ESTATE = 0, 1, 2, 3, 4, 5

MONO_WATER, MONO_SQZ_FIRST, MONO_SQZ_CENT = 1, 2, 4
PREC_FP64, PREC_MIXED = 0, 1

TERM_BOND, TERM_ANGLE, TERM_TORS, TERM_IMP = 1, 2, 4, 8
TERM_VDW, TERM_EL, TERM_EL14, TERM_VDW14, TERM_HVDW0 = 16, 32, 64, 128, 256
TERM_ARITH = 512  # MOIL's arith rule: Lorentz-Berthelot LJ, AMBER torsion/improper forms
TERM_DEFAULT = 0x1FF

N_ENERGIES = 15
ENERGY_NAMES = ('e_total', 'e_bond', 'e_theta', 'e_tors', 'e_imp', 'e_vdw', 'e_el',
                'e_vsym', 'e_lsym', 'e_el14', 'e_vdw14', 'e_corr', 'e_self',
                'e_ew_receip', 'e_dir')

N_VIRIAL = 13
VIRIAL_NAMES = ('virial', 'virXX', 'virYY', 'virZZ', 'V_PIdirXX', 'V_PIdirYY', 'V_PIdirZZ',
                'V_PIrecXX', 'V_PIrecYY', 'V_PIrecZZ', 'V_PIrelXX', 'V_PIrelYY', 'V_PIrelZZ')

_pi = C.POINTER(C.c_int32)
_pd = C.POINTER(C.c_double)


class MoilTopology(C.Structure):
    _fields_ = [
        ('npt', C.c_int32), ('totmon', C.c_int32),
        ('poipt', _pi), ('monoflag', _pi),
        ('ptchg', _pd), ('epsgm6', _pd), ('epsgm12', _pd),
        ('flagchr', _pi), ('exc1', _pi), ('exc2', _pi),
        ('nb', C.c_int32), ('ib1', _pi), ('ib2', _pi), ('kbond', _pd), ('req', _pd),
        ('nangl', C.c_int32), ('iangl1', _pi), ('iangl2', _pi), ('iangl3', _pi),
        ('kangl', _pd), ('angleq', _pd),
        ('ntors', C.c_int32), ('itor1', _pi), ('itor2', _pi), ('itor3', _pi), ('itor4', _pi),
        ('period', _pi), ('ktors1', _pd), ('ktors2', _pd), ('ktors3', _pd),
        ('phase1', _pd), ('phase2', _pd), ('phase3', _pd),
        ('nimp', C.c_int32), ('iimp1', _pi), ('iimp2', _pi), ('iimp3', _pi), ('iimp4', _pi),
        ('kimp', _pd), ('impeq', _pd),
        ('totspe', C.c_int32), ('spec1', _pi), ('spec2', _pi), ('p14', _pd),
        ('nwaters', C.c_int32), ('idxtip3', _pi),
    ]


class MoilParams(C.Structure):
    _fields_ = [
        ('a', C.c_double), ('b', C.c_double), ('c', C.c_double),
        ('cutvdw2', C.c_double), ('cutele2', C.c_double), ('cutvbig2', C.c_double),
        ('cutebig2', C.c_double), ('cutmono2', C.c_double),
        ('eps', C.c_double), ('dtol', C.c_double),
        ('nfft1', C.c_int32), ('nfft2', C.c_int32), ('nfft3', C.c_int32),
        ('order', C.c_int32), ('precision', C.c_int32), ('terms', C.c_uint32),
    ]


class MoilVirial(C.Structure):
    _fields_ = [(n, C.c_double) for n in VIRIAL_NAMES]

    def as_array(self):
        return np.array([getattr(self, n) for n in VIRIAL_NAMES])


class MoilEnergies(C.Structure):
    _fields_ = [(n, C.c_double) for n in ENERGY_NAMES]

    def as_dict(self):
        return {n: getattr(self, n) for n in ENERGY_NAMES}


def _arr(a, dt):
    a = np.ascontiguousarray(a, dtype=dt)
    if a.size == 0:
        a = np.zeros(1, dtype=dt)
    return a


def make_topology(t):
    """Build a MoilTopology from a systems.Topology; returns (struct,
    keepalive-list).  The numpy arrays in keepalive must outlive calls."""
    keep = []

    def I(a):
        x = _arr(a, np.int32)
        keep.append(x)
        return x.ctypes.data_as(_pi)

    def D(a):
        x = _arr(a, np.float64)
        keep.append(x)
        return x.ctypes.data_as(_pd)

    s = MoilTopology()
    s.npt = t.npt
    s.totmon = t.totmon
    s.poipt = I(t.poipt)
    s.monoflag = I(t.monoflag)
    s.ptchg, s.epsgm6, s.epsgm12 = D(t.ptchg), D(t.epsgm6), D(t.epsgm12)
    s.flagchr = I(t.flagchr.astype(np.int32))
    s.exc1, s.exc2 = I(t.exc1), I(t.exc2)
    s.nb = len(t.ib1)
    s.ib1, s.ib2, s.kbond, s.req = I(t.ib1), I(t.ib2), D(t.kbond), D(t.req)
    s.nangl = len(t.iangl1)
    s.iangl1, s.iangl2, s.iangl3 = I(t.iangl1), I(t.iangl2), I(t.iangl3)
    s.kangl, s.angleq = D(t.kangl), D(t.angleq)
    s.ntors = len(t.itor1)
    s.itor1, s.itor2, s.itor3, s.itor4 = I(t.itor1), I(t.itor2), I(t.itor3), I(t.itor4)
    s.period = I(t.period)
    s.ktors1, s.ktors2, s.ktors3 = D(t.ktors1), D(t.ktors2), D(t.ktors3)
    s.phase1, s.phase2, s.phase3 = D(t.phase1), D(t.phase2), D(t.phase3)
    s.nimp = len(t.iimp1)
    s.iimp1, s.iimp2, s.iimp3, s.iimp4 = I(t.iimp1), I(t.iimp2), I(t.iimp3), I(t.iimp4)
    s.kimp, s.impeq = D(t.kimp), D(t.impeq)
    s.totspe = len(t.spec1)
    s.spec1, s.spec2 = I(t.spec1), I(t.spec2)
    s.p14 = D(np.asarray(t.p14, np.float64).reshape(-1))
    s.nwaters = t.nwaters
    s.idxtip3 = I(t.idxtip3)
    return s, keep


def make_params(box, *, rvmx, relx, rvbg, rebg, cutm, dtol, nfft, order=4, eps=1.0,
                precision=PREC_FP64, terms=TERM_DEFAULT):
    p = MoilParams()
    p.a, p.b, p.c = box
    # squared exactly as pot/energy.f:392-400
    p.cutvdw2 = rvmx * rvmx
    p.cutele2 = relx * relx
    p.cutvbig2 = rvbg * rvbg
    p.cutebig2 = rebg * rebg
    p.cutmono2 = cutm * cutm
    p.eps = eps
    p.dtol = dtol
    p.nfft1, p.nfft2, p.nfft3 = nfft
    p.order = order
    p.precision = precision
    p.terms = terms
    return p


class MoilDynaParams(C.Structure):
    _fields_ = [('dt', C.c_double), ('nlist', C.c_int32), ('mshake', C.c_int32),
                ('cmat', C.c_double * 9), ('tolcons', C.c_double), ('reqoh2', C.c_double),
                ('reqhh2', C.c_double), ('invmh', C.c_double), ('invmo', C.c_double)]


class MoilTuning(C.Structure):
    """Launch shapes of a context (include/moil_gpu.h moil_tuning)."""
    _fields_ = [('wchunk_pairs', C.c_int32), ('achunk_pairs', C.c_int32),
                ('water_wg_per_cu', C.c_int32), ('pme_tile', C.c_int32),
                ('us_default', C.c_double), ('us_tuned', C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# MOIL's time unit in picoseconds (COMMON/CONVERT.BLOCK tconv) and kB in
# kcal/mol/K (kboltzmann)
TCONV = 0.04888821290839616117
KBOLTZMANN = 1.987e-3
