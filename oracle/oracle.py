"""ctypes wrapper of the C restatement (oracle/moil_oracle.c).

Test infrastructure only: the checker the GPU path is compared with.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from moil_fnrfd_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libmoil_oracle.so')
_lib = None


def build():
    subprocess.run(['make', '-s', '-C', HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(abi.MoilTopology), C.POINTER(abi.MoilParams),
                                 C.c_char_p, C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_ewald_info.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_int), C.c_void_p, C.c_int, C.c_void_p,
                                     C.c_void_p, C.c_void_p]
        L.orc_squeeze.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_build_lists.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_dump_water_pairs.restype = C.c_int64
        L.orc_dump_water_pairs.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        L.orc_dump_atom_pairs.restype = C.c_int64
        L.orc_dump_atom_pairs.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        L.orc_eforce.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_set_molecules.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p]
        L.orc_virial.argtypes = [C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


class Oracle:
    """FP64 CPU restatement of MOIL's eforce path on one system."""

    def __init__(self, topo, params: abi.MoilParams):
        L = lib()
        self._ts, self._keep = abi.make_topology(topo)
        self.params = params
        err = C.create_string_buffer(512)
        self._h = L.orc_create(C.byref(self._ts), C.byref(params), err, 512)
        if not self._h:
            raise ValueError(err.value.decode())
        self.npt = topo.npt

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ewald_info(self):
        beta, es, nt = C.c_double(), C.c_double(), C.c_int()
        lib().orc_ewald_info(self._h, C.byref(beta), C.byref(es), C.byref(nt), None, 0,
                             None, None, None)
        tab = np.zeros((nt.value, 4))
        p = self.params
        b1, b2, b3 = np.zeros(p.nfft1), np.zeros(p.nfft2), np.zeros(p.nfft3)
        lib().orc_ewald_info(self._h, None, None, None, tab.ctypes.data, nt.value,
                             b1.ctypes.data, b2.ctypes.data, b3.ctypes.data)
        return dict(beta=beta.value, e_self=es.value, ntab=nt.value, erf_arr=tab,
                    bsp=(b1, b2, b3))

    def squeeze(self, coor):
        x = np.ascontiguousarray(coor, np.float64).copy()
        lib().orc_squeeze(self._h, x.ctypes.data)
        return x

    def build_lists(self, coor):
        x = np.ascontiguousarray(coor, np.float64)
        lib().orc_build_lists(self._h, x.ctypes.data)

    def water_pairs(self):
        n = lib().orc_dump_water_pairs(self._h, None, 0)
        rows = np.zeros((max(n, 1), 4), np.int32)
        lib().orc_dump_water_pairs(self._h, rows.ctypes.data, n)
        return rows[:n]

    def atom_pairs(self):
        n = lib().orc_dump_atom_pairs(self._h, None, 0)
        rows = np.zeros((max(n, 1), 4), np.int32)
        lib().orc_dump_atom_pairs(self._h, rows.ctypes.data, n)
        return rows[:n]

    def set_molecules(self, patom, ordlist, pmol, ptms):
        """Molecules for the virial (MOIL's PRESS.BLOCK arrays, 1-based);
        an empty patom switches it off."""
        self._mol = [np.ascontiguousarray(patom, np.int32), np.ascontiguousarray(ordlist, np.int32),
                     np.ascontiguousarray(pmol, np.int32), np.ascontiguousarray(ptms, np.float64)]
        lib().orc_set_molecules(self._h, len(self._mol[0]), *[a.ctypes.data for a in self._mol])

    def virial(self):
        """The virial terms of the last eforce (abi.VIRIAL_NAMES order)."""
        v = np.zeros(13)
        if lib().orc_virial(self._h, v.ctypes.data):
            raise RuntimeError('orc_virial: no molecules set')
        return v

    def eforce(self, coor):
        x = np.ascontiguousarray(coor, np.float64)
        dpot = np.zeros((self.npt, 3))
        en = abi.MoilEnergies()
        rc = lib().orc_eforce(self._h, x.ctypes.data, dpot.ctypes.data, C.byref(en))
        if rc:
            raise RuntimeError('orc_eforce failed %d' % rc)
        return en.as_dict(), dpot


def ref_canonical_water_pairs(d: dict) -> np.ndarray:
    """Canonical (i, j, symop, list) rows from a MOIL dump (oracle/_ref)."""
    rows = []
    nw = int(d['COUNTS'][1])
    indx = d.get('INDXWAT')
    for kind in (1, 2):
        poin = d.get('POINWAT%d' % kind)
        lst = d.get('LISTWT%d' % kind)
        if lst is None:
            continue
        nreal = int(poin[nw]) if poin is not None else 0
        if poin is not None:
            for iw in range(1, nw):
                for k in range(poin[iw], poin[iw + 1]):
                    rows.append((indx[iw - 1], lst[k], 0, kind))
        ib = d['IBLCKWT%d' % kind]
        sw = d.get('SYMMWAT%d' % kind)
        pw = d.get('PSYMWT%d' % kind)
        jb = nreal
        for s in range(1, len(ib)):
            for w in range(ib[s - 1], ib[s]):
                for k in range(jb, pw[w]):
                    rows.append((sw[w], lst[k], s, kind))
                jb = pw[w]
    return np.array(rows, np.int32).reshape(-1, 4)


def ref_canonical_atom_pairs(d: dict) -> np.ndarray:
    rows = []
    npt = int(d['COUNTS'][0])
    for kind in (1, 2, 3):
        pt = d['POINT%d' % kind]
        lst = d.get('LIST%d' % kind)
        if lst is None:
            continue
        nreal = int(pt[npt])
        for l in range(1, npt + 1):
            for k in range(pt[l - 1], pt[l]):
                rows.append((l, lst[k], 0, kind))
        ib = d['IBLOCK%d' % kind]
        sr = d.get('SYMREAL%d' % kind)
        ps = d.get('PSYM%d' % kind)
        jb = nreal
        prev = 0
        for s in range(1, len(ib) + 1):
            for w in range(prev, ib[s - 1]):
                for k in range(jb, ps[w]):
                    rows.append((sr[w], lst[k], s, kind))
                jb = ps[w]
            prev = ib[s - 1]
    return np.array(rows, np.int32).reshape(-1, 4)


def sort_rows(r: np.ndarray) -> np.ndarray:
    if len(r) == 0:
        return r.reshape(-1, 4)
    idx = np.lexsort((r[:, 3], r[:, 2], r[:, 1], r[:, 0]))
    return r[idx]
