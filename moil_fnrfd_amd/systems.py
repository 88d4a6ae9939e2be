"""Synthetic MOIL systems and the MOIL `.con` connectivity format.

A :class:`Topology` holds exactly what MOIL keeps in its CONNECT COMMON
block for the energy path (1-based atom/monomer indices, MOIL's array
names), so the same arrays can be handed to the GPU library, to the CPU
oracle and -- written as a `.con` file -- to MOIL's own reader
(generic/rconn.f:46-497).

MOIL's `.con` text format quantises every parameter (charges f9.5,
epsgm e12.5, bonds f10.4, angles f10.5 degrees, torsions f10.4/f6.2,
impropers e15.8, 1-4 terms f30.5/f15.5; generic/rconn.f:107-120).
Every generator below therefore rounds each parameter through the text
field it is written with (`_q`) before it is used anywhere, so that MOIL
and this package always work on identical float64 numbers.

Generators follow the synthetic-system recipes of the energy-path
benchmark: rigid 3-site water boxes (lattice + jitter, random
orientations) and random-walk heteropolymers in water.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

KOFDIE = 332.0716                       # COMMON/CONVERT.BLOCK:13
PI180 = 180.0 / (4.0 * math.atan(1.0))  # pot/init_var.f:8

# TIP3-style rigid water (O-H 0.9572 A, H-O-H 104.52 deg)
WAT_ROH = 0.9572
WAT_THETA = 104.52
WAT_QO = -0.834
WAT_QH = 0.417
WAT_EPS_O = 0.1521
WAT_RMIN2_O = 1.7682

# monomer flags handed to the library (derived from MOIL monomer names)
MONO_WATER = 1      # 'TIP3' or 'SPC*': water-group lists (pot/nbondm.f:81-82)
MONO_SQZ_FIRST = 2  # squeezed by its first atom (symm/squeeze.f:75-79)
MONO_SQZ_CENT = 4   # squeezed by its centroid (symm/squeeze.f:70-72)

_SQZ_FIRST_NAMES = {'TIP3', 'NAAT', 'NAO', 'CL', 'MG', 'CMO', 'CO', 'SUL',
                    'K', 'AR'}
_SQZ_CENT_NAMES = {'DMPC', 'DOPC', 'NATA'}


def _q(fmt: str, x: float) -> float:
    """Round `x` through a Fortran-style text field (see module doc)."""
    return float(format(float(x), fmt))


def lj_epsgm(eps: float, rmin_half: float):
    """MOIL's per-atom LJ factors (connect/fillpt.f:150-179, geometric):
    epsgm12 = 2*sqrt(eps)*sigma^6, epsgm6 = 2*sqrt(eps)*sigma^3 so that
    A_ij = epsgm12_i*epsgm12_j and B_ij = epsgm6_i*epsgm6_j."""
    if eps <= 0.0:
        return 0.0, 0.0
    sigma = 2.0 * rmin_half / 2.0 ** (1.0 / 6.0)
    e12 = 2.0 * math.sqrt(eps) * sigma ** 6
    e6 = 2.0 * math.sqrt(eps) * sigma ** 3
    return e12, e6


def to_arith(t: 'Topology', seed: int, v14: float = 0.5, dup_frac: float = 0.25) -> 'Topology':
    """The same system under MOIL's arith rule (its AMBER-style parameter
    sets): per-atom epsgm12 = sqrt(4 eps) and epsgm6 = sigma recovered from
    the geometric factors (connect/fillpt.f:136-179; an atom without LJ
    keeps epsgm12 = 0 and takes sigma 1 A), 1-4 coefficients A = v14 a
    sigma^12, B = v14 a sigma^6 from the arith pair rule with the 1-4
    charge products kept, torsions of one cosine each with period 1..4,
    amplitude k2 and phase2 = +-1 (ktors1, ktors3 and the other phases
    zero; pot/ephi.f:113-138) plus a second term of another period on a
    fraction dup_frac of them, as MOIL's AMBER reader lists several terms
    of one dihedral (connect/angtor.f:251-259).  Impropers keep their
    force constants (the arith form k - k cos 2 phi has no phase)."""
    rng = np.random.default_rng(seed)
    e12, e6 = t.epsgm12, t.epsgm6
    has = e12 > 0
    sig = np.ones_like(e6)
    sig[has] = (e12[has] / e6[has]) ** (1.0 / 3.0)
    a12 = np.zeros_like(e12)
    a12[has] = e6[has] ** 2 / e12[has]
    a12 = np.array([_q('12.5E', v) for v in a12])
    sig = np.array([_q('12.5E', v) for v in sig])
    p14 = t.p14.copy()
    for m in range(len(t.spec1)):
        i, j = int(t.spec1[m]) - 1, int(t.spec2[m]) - 1
        a = a12[i] * a12[j]
        sg = 0.5 * (sig[i] + sig[j])
        p14[m, 0] = _q('30.5f', v14 * a * sg ** 12)
        p14[m, 1] = _q('15.5f', v14 * a * sg ** 6)
    nt = len(t.itor1)
    per = rng.integers(1, 5, nt)
    dup = np.flatnonzero(rng.random(nt) < dup_frac)
    per2 = (per[dup] % 4) + 1
    rows = np.concatenate([np.arange(nt), dup])
    order = np.argsort(rows, kind='stable')
    rows = rows[order]
    period = np.concatenate([per, per2])[order].astype(np.int32)
    k2 = np.array([_q('10.4f', v) for v in rng.uniform(0.1, 2.0, len(rows))])
    ph2 = np.array([_q('6.2f', v) for v in rng.choice([-1.0, 1.0], len(rows))])
    z = np.zeros(len(rows))
    return dataclasses.replace(
        t, epsgm12=a12, epsgm6=sig, p14=p14,
        itor1=t.itor1[rows].copy(), itor2=t.itor2[rows].copy(), itor3=t.itor3[rows].copy(),
        itor4=t.itor4[rows].copy(), period=period, ktors1=z.copy(), ktors2=k2,
        ktors3=z.copy(), phase1=z.copy(), phase2=ph2, phase3=z.copy(), arith=True)


def monomer_flags(monames) -> np.ndarray:
    out = np.zeros(len(monames), np.int32)
    for k, nm in enumerate(monames):
        nm = nm.strip()
        f = 0
        if nm == 'TIP3' or nm[:3] == 'SPC':
            f |= MONO_WATER
        if nm in _SQZ_FIRST_NAMES or nm[:3] == 'SPC':
            f |= MONO_SQZ_FIRST
        if nm in _SQZ_CENT_NAMES:
            f |= MONO_SQZ_CENT
        out[k] = f
    return out


@dataclasses.dataclass
class Topology:
    """MOIL CONNECT-block content for the energy path (1-based indices)."""
    npt: int
    moname: List[str]            # per monomer, 4 chars
    poipt: np.ndarray            # int32 [totmon+1], poipt[0]=0, last atom
    ptnm: List[str]
    ptms: np.ndarray
    ptchg: np.ndarray
    epsgm6: np.ndarray
    epsgm12: np.ndarray
    exc1: np.ndarray             # int32 [npt+1], exc1[0]=0
    exc2: np.ndarray             # int32 [totex]
    ib1: np.ndarray
    ib2: np.ndarray
    kbond: np.ndarray
    req: np.ndarray
    iangl1: np.ndarray
    iangl2: np.ndarray
    iangl3: np.ndarray
    kangl: np.ndarray
    angleq: np.ndarray           # radians (as MOIL holds it after rconn)
    angleq_deg: np.ndarray       # the quantised degrees written to .con
    itor1: np.ndarray
    itor2: np.ndarray
    itor3: np.ndarray
    itor4: np.ndarray
    period: np.ndarray
    ktors1: np.ndarray
    ktors2: np.ndarray
    ktors3: np.ndarray
    phase1: np.ndarray
    phase2: np.ndarray
    phase3: np.ndarray
    iimp1: np.ndarray
    iimp2: np.ndarray
    iimp3: np.ndarray
    iimp4: np.ndarray
    kimp: np.ndarray
    impeq: np.ndarray            # radians
    impeq_deg: np.ndarray
    spec1: np.ndarray
    spec2: np.ndarray
    p14: np.ndarray              # float64 [totspe, 3]
    idxtip3: np.ndarray          # monomer numbers of waters (1-based)
    arith: bool = False          # MOIL's arith rule (the .con header flag)

    @property
    def totmon(self) -> int:
        return len(self.moname)

    @property
    def nwaters(self) -> int:
        return len(self.idxtip3)

    @property
    def flagchr(self) -> np.ndarray:
        # connect/chsort.f:20-24
        return (np.abs(self.ptchg) > 1e-6).astype(np.uint8)

    @property
    def monoflag(self) -> np.ndarray:
        return monomer_flags(self.moname)

    def net_charge(self) -> float:
        return float(np.sum(self.ptchg))


class _TopoBuilder:
    """Accumulates atoms/monomers/terms; every parameter is quantised."""

    def __init__(self):
        self.moname, self.poipt = [], [0]
        self.ptnm, self.ptms, self.ptchg, self.e6, self.e12 = [], [], [], [], []
        self.bonds, self.angles, self.tors, self.imps, self.spec = [], [], [], [], []
        self.excl = {}
        self.idxtip3 = []
        self.atype = []  # LJ type per atom (-1: no LJ; WATER_O_TYPE: water O)

    @property
    def npt(self):
        return len(self.ptnm)

    def add_monomer(self, name, atoms, types=None):
        """atoms: list of (name, mass, charge, eps, rmin_half); types: the
        atoms' LJ type numbers (default -1)."""
        self.atype.extend(types if types is not None else [-1] * len(atoms))
        for (nm, m, q, eps, rh) in atoms:
            e12, e6 = lj_epsgm(eps, rh)
            self.ptnm.append(nm)
            self.ptms.append(_q('7.3f', m))
            self.ptchg.append(_q('9.5f', q))
            self.e6.append(_q('12.5E', e6))
            self.e12.append(_q('12.5E', e12))
        self.moname.append(name)
        self.poipt.append(self.npt)
        if name == 'TIP3':
            self.idxtip3.append(len(self.moname))

    def bond(self, i, j, k, r0):
        self.bonds.append((i, j, _q('10.4f', k), _q('10.4f', r0)))

    def angle(self, i, j, k, kk, th_deg):
        self.angles.append((i, j, k, _q('10.5f', kk), _q('10.5f', th_deg)))

    def torsion(self, i, j, k, l, period, k1, k2, k3, p1, p2, p3):
        self.tors.append((i, j, k, l, period, _q('10.4f', k1), _q('10.4f', k2),
                          _q('10.4f', k3), _q('6.2f', p1), _q('6.2f', p2),
                          _q('6.2f', p3)))

    def improper(self, i, j, k, l, kk, th_deg):
        self.imps.append((i, j, k, l, _q('15.8E', kk), _q('15.8E', th_deg)))

    def exclude(self, i, j):
        lo, hi = (i, j) if i < j else (j, i)
        self.excl.setdefault(lo, set()).add(hi)

    def special14(self, i, j, v14, el14, eps=1.0, hvdw0=True):
        """1-4 pair parameters as generic/wconn.f:296-325 (geometric)."""
        e12i, e12j = self.e12[i - 1], self.e12[j - 1]
        if (e12i < 1e-3 or e12j < 1e-3) and hvdw0:
            a = b = 0.0
        else:
            a = e12i * e12j * v14
            b = self.e6[i - 1] * self.e6[j - 1] * v14
        c = self.ptchg[i - 1] * self.ptchg[j - 1] * eps * el14 * KOFDIE
        self.spec.append((i, j, _q('30.5f', a), _q('15.5f', b), _q('15.5f', c)))

    def build(self) -> Topology:
        npt = self.npt
        exc1 = np.zeros(npt + 1, np.int32)
        exc2 = []
        for i in range(1, npt + 1):
            lst = sorted(self.excl.get(i, ()))
            exc2.extend(lst)
            exc1[i] = len(exc2)
        i32 = lambda v: np.asarray(v, np.int32)
        f64 = lambda v: np.asarray(v, np.float64)
        B = self.bonds or []
        A = self.angles or []
        T = self.tors or []
        I = self.imps or []
        S = self.spec or []
        col = lambda L, k, n: [t[k] for t in L] if L else []
        adeg = f64(col(A, 4, 5))
        ideg = f64(col(I, 5, 6))
        return Topology(
            npt=npt, moname=list(self.moname), poipt=i32(self.poipt),
            ptnm=list(self.ptnm), ptms=f64(self.ptms), ptchg=f64(self.ptchg),
            epsgm6=f64(self.e6), epsgm12=f64(self.e12),
            exc1=exc1, exc2=i32(exc2),
            ib1=i32(col(B, 0, 4)), ib2=i32(col(B, 1, 4)), kbond=f64(col(B, 2, 4)),
            req=f64(col(B, 3, 4)),
            iangl1=i32(col(A, 0, 5)), iangl2=i32(col(A, 1, 5)), iangl3=i32(col(A, 2, 5)),
            kangl=f64(col(A, 3, 5)), angleq=adeg / PI180, angleq_deg=adeg,
            itor1=i32(col(T, 0, 11)), itor2=i32(col(T, 1, 11)), itor3=i32(col(T, 2, 11)),
            itor4=i32(col(T, 3, 11)), period=i32(col(T, 4, 11)),
            ktors1=f64(col(T, 5, 11)), ktors2=f64(col(T, 6, 11)), ktors3=f64(col(T, 7, 11)),
            phase1=f64(col(T, 8, 11)), phase2=f64(col(T, 9, 11)), phase3=f64(col(T, 10, 11)),
            iimp1=i32(col(I, 0, 6)), iimp2=i32(col(I, 1, 6)), iimp3=i32(col(I, 2, 6)),
            iimp4=i32(col(I, 3, 6)), kimp=f64(col(I, 4, 6)), impeq=ideg / PI180,
            impeq_deg=ideg,
            spec1=i32(col(S, 0, 5)), spec2=i32(col(S, 1, 5)),
            p14=f64([s[2:5] for s in S]).reshape(-1, 3),
            idxtip3=i32(self.idxtip3),
        )


def _water_frame():
    half = math.radians(WAT_THETA) / 2.0
    o = np.zeros(3)
    h1 = np.array([WAT_ROH * math.sin(half), 0.0, WAT_ROH * math.cos(half)])
    h2 = np.array([-WAT_ROH * math.sin(half), 0.0, WAT_ROH * math.cos(half)])
    return np.stack([o, h1, h2])


def _random_rotations(rng, n):
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


WATER_O_TYPE = -2  # LJ type marker of the water O (one shared type)


def _add_water(tb: _TopoBuilder):
    base = tb.npt + 1
    tb.add_monomer('TIP3', [('OH2', 15.999, WAT_QO, WAT_EPS_O, WAT_RMIN2_O),
                            ('H1', 1.008, WAT_QH, 0.0, 0.0),
                            ('H2', 1.008, WAT_QH, 0.0, 0.0)], [WATER_O_TYPE, -1, -1])
    o, h1, h2 = base, base + 1, base + 2
    tb.bond(o, h1, 450.0, WAT_ROH)
    tb.bond(o, h2, 450.0, WAT_ROH)
    tb.angle(h1, o, h2, 55.0, WAT_THETA)
    tb.exclude(o, h1)
    tb.exclude(o, h2)
    tb.exclude(h1, h2)


def water_positions(centres, rng):
    frame = _water_frame()
    R = _random_rotations(rng, len(centres))
    pos = centres[:, None, :] + np.einsum('nij,kj->nki', R, frame)
    return pos.reshape(-1, 3)


@dataclasses.dataclass
class System:
    topo: Topology
    coor: np.ndarray          # float64 [npt, 3]
    box: tuple                # (a, b, c)
    meta: dict


def water_box(n_side: int, L: float, seed: int, jitter: float = 0.3) -> System:
    """W = n_side^3 rigid waters: O on an n_side^3 lattice of the cubic
    box [-L/2, L/2]^3 plus U(-jitter, jitter) per coordinate, uniformly
    random orientations."""
    ss = np.random.SeedSequence(seed)
    r_jit, r_rot = [np.random.default_rng(s) for s in ss.spawn(2)]
    sp = L / n_side
    g = (np.arange(n_side) + 0.5) * sp - 0.5 * L
    X, Y, Z = np.meshgrid(g, g, g, indexing='ij')
    cen = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    cen = cen + r_jit.uniform(-jitter, jitter, cen.shape)
    coor = water_positions(cen, r_rot)
    tb = _TopoBuilder()
    for _ in range(len(cen)):
        _add_water(tb)
    topo = tb.build()
    return System(topo, coor, (L, L, L),
                  dict(kind='water_box', n_side=n_side, L=L, seed=seed, jitter=jitter))


# ---------------------------------------------------------------------------
# Heteropolymer in water (random-walk chains + 3-site water)
# ---------------------------------------------------------------------------

def _rand_unit(rng):
    v = rng.standard_normal(3)
    return v / np.linalg.norm(v)


def _grow_chain(rng, n, start, L, dmin=0.0):
    """Random walk with bond U(1.0,1.5) A, angle U(100,120) deg and
    uniform torsion; with dmin > 0 each step is redrawn (up to 100 times)
    until the new atom is at least dmin from every atom four or more bonds
    back along the chain."""
    pts = [start]
    d = _rand_unit(rng)
    pts.append(start + rng.uniform(1.0, 1.5) * d)
    for k in range(2, n):
        for _try in range(100):
            x = _chain_step(rng, pts, k)
            if dmin <= 0 or k < 4:
                break
            d2 = ((np.array(pts[:k - 3]) - x) ** 2).sum(1)
            if d2.min() >= dmin * dmin:
                break
        pts.append(x)
    return np.array(pts)


def _chain_step(rng, pts, k):
    """Next atom of the walk: bond U(1.0,1.5) A, angle U(100,120) deg,
    uniform torsion."""
    b = rng.uniform(1.0, 1.5)
    th = math.radians(rng.uniform(100.0, 120.0))
    phi = rng.uniform(-math.pi, math.pi)
    p1, p2 = pts[-2], pts[-1]
    u = p2 - p1
    u /= np.linalg.norm(u)
    ref = _rand_unit(rng) if k == 2 else pts[-3] - p1
    n1 = np.cross(u, ref)
    if np.linalg.norm(n1) < 1e-8:
        n1 = np.cross(u, _rand_unit(rng))
    n1 /= np.linalg.norm(n1)
    n2 = np.cross(n1, u)
    dvec = (-math.cos(th) * u + math.sin(th) * (math.cos(phi) * n2 + math.sin(phi) * n1))
    return p2 + b * dvec


class _CellHash:
    """Periodic cell hash of placed atoms for minimum-image contact checks."""

    def __init__(self, L, h):
        self.L = L
        self.n = max(1, int(L // h))
        self.h = L / self.n
        self.cells = {}

    def _key(self, x):
        return tuple(int(np.floor((x[k] + 0.5 * self.L) / self.h)) % self.n for k in range(3))

    def add(self, x):
        self.cells.setdefault(self._key(x), []).append(np.array(x, float))

    def too_close(self, x, dmin):
        k = self._key(x)
        d2min = dmin * dmin
        r = range(-1, 2) if self.n >= 3 else range(0, self.n)
        for a in r:
            for b in r:
                for c in r:
                    for y in self.cells.get(((k[0] + a) % self.n, (k[1] + b) % self.n,
                                             (k[2] + c) % self.n), ()):
                        d = x - y
                        d -= self.L * np.round(d / self.L)
                        if float(d @ d) < d2min:
                            return True
        return False


def _chain_contacts_ok(xyz, hashed, L, dmin):
    """True if no atom of the chain is closer than dmin (minimum image) to
    an atom of a chain already placed, or to an atom of its own chain four
    or more bonds away."""
    for a in range(len(xyz)):
        if hashed.too_close(xyz[a], dmin):
            return False
        for b in range(a + 4, len(xyz)):
            d = xyz[a] - xyz[b]
            d -= L * np.round(d / L)
            if float(d @ d) < dmin * dmin:
                return False
    return True


def _chain_topology(tb, rng_par, xyz, c, types_eps, types_rmh, n_types, mono_size,
                    frac_neutral, frac_hlike, imp_general=False):
    """Atoms, monomers and bonded terms of one chain at coordinates xyz
    (chain number c): charges N(0,0.4) shifted to net zero (quantised
    exactly to zero per chain), LJ types, bonds/angles at their current
    geometry, torsions, impropers and the 1-2/1-3/1-4 exclusions.
    Improper phases are 0 or 180 deg, or with imp_general drawn from
    +-U(10, 170) deg (MOIL's general k (phi - phi0)^2 branch,
    pot/eimphi.f:188-249).  Returns the chain's 1-based atom numbers."""
    n = len(xyz)
    q = rng_par.normal(0.0, 0.4, n)
    neutral = rng_par.random(n) < frac_neutral
    neutral[-1] = False  # the last atom absorbs the quantisation residual
    q[neutral] = 0.0
    q[~neutral] -= q[~neutral].mean()
    q = np.array([_q('9.5f', v) for v in q])
    # make the quantised charges sum to exactly zero per chain
    resid = round(float(np.sum(q)) * 1e5)
    q[-1] = _q('9.5f', q[-1] - resid * 1e-5)
    typ = rng_par.integers(0, n_types, n)
    hlike = rng_par.random(n) < frac_hlike
    base = tb.npt + 1
    k = 0
    while k < n:
        m = min(mono_size, n - k)
        atoms = [('C%d' % ((k + t) % 10), 12.011, q[k + t],
                  0.0 if hlike[k + t] else types_eps[typ[k + t]],
                  types_rmh[typ[k + t]]) for t in range(m)]
        tb.add_monomer('PLY%d' % (c % 10) if m == mono_size else 'PLE%d' % (c % 10), atoms,
                       [-1 if hlike[k + t] else int(typ[k + t]) for t in range(m)])
        k += m
    idx = np.arange(base, base + n)
    for a in range(n - 1):
        tb.bond(idx[a], idx[a + 1], rng_par.uniform(300, 500),
                float(np.linalg.norm(xyz[a + 1] - xyz[a])))
        tb.exclude(idx[a], idx[a + 1])
    for a in range(n - 2):
        v1 = xyz[a] - xyz[a + 1]
        v2 = xyz[a + 2] - xyz[a + 1]
        th = math.degrees(math.acos(np.dot(v1, v2) / np.linalg.norm(v1) / np.linalg.norm(v2)))
        tb.angle(idx[a], idx[a + 1], idx[a + 2], rng_par.uniform(50, 100), th)
        tb.exclude(idx[a], idx[a + 2])
    for a in range(n - 3):
        per = int(rng_par.integers(1, 4))
        kk = [0.0, 0.0, 0.0]
        kk[per - 1] = rng_par.uniform(0.0, 2.0)
        ph = [1.0, 1.0, 1.0]
        ph[per - 1] = float(rng_par.choice([-1.0, 1.0]))
        tb.torsion(idx[a], idx[a + 1], idx[a + 2], idx[a + 3], per, kk[0], kk[1], kk[2],
                   ph[0], ph[1], ph[2])
        tb.exclude(idx[a], idx[a + 3])
        kimp = rng_par.uniform(5.0, 20.0)
        if imp_general:
            ph0 = float(rng_par.uniform(10.0, 170.0) * rng_par.choice([-1.0, 1.0]))
        else:
            ph0 = float(rng_par.choice([0.0, 180.0]))
        tb.improper(idx[a], idx[a + 1], idx[a + 2], idx[a + 3], kimp, ph0)
    return idx


def hetero_system(n_chains: int, L: float, seed: int, n_water: Optional[int] = None,
                  chain_len=(10, 20), n_types: int = 24, mono_size: int = 4,
                  v14: float = 0.5, el14: float = 0.5,
                  water_min_dist: float = 2.5, chain_min_dist: float = 3.0,
                  frac_neutral: float = 0.0, frac_hlike: float = 0.0,
                  imp_general: bool = False) -> System:
    """Random-walk heteropolymer chains plus 3-site water in a cubic box
    (centred at the origin).  Bonds/angles/proper torsions from chain
    topology (k_b U(300,500), k_theta U(50,100), torsion period n in
    {1,2,3} with barrier U(0,2)), one improper per interior quadruple
    (i-1,i,i+1,i+2) with phi0 = 0 or 180 deg, n_types LJ types with
    eps~U(0.02,0.2), Rmin/2~U(1.0,2.1); charges N(0,0.4) shifted to net
    zero per chain.  Monomers hold <= mono_size consecutive chain atoms.
    Exclusions 1-2, 1-3 and 1-4 (MOIL excludes 1-4 and re-adds them via
    the scaled special list, connect/angtor.f:483-577).

    Chains are redrawn until no chain atom is closer than chain_min_dist
    (minimum image) to another chain atom four or more bonds away.
    frac_neutral of the chain atoms get charge 0 (uncharged: MOIL's list2)
    and frac_hlike get zero LJ (MOIL's hydrogen rule, pot/nbond.f:115-126);
    both default to 0, the plain heteropolymer recipe.  imp_general draws
    general improper phases (see _chain_topology)."""
    ss = np.random.SeedSequence(seed)
    rng_walk, rng_par, rng_wat, rng_rot = [np.random.default_rng(s) for s in ss.spawn(4)]
    types_eps = rng_par.uniform(0.02, 0.2, n_types)
    types_rmh = rng_par.uniform(1.0, 2.1, n_types)
    tb = _TopoBuilder()
    chain_xyz = []
    chain_atoms = []
    hashed = _CellHash(L, max(chain_min_dist, 1.0))
    for c in range(n_chains):
        n = int(rng_walk.integers(chain_len[0], chain_len[1] + 1))
        for _attempt in range(200):
            start = rng_walk.uniform(-0.5 * L + 3.0, 0.5 * L - 3.0, 3)
            xyz = _grow_chain(rng_walk, n, start, L, dmin=chain_min_dist)
            # keep the chain's atoms inside the primary box
            lo, hi = xyz.min(0), xyz.max(0)
            shift = np.where(hi > 0.5 * L - 0.5, 0.5 * L - 0.5 - hi, 0.0)
            shift = np.where(lo + shift < -0.5 * L + 0.5, -0.5 * L + 0.5 - lo, shift)
            xyz = xyz + shift
            if chain_min_dist <= 0 or _chain_contacts_ok(xyz, hashed, L, chain_min_dist):
                break
        else:
            raise RuntimeError('hetero_system: no contact-free placement for chain %d' % c)
        for x in xyz:
            hashed.add(x)
        idx = _chain_topology(tb, rng_par, xyz, c, types_eps, types_rmh, n_types, mono_size,
                              frac_neutral, frac_hlike, imp_general)
        chain_xyz.append(xyz)
        chain_atoms.append(idx)
    # 1-4 specials (scaled) for torsion end atoms
    seen = set()
    for idx in chain_atoms:
        for a in range(len(idx) - 3):
            key = (int(idx[a]), int(idx[a + 3]))
            if key not in seen:
                seen.add(key)
                tb.special14(key[0], key[1], v14, el14)
    poly = np.concatenate(chain_xyz) if chain_xyz else np.zeros((0, 3))
    # water: lattice sites at ~ liquid spacing, dropped near polymer atoms
    nw_target = n_water
    if nw_target is None:
        nw_target = int(round((L ** 3 - 0.0) * 0.0334)) - len(poly) // 3
    n_side = int(math.ceil(nw_target ** (1.0 / 3.0) * 1.15)) + 1
    sp = L / n_side
    g = (np.arange(n_side) + 0.5) * sp - 0.5 * L
    X, Y, Z = np.meshgrid(g, g, g, indexing='ij')
    cen = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    cen = cen + rng_wat.uniform(-0.2, 0.2, cen.shape)
    if len(poly):
        keep = np.ones(len(cen), bool)
        cell = max(water_min_dist, 1.0)
        # brute-force in chunks (sizes here are modest)
        for s in range(0, len(cen), 4096):
            d2 = ((cen[s:s + 4096, None, :] - poly[None, :, :]) ** 2).sum(-1)
            keep[s:s + 4096] = d2.min(1) >= water_min_dist ** 2
        cen = cen[keep]
    order = rng_wat.permutation(len(cen))[:nw_target]
    cen = cen[np.sort(order)]
    wxyz = water_positions(cen, rng_rot)
    for _ in range(len(cen)):
        _add_water(tb)
    topo = tb.build()
    coor = np.concatenate([poly, wxyz]) if len(poly) else wxyz
    return System(topo, coor, (L, L, L),
                  dict(kind='hetero', n_chains=n_chains, L=L, seed=seed, n_water=len(cen)))


def _wrap01(x, L):
    """Coordinates of the centred box mapped into [0, L) (periodic)."""
    w = np.mod(x + 0.5 * L, L)
    return np.where(w >= L, w - L, w)


def _unit_rows(rng, shape):
    v = rng.standard_normal(shape + (3,))
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _grow_chains_packed(rng, lengths, L, dmin, ncand=16, retries=12, groups=8):
    """All chains grown together, one atom per chain per step, by the
    random-walk rule of _chain_step (bond U(1.0,1.5) A, angle U(100,120)
    deg, uniform torsion).  Each step draws ncand candidates per chain and
    keeps the first one at least dmin (minimum image) from every atom
    already placed except the chain's own last three; chains with no such
    candidate draw new sets (up to `retries` times) and finally keep the
    candidate farthest from them.  The atoms of one step are placed in
    `groups` random groups, each checked against the atoms of the earlier
    groups too; inside a group, the later chain of a pair closer than dmin
    moves to another qualifying candidate.  Returns a list of (n_i, 3)
    arrays (chains continuous, not wrapped)."""
    from scipy.spatial import cKDTree
    lengths = np.asarray(lengths)
    nc, nmax = len(lengths), int(lengths.max())
    X = np.zeros((nc, nmax, 3))
    P = np.zeros((0, 3))                 # placed atoms, wrapped
    owner = np.zeros(0, np.int64)        # chain of each placed atom
    pos = np.zeros(0, np.int64)          # its position along the chain

    def candidates(ch, k):
        n = len(ch)
        if k == 0:
            return rng.uniform(-0.5 * L, 0.5 * L, (n, ncand, 3))
        b = rng.uniform(1.0, 1.5, (n, ncand))
        if k == 1:
            d = _unit_rows(rng, (n, ncand))
        else:
            th = np.radians(rng.uniform(100.0, 120.0, (n, ncand)))
            phi = rng.uniform(-math.pi, math.pi, (n, ncand))
            p1, p2 = X[ch, k - 2], X[ch, k - 1]
            u = p2 - p1
            u /= np.linalg.norm(u, axis=1, keepdims=True)
            ref = X[ch, k - 3] - p1 if k >= 3 else _unit_rows(rng, (n,))
            n1 = np.cross(u, ref)
            nn = np.linalg.norm(n1, axis=1)
            bad = nn < 1e-8
            if bad.any():
                n1[bad] = np.cross(u[bad], _unit_rows(rng, (int(bad.sum()),)))
                nn[bad] = np.linalg.norm(n1[bad], axis=1)
            n1 /= nn[:, None]
            n2 = np.cross(n1, u)
            d = (-np.cos(th)[..., None] * u[:, None, :] + np.sin(th)[..., None] *
                 (np.cos(phi)[..., None] * n2[:, None, :] + np.sin(phi)[..., None] *
                  n1[:, None, :]))
        return X[ch, k - 1][:, None, :] + b[..., None] * d

    def nearest(tree, ch, cand, k, own_check):
        n = len(ch)
        if tree is None:
            return np.full((n, ncand), np.inf)
        npts = tree.n
        dist, idx = tree.query(_wrap01(cand.reshape(-1, 3), L), k=4, distance_upper_bound=dmin)
        dist = dist.reshape(n, ncand, 4)
        idx = idx.reshape(n, ncand, 4)
        hit = idx < npts
        if own_check:
            safe_idx = np.where(hit, idx, 0)
            own = hit & (owner[safe_idx] == ch[:, None, None]) & (pos[safe_idx] >= k - 3)
            hit = hit & ~own
        return np.where(hit, dist, np.inf).min(axis=2)

    for k in range(nmax):
        act = np.nonzero(lengths > k)[0]
        tree = cKDTree(P, boxsize=L) if len(P) else None
        # the step's atoms are placed in groups; each group is checked
        # against the atoms placed before the step and by earlier groups
        perm = rng.permutation(len(act))
        step_pts = np.zeros((0, 3))
        for grp in np.array_split(perm, min(groups, len(act))):
            ch = act[np.sort(grp)]
            na = len(ch)

            def near_of(chs, cc):
                nr = nearest(tree, chs, cc, k, True)
                if len(step_pts):
                    nr = np.minimum(nr, nearest(cKDTree(step_pts, boxsize=L), chs, cc, k, False))
                return nr
            cand = candidates(ch, k)
            near = near_of(ch, cand)
            for _ in range(retries):
                stuck = np.nonzero(~(near >= dmin).any(axis=1))[0]
                if len(stuck) == 0:
                    break
                c2 = candidates(ch[stuck], k)
                n2 = near_of(ch[stuck], c2)
                better = n2.max(axis=1) > near[stuck].max(axis=1)
                cand[stuck[better]] = c2[better]
                near[stuck[better]] = n2[better]
            okc = near >= dmin
            choice = np.where(okc.any(axis=1), okc.argmax(axis=1), near.argmax(axis=1))
            new = cand[np.arange(na), choice]
            # conflicts inside the group: move the later chain of a pair to
            # another qualifying candidate
            tried = np.zeros((na, ncand), bool)
            tried[np.arange(na), choice] = True
            for _ in range(8):
                pairs = cKDTree(_wrap01(new, L), boxsize=L).query_pairs(dmin, output_type='ndarray')
                if len(pairs) == 0:
                    break
                for m in np.unique(pairs.max(axis=1)):
                    alt = np.nonzero(okc[m] & ~tried[m])[0]
                    if len(alt) == 0:
                        continue
                    tried[m, alt[0]] = True
                    choice[m] = alt[0]
                    new[m] = cand[m, alt[0]]
            X[ch, k] = new
            step_pts = np.concatenate([step_pts, _wrap01(new, L)])
            P = np.concatenate([P, _wrap01(new, L)])
            owner = np.concatenate([owner, ch])
            pos = np.concatenate([pos, np.full(na, k)])
    out = []
    for c in range(nc):
        xyz = X[c, :lengths[c]].copy()
        # whole-box shifts only (distances unchanged): centroid in the box
        cen = xyz.mean(axis=0)
        xyz -= L * np.round(cen / L)
        out.append(xyz)
    return out


def _lj_type_table(tb, types_eps, types_rmh):
    """LJ types of a heteropolymer system: the chain types, then the water
    O as one more type; per atom its type number (-1: no LJ)."""
    nt = len(types_eps)
    at = np.asarray(tb.atype, np.int64)
    at = np.where(at == WATER_O_TYPE, nt, at)
    eps = np.concatenate([np.asarray(types_eps, float), [WAT_EPS_O]])
    rmh = np.concatenate([np.asarray(types_rmh, float), [WAT_RMIN2_O]])
    return dict(lj_type=at, lj_eps=eps, lj_rmin_half=rmh)


def hetero_packed(n_chains: int, L: float, seed: int, n_atoms: Optional[int] = None,
                  chain_len=(10, 20), n_types: int = 24, mono_size: int = 4,
                  v14: float = 0.5, el14: float = 0.5, chain_min_dist: float = 2.0,
                  water_min_dist: float = 2.2) -> System:
    """The heteropolymer-in-water recipe at scale (hundreds of thousands of
    chain atoms): n_chains random-walk chains of U{10..20} atoms with the
    same topology rules as hetero_system, grown together by
    _grow_chains_packed with chain_min_dist between non-neighbouring chain
    atoms; the rest of the box filled with 3-site waters on a jittered
    lattice whose spacing is chosen so that enough sites lie at least
    water_min_dist from every chain atom (minimum image) to reach n_atoms
    atoms in total (default: 0.1006 atoms/A^3)."""
    from scipy.spatial import cKDTree
    ss = np.random.SeedSequence(seed)
    rng_walk, rng_par, rng_wat, rng_rot = [np.random.default_rng(s) for s in ss.spawn(4)]
    types_eps = rng_par.uniform(0.02, 0.2, n_types)
    types_rmh = rng_par.uniform(1.0, 2.1, n_types)
    lengths = rng_walk.integers(chain_len[0], chain_len[1] + 1, n_chains)
    chains = _grow_chains_packed(rng_walk, lengths, L, chain_min_dist)
    tb = _TopoBuilder()
    chain_atoms = []
    for c, xyz in enumerate(chains):
        chain_atoms.append(_chain_topology(tb, rng_par, xyz, c, types_eps, types_rmh, n_types,
                                           mono_size, 0.0, 0.0))
    for idx in chain_atoms:
        for a in range(len(idx) - 3):
            tb.special14(int(idx[a]), int(idx[a + 3]), v14, el14)
    poly = np.concatenate(chains)
    if n_atoms is None:
        n_atoms = int(round(0.1006 * L ** 3))
    nw = max(0, (n_atoms - len(poly)) // 3)
    tree = cKDTree(_wrap01(poly, L), boxsize=L)
    n_side = max(2, int(math.ceil((nw * 1.2) ** (1.0 / 3.0))))
    while True:
        sp = L / n_side
        g = (np.arange(n_side) + 0.5) * sp - 0.5 * L
        X, Y, Z = np.meshgrid(g, g, g, indexing='ij')
        cen = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
        cen = cen + rng_wat.uniform(-0.2, 0.2, cen.shape)
        dist, _ = tree.query(_wrap01(cen, L), k=1, distance_upper_bound=water_min_dist)
        free = cen[~np.isfinite(dist)]
        if len(free) >= nw:
            break
        n_side += max(1, n_side // 20)
    order = rng_wat.permutation(len(free))[:nw]
    cen = free[np.sort(order)]
    wxyz = water_positions(cen, rng_rot)
    for _ in range(len(cen)):
        _add_water(tb)
    topo = tb.build()
    coor = np.concatenate([poly, wxyz])
    return System(topo, coor, (L, L, L),
                  dict(kind='hetero_packed', n_chains=n_chains, L=L, seed=seed, n_water=len(cen),
                       water_spacing=sp, **_lj_type_table(tb, types_eps, types_rmh)))


# ---------------------------------------------------------------------------
# MOIL .con writer (layout read by generic/rconn.f:46-497)
# ---------------------------------------------------------------------------

def _fmt_rows(vals, per, fmt):
    out = []
    for s in range(0, len(vals), per):
        out.append(''.join(fmt % v for v in vals[s:s + per]))
    return out


def write_con(path: str, t: Topology, arith: Optional[bool] = None):
    L = []
    L.append('12.1')
    L.append('~ CONNECTIVITY FILE FOR MOLECULES: ')
    L.append('~ totmon npt nb nangl ntors nimp totex totspe lestyp NBULK nmb nchg nbeta nwaters arith prll mdivyes')
    nchg = int(np.sum(t.flagchr))
    hdr = [t.totmon, t.npt, len(t.ib1), len(t.iangl1), len(t.itor1), len(t.iimp1),
           len(t.exc2), len(t.spec1), 0, 1, 0, nchg, 0, t.nwaters]
    if arith is None:
        arith = t.arith
    L.append(' ' + ''.join('%8d' % v for v in hdr) + ' %s F F' % ('T' if arith else 'F'))
    L.append('SYNT')
    L.append('~ Pointers to last particle of BULK')
    L.append(' %7d' % t.npt)
    L.append('~ Monomer names ')
    L.extend(_fmt_rows([('%-4s' % m)[:4] for m in t.moname], 10, ' %s'))
    L.append('~ Pointers to last particle of monomer')
    L.extend(_fmt_rows(list(t.poipt[1:]), 10, ' %7d'))
    L.append('~ Properties of particles list : ')
    L.append('~pt mono ptid lesid mutaid ptnm   ptms   ptchg   epsgm6 epsgm12 ptwei ptsaid hbond')
    mono_of = np.searchsorted(t.poipt[1:], np.arange(1, t.npt + 1))  # 0-based monomer
    for i in range(t.npt):
        L.append(' %7d %7d %3d %3d %3d %-4s %7.3f  %9.5f %12.5E%12.5E%10.3E %3d  ' % (
            i + 1, mono_of[i] + 1, 1, 0, 0, t.ptnm[i][:4], t.ptms[i], t.ptchg[i],
            t.epsgm6[i], t.epsgm12[i], 1.0, 0))
    if len(t.ib1):
        L.append('~ Bonds list: ')
        L.append('~ ib1 ib2 kbond req ')
        for a, b, k, r in zip(t.ib1, t.ib2, t.kbond, t.req):
            L.append(' %7d %7d %10.4f %10.4f' % (a, b, k, r))
    if len(t.iangl1):
        L.append('~ Angles list: ')
        L.append('~ iangl1 iangl2 iangl3 kangl angleq ')
        for a, b, c, k, th in zip(t.iangl1, t.iangl2, t.iangl3, t.kangl, t.angleq_deg):
            L.append(' %7d %7d %7d %10.5f %10.5f' % (a, b, c, k, th))
    if len(t.itor1):
        L.append('~ Torsions list: ')
        L.append('~ itor1 itor2 itor3 itor4 period ktors1 ktors2 ktors3 phase1 phase2 phase3 ')
        for r in zip(t.itor1, t.itor2, t.itor3, t.itor4, t.period, t.ktors1, t.ktors2,
                     t.ktors3, t.phase1, t.phase2, t.phase3):
            L.append(' %7d %7d %7d %7d %7d %10.4f %10.4f %10.4f %6.2f%6.2f%6.2f' % r)
    if len(t.iimp1):
        L.append('~ Improper torsion properties: ')
        L.append('~iimp1 iimp2 iimp3 iimp4 kimp impeq ')
        for r in zip(t.iimp1, t.iimp2, t.iimp3, t.iimp4, t.kimp, t.impeq_deg):
            L.append(' %7d %7d %7d %7d %15.8E %15.8E' % r)
    if len(t.exc2):
        L.append('~ Exclusion list 1-2 1-3 1-4, set as followed:')
        L.append('~ atom number, number of exclusions and list')
        for i in range(1, t.npt + 1):
            n = int(t.exc1[i] - t.exc1[i - 1])
            if n > 0:
                L.append(' %7d %7d' % (i, n))
                L.extend(_fmt_rows(list(t.exc2[t.exc1[i - 1]:t.exc1[i]]), 10, ' %7d'))
    if len(t.spec1):
        L.append('~ Special list 1-4  set as followed:')
        L.append('~ atom number, atom number, esp12*v14 eps6*v14 qiqj*el14 ')
        for a, b, p in zip(t.spec1, t.spec2, t.p14):
            L.append(' %7d %7d %30.5f %15.5f %15.5f' % (a, b, p[0], p[1], p[2]))
    L.append('~ Charge flags. T = charged atom , F = uncharged')
    fl = ['T' if f else 'F' for f in t.flagchr]
    L.extend(_fmt_rows(fl, 15, ' %s'))
    L.append('~ LES copy labels ')
    L.extend([' ' + r for r in _fmt_rows([0] * t.npt, 15, '%6d')])
    if t.nwaters:
        L.append(' ~ The monomers divisions below are TIP3/SPC/SPCE waters')
        L.extend(_fmt_rows(list(t.idxtip3), 10, '%8d'))
    with open(path, 'w') as f:
        f.write('\n'.join(L) + '\n')
