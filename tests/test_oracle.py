"""The CPU oracle (oracle/moil_oracle.c) pinned against MOIL itself and
against closed forms.

MOIL ships no tests or golden data for this path; the golden fixtures
under tests/golden were written by MOIL's own Fortran (oracle/_ref, built
from the reference sources) with tests/golden/gen_golden.py.
"""
import math

import numpy as np
import pytest
from scipy.special import erfc

from moil_fnrfd_amd import abi, configs
from oracle.oracle import Oracle
from tests._util import (energy_errors, golden_energies, load_golden, rms_rel, rows_sha,
                         sort_rows)
from tests._cases import CASES, make_case
from tests.golden.gen_golden import topo_checksum

def _case(name):
    return make_case(name)


@pytest.mark.parametrize('name', sorted(CASES))
def test_oracle_matches_moil(name):
    g = load_golden(name)
    s, cut, ew, sqz = _case(name)
    assert topo_checksum(s.topo) == bytes(g['topo_sha']).decode()
    if 'coor' in g:
        assert np.array_equal(s.coor, g['coor'])
    o = Oracle(s.topo, configs.params_for(s, cut, ew))
    x = o.squeeze(s.coor) if sqz else s.coor
    if 'coor_after' in g:
        assert np.array_equal(x, g['coor_after'])
    o.build_lists(x)
    rows = sort_rows(o.water_pairs())
    if 'pairs' in g:
        assert np.array_equal(rows, g['pairs'])
    assert rows_sha(rows) == bytes(g['pairs_sha']).decode()
    arows = sort_rows(o.atom_pairs())
    if 'apairs' in g:
        assert np.array_equal(arows, g['apairs'])
    assert rows_sha(arows) == bytes(g['apairs_sha']).decode()
    e, dp = o.eforce(x)
    ref = golden_energies(g)
    info = o.ewald_info()
    assert info['beta'] == ref['ewaldcof']          # bit-identical beta
    assert info['e_self'] == ref['e_self']
    err = energy_errors(e, ref)
    assert max(err.values()) < 1e-12, err
    assert rms_rel(dp, g['dpot']) < 1e-10


@pytest.mark.slow
def test_oracle_matches_moil_32k():
    g = load_golden('water32k')
    s, cut, ew = configs.water32k()
    o = Oracle(s.topo, configs.params_for(s, cut, ew))
    o.build_lists(s.coor)
    assert rows_sha(o.water_pairs()) == bytes(g['pairs_sha']).decode()
    e, dp = o.eforce(s.coor)
    err = energy_errors(e, golden_energies(g))
    assert max(err.values()) < 1e-12, err
    assert rms_rel(dp[g['dpot_idx']], g['dpot']) < 1e-10


def test_erfc_table_closed_form():
    """Table nodes are erfc(i*del); the spline reproduces erfc and its
    derivative between nodes (error bound (5/384) h^4 max|f''''|)."""
    s, cut, ew = configs.water216()
    o = Oracle(s.topo, configs.params_for(s, cut, ew))
    info = o.ewald_info()
    tab, n = info['erf_arr'], info['ntab']
    dl = 1.0 / 5000.0
    x = np.arange(n) * dl
    assert np.max(np.abs(tab[:, 0] - erfc(x)) / erfc(x)) < 1e-14  # Cody's erfc vs scipy's
    # mid-interval values against scipy
    xm = (np.arange(n - 1) + 0.37) * dl
    i = (5000.0 * xm).astype(int)
    dx = xm - i * dl
    t = tab[i]
    val = dx * (t[:, 1] + 0.5 * dx * (t[:, 2] + dx * t[:, 3] / 3.0)) + t[:, 0]
    der = -t[:, 1] - dx * (t[:, 2] + 0.5 * dx * t[:, 3])
    assert np.max(np.abs(val - erfc(xm))) < 1e-15
    assert np.max(np.abs(der - 2 / math.sqrt(math.pi) * np.exp(-xm ** 2))) < 1e-11
    assert int(load_golden('water216')['erftab_rows'][0]) == n


def _tab_sha(tab):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(tab, np.float64).tobytes()).hexdigest()


@pytest.mark.parametrize('name', sorted(CASES))
def test_erfc_table_identical_to_moil(name):
    """The whole erfc spline table -- Cody's erfc at the nodes (the
    oracle's restatement of symm/erfcfun.f) and de Boor's clamped spline
    (symm/ewald_init.f:418-574) -- is MOIL's table byte for byte, for the
    oracle and for the library's host set-up (moil_ewald_setup, the
    numbers moil_gpu_init uploads), and beta is MOIL's to the bit."""
    from moil_fnrfd_amd import eforce
    g = load_golden(name)
    s, cut, ew, sqz = _case(name)
    info = Oracle(s.topo, configs.params_for(s, cut, ew)).ewald_info()
    beta, dens, tab = eforce.ewald_setup(cut['relx'], ew['dtol'])
    ref = golden_energies(g)
    for t, b in ((info['erf_arr'], info['beta']), (tab, beta)):
        assert b == ref['ewaldcof']
        assert len(t) == int(g['erftab_rows'][0])
        assert np.array_equal(t[:64], g['erftab_head'])
        assert np.array_equal(t[::97], g['erftab_sample'])
        assert _tab_sha(t) == bytes(g['erftab_sha']).decode()
    assert dens == ref['erftbdns']


def test_cody_erfc_against_scipy():
    """The library's erfc (Cody's three rational approximations, the table
    nodes) against scipy on each of its intervals: |x| <= 0.46875 (erf),
    0.46875..4, and above 4 (rational in 1/x^2, where the relative error of
    exp(-x^2) grows with x^2)."""
    from moil_fnrfd_amd import eforce
    # dtol below MOIL's floor is raised to 1e-10 (symm/ewald_init.f:78-79):
    # beta 0.48 at rc 9, a table reaching x = 6.5 across every interval
    b, dens, tab = eforce.ewald_setup(9.0, 1e-300)
    x = np.arange(len(tab)) / dens
    assert x[-1] > 6.0
    rel = np.abs(tab[:, 0] - erfc(x)) / erfc(x)
    for lo, hi, bar in ((0.0, 0.46875, 1e-15), (0.46875, 4.0, 1e-14), (4.0, 7.0, 3e-14)):
        m = (x >= lo) & (x < hi)
        assert m.sum() > 100 and rel[m].max() < bar, (lo, hi, rel[m].max())
    b2, _, _ = eforce.ewald_setup(9.0, 1e-10)
    assert b2 == b


def test_beta_closed_form():
    """beta solves erfc(beta rc)/rc = dtol (symm/ewald_init.f:298-335)."""
    for mk in (configs.water216, configs.water32k):
        s, cut, ew = mk()
        o = Oracle(s.topo, configs.params_for(s, cut, ew))
        b = o.ewald_info()['beta']
        rc = cut['relx']
        assert abs(erfc(b * rc) / rc - ew['dtol']) / ew['dtol'] < 1e-13


def test_e_self_closed_form():
    s, cut, ew = configs.water216()
    o = Oracle(s.topo, configs.params_for(s, cut, ew))
    info = o.ewald_info()
    es = -info['beta'] / math.sqrt(math.pi) * 332.0716 * float(np.sum(s.topo.ptchg ** 2))
    assert abs(info['e_self'] - es) < 1e-12 * abs(es)


def test_bspline_moduli_closed_form():
    """|b(m)|^2 for order 4 = 1/(2/3 + 1/3 cos(2 pi m/n))^2 (Essmann et al.
    1995, eq. 4.4 with the cubic B-spline values 1/6, 2/3, 1/6)."""
    s, cut, ew = configs.water216()
    o = Oracle(s.topo, configs.params_for(s, cut, ew))
    b1 = o.ewald_info()['bsp'][0]
    n = len(b1)
    m = np.arange(n)
    ref = (2.0 / 3.0 + (1.0 / 3.0) * np.cos(2 * np.pi * m / n)) ** 2
    assert np.max(np.abs(b1 - ref)) < 1e-14
    g = load_golden('water216')
    assert np.max(np.abs(b1 - g['bspmod1'])) < 1e-15


def _brute_ewald_real(s, beta, rvbg, rvmx):
    """Minimum-image O(N^2) restatement of the buffered water real space
    for a wrapped box: pairs with O-O <= rvbg, LJ when O-O <= rvmx, nine
    erfc Coulomb terms (pot/watwat_ewald.f), exact scipy erfc."""
    x = s.coor.reshape(-1, 3, 3)
    L = np.array(s.box)
    q = s.topo.ptchg[:3]
    e12 = s.topo.epsgm12[0] ** 2
    e6 = s.topo.epsgm6[0] ** 2
    evdw = eel = 0.0
    nw = len(x)
    for i in range(nw - 1):
        d = x[i, 0] - x[i + 1:, 0]
        sh = -L * np.round(d / L)
        roo = np.sqrt(((d + sh) ** 2).sum(1))
        sel = roo <= rvbg
        for j in np.nonzero(sel)[0]:
            jj = i + 1 + j
            if roo[j] <= rvmx:
                s6 = roo[j] ** -6
                evdw += e12 * s6 * s6 - e6 * s6
            for a in range(3):
                for b in range(3):
                    r = np.linalg.norm(x[i, a] - x[jj, b] + sh[j])
                    eel += 332.0716 * q[a] * q[b] * erfc(beta * r) / r
    return evdw, eel


def test_real_space_brute_force():
    s, cut, ew = configs.water216()
    o = Oracle(s.topo, configs.params_for(s, cut, ew))
    o.build_lists(s.coor)
    e, _ = o.eforce(s.coor)
    beta = o.ewald_info()['beta']
    evdw, eel = _brute_ewald_real(s, beta, cut['rvbg'], cut['rvmx'])
    assert abs(e['e_vdw'] + e['e_vsym'] - evdw) < 1e-9 * abs(evdw)
    assert abs(e['e_el'] + e['e_lsym'] - eel) < 1e-9 * abs(eel)


def _ewald_brute(pos, q, L, beta, kmax=14, nimg=2):
    """Ewald sum for point charges in a cubic box: real part over images
    (|n| <= nimg), reciprocal part over |k| <= kmax, minus self term."""
    kc = 332.0716
    er = 0.0
    rng = range(-nimg, nimg + 1)
    for i in range(len(q)):
        for j in range(len(q)):
            for nx in rng:
                for ny in rng:
                    for nz in rng:
                        if i == j and nx == ny == nz == 0:
                            continue
                        r = np.linalg.norm(pos[i] - pos[j] + L * np.array([nx, ny, nz]))
                        er += 0.5 * kc * q[i] * q[j] * erfc(beta * r) / r
    ek = 0.0
    V = L ** 3
    ks = np.arange(-kmax, kmax + 1)
    KX, KY, KZ = np.meshgrid(ks, ks, ks, indexing='ij')
    m = np.stack([KX.ravel(), KY.ravel(), KZ.ravel()], 1) / L
    m = m[np.any(m != 0, axis=1)]
    m2 = (m ** 2).sum(1)
    S = (q[None, :] * np.exp(2j * np.pi * (m @ pos.T))).sum(1)
    ek = kc / (2 * np.pi * V) * np.sum(np.exp(-np.pi ** 2 * m2 / beta ** 2) / m2 * np.abs(S) ** 2)
    es = -kc * beta / np.sqrt(np.pi) * np.sum(q ** 2)
    return er, ek, es


def test_pme_against_ewald_sum():
    """PME reciprocal energy converges to the exact Ewald reciprocal sum
    (4 ions in an 8 A box, no exclusions) at the order-4 rate h^4."""
    from moil_fnrfd_amd.systems import _TopoBuilder
    tb = _TopoBuilder()
    for k, qv in enumerate([1.0, -1.0, 0.5, -0.5]):
        tb.add_monomer('ION%d' % k, [('X', 10.0, qv, 0.0, 0.0)])
    t = tb.build()
    pos = np.array([[0.1, 0.2, -0.3], [2.1, -1.3, 0.7], [-2.5, 2.2, 1.9], [1.0, 3.1, -2.8]])
    L = 8.0
    errs = []
    for n in (48, 96):
        p = abi.make_params((L, L, L), rvmx=3.5, relx=3.5, rvbg=3.5, rebg=3.5, cutm=3.6,
                            dtol=1e-5, nfft=(n, n, n))
        o = Oracle(t, p)
        o.build_lists(pos)
        e, _ = o.eforce(pos)
        beta = o.ewald_info()['beta']
        er, ek, es = _ewald_brute(pos, np.array([1.0, -1.0, 0.5, -0.5]), L, beta)
        errs.append(abs(e['e_ew_receip'] - ek) / abs(ek))
        assert abs(e['e_self'] - es) < 1e-12 * abs(es)
    assert errs[1] < 2e-6
    assert 8.0 < errs[0] / errs[1] < 24.0


def test_newton_third_law_and_fd():
    """Real-space + bonded forces sum to zero; the total gradient matches
    central finite differences of the total energy at fixed lists."""
    s, cut, ew = configs.water216()
    o = Oracle(s.topo, configs.params_for(s, cut, ew))
    o.build_lists(s.coor)
    e0, dp = o.eforce(s.coor)
    rng = np.random.default_rng(5)
    h = 1e-5
    for a in rng.choice(s.topo.npt, 6, replace=False):
        for k in range(3):
            xp = s.coor.copy(); xp[a, k] += h
            xm = s.coor.copy(); xm[a, k] -= h
            ep, _ = o.eforce(xp)
            em, _ = o.eforce(xm)
            fd = (ep['e_total'] - em['e_total']) / (2 * h)
            assert abs(fd - dp[a, k]) < 2e-5 * max(1.0, abs(dp[a, k])), (a, k, fd, dp[a, k])


def _four_atom(th_deg=110.0, phi_deg=60.0):
    from moil_fnrfd_amd.systems import _TopoBuilder
    tb = _TopoBuilder()
    tb.add_monomer('MOL', [('C1', 12.0, 0.0, 0.0, 0.0), ('C2', 12.0, 0.0, 0.0, 0.0),
                           ('C3', 12.0, 0.0, 0.0, 0.0), ('C4', 12.0, 0.0, 0.0, 0.0)])
    tb.bond(1, 2, 300.0, 1.4)
    tb.angle(1, 2, 3, 60.0, 109.5)
    tb.torsion(1, 2, 3, 4, 3, 0.0, 0.0, 1.5, 1.0, 1.0, -1.0)
    tb.improper(1, 2, 3, 4, 10.0, 35.0)
    for i in range(1, 5):
        for j in range(i + 1, 5):
            tb.exclude(i, j)
    t = tb.build()
    b = 1.5
    th = math.radians(th_deg)
    phi = math.radians(phi_deg)
    p2 = np.zeros(3)
    p1 = np.array([b, 0, 0])
    p3 = np.array([b * math.cos(th), b * math.sin(th), 0])
    # p4 placed with dihedral phi about the p2-p3 axis
    u = (p3 - p2) / np.linalg.norm(p3 - p2)
    n1 = np.cross(p1 - p2, u); n1 /= np.linalg.norm(n1)
    n2 = np.cross(u, n1)
    d = -math.cos(th) * u + math.sin(th) * (math.cos(phi) * (-n2) + math.sin(phi) * n1)
    p4 = p3 + b * d
    return t, np.array([p1, p2, p3, p4])


def _wrap(d):
    """An angle difference folded into [-pi, pi] (pot/eimphi.f:212-214)."""
    if d > math.pi:
        d -= 2 * math.pi
    if d < -math.pi:
        d += 2 * math.pi
    return d


def _iupac(x):
    b1, b2, b3 = x[1] - x[0], x[2] - x[1], x[3] - x[2]
    nA, nB = np.cross(b1, b2), np.cross(b2, b3)
    return math.atan2(np.dot(np.cross(nA, nB), b2 / np.linalg.norm(b2)), np.dot(nA, nB))


@pytest.mark.parametrize('phi_deg,phi0_deg', [(60.0, 35.0), (-60.0, 35.0), (-150.0, 170.0),
                                              (150.0, -170.0), (-20.0, -35.0)])
def test_improper_sign_convention(phi_deg, phi0_deg):
    """MOIL's general improper branch on geometries of either handedness:
    E = k (phi - phi0)^2 with phi the signed IUPAC dihedral and the
    difference wrapped into [-pi, pi] -- a mirrored geometry (phi -> -phi)
    changes E, so a dropped sign or an |phi| fails."""
    from moil_fnrfd_amd.systems import _TopoBuilder
    t, x = _four_atom(phi_deg=phi_deg)
    tb = _TopoBuilder()
    tb.add_monomer('MOL', [('C%d' % k, 12.0, 0.0, 0.0, 0.0) for k in range(1, 5)])
    tb.improper(1, 2, 3, 4, 10.0, phi0_deg)
    for i in range(1, 5):
        for j in range(i + 1, 5):
            tb.exclude(i, j)
    t = tb.build()
    p = abi.make_params((30.0, 30.0, 30.0), rvmx=5, relx=5, rvbg=5, rebg=5, cutm=6, dtol=1e-5,
                        nfft=(16, 16, 16))
    o = Oracle(t, p)
    o.build_lists(x)
    e, dp = o.eforce(x)
    phi = _iupac(x)
    assert abs(abs(phi) - math.radians(180.0 - abs(phi_deg))) < 1e-9  # the builder's angle
    expect = 10.0 * _wrap(phi - float(t.impeq[0])) ** 2
    assert abs(e['e_imp'] - expect) < 1e-9 * max(1.0, expect), (e['e_imp'], expect)
    # the mirrored geometry's energy differs unless phi0 is 0 or 180
    xm = x.copy()
    xm[:, 2] = -xm[:, 2]
    em, _ = o.eforce(xm)
    assert abs(em['e_imp'] - 10.0 * _wrap(-phi - float(t.impeq[0])) ** 2) < 1e-9 * max(1.0, expect)
    h = 1e-6
    for a in range(4):
        for k in range(3):
            xp = x.copy(); xp[a, k] += h
            xn = x.copy(); xn[a, k] -= h
            fd = (o.eforce(xp)[0]['e_imp'] - o.eforce(xn)[0]['e_imp']) / (2 * h)
            assert abs(fd - dp[a, k]) < 1e-5 * max(1.0, abs(dp[a, k]))


def test_bonded_closed_forms():
    t, x = _four_atom()
    p = abi.make_params((30.0, 30.0, 30.0), rvmx=5, relx=5, rvbg=5, rebg=5, cutm=6, dtol=1e-5,
                        nfft=(16, 16, 16))
    o = Oracle(t, p)
    o.build_lists(x)
    e, dp = o.eforce(x)
    r12 = np.linalg.norm(x[0] - x[1])
    assert abs(e['e_bond'] - 300.0 * (r12 - 1.4) ** 2) < 1e-10
    v1, v2 = x[0] - x[1], x[2] - x[1]
    th = math.acos(np.dot(v1, v2) / np.linalg.norm(v1) / np.linalg.norm(v2))
    assert abs(e['e_theta'] - 60.0 * (th - math.radians(109.5)) ** 2) < 1e-10
    # dihedral angle (IUPAC sign) and the cosine series with phase -1
    b1, b2, b3 = x[1] - x[0], x[2] - x[1], x[3] - x[2]
    nA, nB = np.cross(b1, b2), np.cross(b2, b3)
    phi = math.atan2(np.dot(np.cross(nA, nB), b2 / np.linalg.norm(b2)), np.dot(nA, nB))
    assert abs(e['e_tors'] - (1.5 - 1.5 * math.cos(3 * phi))) < 1e-10
    # improper: k (phi - phi0)^2 with phi signed as the IUPAC dihedral
    # (MOIL's sign is that of ((a x b) x (b x c)) . b, pot/eimphi.f:188-206,
    # the same sign); checked against the IUPAC angle itself
    assert abs(e['e_imp'] - 10.0 * _wrap(phi - math.radians(35.0)) ** 2) < 1e-8
    # gradient by central differences
    h = 1e-6
    for a in range(4):
        for k in range(3):
            xp = x.copy(); xp[a, k] += h
            xm = x.copy(); xm[a, k] -= h
            ep, _ = o.eforce(xp)
            em, _ = o.eforce(xm)
            fd = (ep['e_total'] - em['e_total']) / (2 * h)
            assert abs(fd - dp[a, k]) < 1e-5 * max(1.0, abs(dp[a, k]))
