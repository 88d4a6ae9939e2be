"""Workload definitions: the synthetic systems and MOIL run parameters of
the energy-path benchmark, plus small parity cases.

Each entry returns (System, cut, ewald) where ``cut`` holds MOIL's
cutoff keywords in Angstrom (rvmx relx rvbg rebg cutm; squared later as
pot/energy.f:392-400 does) and ``ewald`` the PME keywords (dtol, grid,
order).  dtol values reproduce the target beta through MOIL's own
bisection (symm/ewald_init.f:298-335).
"""
from __future__ import annotations

from . import abi
from .systems import hetero_packed, hetero_system, to_arith, water_box

# dtol giving beta = 0.34 at rc = 9 A and beta = 0.31 at rc = 10 A
DTOL_B034_RC9 = 1.67573104443267911e-06
DTOL_B031_RC10 = 1.16486573671995838e-06


def water216():
    """216 rigid waters (648 atoms), 18.6 A box, 9 A cutoffs, 20^3 grid."""
    s = water_box(6, 18.6, seed=1000)
    cut = dict(rvmx=9.0, relx=9.0, rvbg=9.0, rebg=9.0, cutm=10.0)
    ew = dict(dtol=DTOL_B034_RC9, nfft=(20, 20, 20), order=4)
    return s, cut, ew


def water32k():
    """32,768 waters (98,304 atoms) in a 68.9 A box: 10 A cutoffs with a
    1.5 A list skin, beta 0.31, 72^3 grid."""
    s = water_box(32, 68.9, seed=1001)
    cut = dict(rvmx=10.0, relx=10.0, rvbg=11.5, rebg=11.5, cutm=12.5)
    ew = dict(dtol=DTOL_B031_RC10, nfft=(72, 72, 72), order=4)
    return s, cut, ew


def water_scaled(n_side, seed, L=None):
    """Water boxes at the 32k-water density (0.100 waters/A^3) with the
    32k-water cutoffs; grid spacing ~0.96 A."""
    if L is None:
        L = 68.9 * n_side / 32.0
    s = water_box(n_side, L, seed=seed)
    n = max(8, int(round(L / 0.957)))
    n += n % 2
    cut = dict(rvmx=10.0, relx=10.0, rvbg=11.5, rebg=11.5, cutm=12.5)
    ew = dict(dtol=DTOL_B031_RC10, nfft=(n, n, n), order=4)
    return s, cut, ew


def water_buffered(n_side=10, seed=1010):
    """Liquid-density waters with rvmx < relx < rvbg < rebg so that both
    water lists and the buffered evaluation checks are exercised."""
    L = 3.1 * n_side
    s = water_box(n_side, L, seed=seed)
    cut = dict(rvmx=7.0, relx=8.0, rvbg=8.5, rebg=9.5, cutm=10.5)
    n = int(round(L / 1.0))
    n += n % 2
    ew = dict(dtol=1e-5, nfft=(n, n, n), order=4)
    return s, cut, ew


def hetero_small(seed=2000, n_chains=12, L=24.0):
    """~1.4k-atom heteropolymer-in-water box for the atom-list parity
    cases: 15 % uncharged and 15 % zero-LJ chain atoms so that every MOIL
    atom list (list1/list2/list3) and the hydrogen rule are exercised."""
    s = hetero_system(n_chains, L, seed=seed, frac_neutral=0.15, frac_hlike=0.15)
    cut = dict(rvmx=9.0, relx=9.0, rvbg=10.0, rebg=10.0, cutm=15.0)
    n = int(round(L)) + int(round(L)) % 2
    ew = dict(dtol=1e-5, nfft=(n, n, n), order=4)
    return s, cut, ew


def hetero_small_arith(seed=2500, n_chains=10, L=22.0):
    """hetero_small's recipe converted to MOIL's arith rule (systems.to_arith):
    Lorentz-Berthelot LJ in every pair kernel including the water O-O
    constants, one-cosine torsions of periods 1..4 (some dihedrals with
    two terms), k - k cos 2 phi impropers, no hydrogen rule."""
    s = hetero_system(n_chains, L, seed=seed, frac_neutral=0.15, frac_hlike=0.15)
    s.topo = to_arith(s.topo, seed + 1)
    cut = dict(rvmx=9.0, relx=9.0, rvbg=10.0, rebg=10.0, cutm=15.0)
    n = int(round(L)) + int(round(L)) % 2
    ew = dict(dtol=1e-5, nfft=(n, n, n), order=4)
    return s, cut, ew


def hetero_small_phi0(seed=2400, n_chains=10, L=22.0):
    """hetero_small's recipe with improper phases away from 0 and 180 deg
    (+-U(10, 170)): MOIL's general k (phi - phi0)^2 improper branch."""
    s = hetero_system(n_chains, L, seed=seed, frac_neutral=0.15, frac_hlike=0.15,
                      imp_general=True)
    cut = dict(rvmx=9.0, relx=9.0, rvbg=10.0, rebg=10.0, cutm=15.0)
    n = int(round(L)) + int(round(L)) % 2
    ew = dict(dtol=1e-5, nfft=(n, n, n), order=4)
    return s, cut, ew


def chains_only(seed=2200, n_chains=8, L=22.0, n_water=0):
    """Heteropolymer chains with no water at all (or n_water waters): the
    water lists are empty and every pair goes through MOIL's atom lists;
    15 % uncharged and 15 % zero-LJ chain atoms as in hetero_small."""
    s = hetero_system(n_chains, L, seed=seed, n_water=n_water, frac_neutral=0.15,
                      frac_hlike=0.15)
    cut = dict(rvmx=9.0, relx=9.0, rvbg=10.0, rebg=10.0, cutm=15.0)
    n = int(round(L)) + int(round(L)) % 2
    ew = dict(dtol=1e-5, nfft=(n, n, n), order=4)
    return s, cut, ew


def hetero_box(n_chains, L, seed):
    """Heteropolymer in water by the large-system recipe (hetero_packed):
    n_chains chains of 10-20 atoms in an L-A box at 0.1006 atoms/A^3, 10 A
    cutoffs with a 1.5 A list skin, explicit monomer cutoff 18 A (>= rebg
    + twice the largest monomer radius + 1 A), beta 0.31, PME grid at ~1 A
    spacing."""
    s = hetero_packed(n_chains, L, seed=seed)
    cut = dict(rvmx=10.0, relx=10.0, rvbg=11.5, rebg=11.5, cutm=18.0)
    n = fft_size(L)
    ew = dict(dtol=DTOL_B031_RC10, nfft=(n, n, n), order=4)
    return s, cut, ew


def hetero_1m(seed=2100):
    """The ~1.0M-atom solvated heteropolymer: 20,000 chains of 10-20 atoms
    (~300k chain atoms, 24 LJ types, 1-4 pairs scaled 0.5) and ~233k
    3-site waters in a 215 A box, 10 A cutoffs + 1.5 A skin, 216^3 grid."""
    s, cut, ew = hetero_box(20000, 215.0, seed)
    ew['nfft'] = (216, 216, 216)
    return s, cut, ew


def hetero_10k(seed=4000):
    """The ~10,000-atom system of the same recipe (200 chains, 46.4 A box,
    48^3 grid)."""
    return hetero_box(200, 46.4, seed)


def params_for(sysm, cut, ew, precision=abi.PREC_FP64, terms=abi.TERM_DEFAULT):
    """Library parameters for a workload; cut['hvdw0'] = False switches
    MOIL's hydrogen rule off (its 'hvdw' keyword)."""
    if not cut.get('hvdw0', True):
        terms &= ~abi.TERM_HVDW0
    if getattr(sysm.topo, 'arith', False):
        terms |= abi.TERM_ARITH
    return abi.make_params(sysm.box, rvmx=cut['rvmx'], relx=cut['relx'], rvbg=cut['rvbg'],
                           rebg=cut['rebg'], cutm=cut['cutm'], dtol=ew['dtol'], nfft=ew['nfft'],
                           order=ew.get('order', 4), eps=cut.get('eps', 1.0),
                           precision=precision, terms=terms)


def fft_size(n_min):
    """Smallest 2^a 3^b 5^c >= n_min (and >= 8)."""
    n = max(int(n_min), 8)
    while True:
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            return n
        n += 1


def batch_system(n_atoms, seed, n_chains):
    """One system of the batched workload: a cubic box at liquid water
    density (0.1002 atoms/A^3) holding n_chains heteropolymer chains and
    3-site water, 9 A cutoffs with a 1.5 A list skin, beta 0.34, PME grid
    at ~1 A spacing."""
    L = (n_atoms / (3 * 0.0334)) ** (1.0 / 3.0)
    s = hetero_system(n_chains, L, seed=seed)
    n = fft_size(L)
    cut = dict(rvmx=9.0, relx=9.0, rvbg=10.5, rebg=10.5, cutm=18.0)
    ew = dict(dtol=DTOL_B034_RC9, nfft=(n, n, n), order=4)
    return s, cut, ew


def batch_workload(seed=3000, n_systems=64, n_slices=32, slice_atoms=12000):
    """The batched workload: n_systems independent periodic systems with
    atom counts log-uniform in [3,000, 50,000], each with 1-8 chains, plus
    n_slices slices of a slice_atoms-atom system linearly interpolated
    between two conformations (the second = the first + N(0, 0.5 A)).
    Returns a list of (System, cut, ewald)."""
    import copy

    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_systems):
        n = int(round(np.exp(rng.uniform(np.log(3000.0), np.log(50000.0)))))
        out.append(batch_system(n, seed + 1 + k, int(rng.integers(1, 9))))
    base, cut, ew = batch_system(slice_atoms, seed + 1000, 8)
    far = base.coor + rng.normal(0.0, 0.5, base.coor.shape)
    for k in range(n_slices):
        s = copy.copy(base)
        lam = k / max(n_slices - 1, 1)
        s.coor = (1.0 - lam) * base.coor + lam * far
        out.append((s, cut, ew))
    return out


# MOIL's chain minimiser defaults (chain/chmin.f:134-145)
CHAIN_GAMMA = 100.0
CHAIN_RHO = 100.0
CHAIN_LAMBDA = 2.0


def chain_replicas(base_coor, igrid, seed, end_sigma=0.5, noise=0.05):
    """igrid replicas of one system for the chain functional: replica j
    lies at fraction j/(igrid-1) of the straight line from base_coor to
    base_coor + N(0, end_sigma) per coordinate, interior and last replicas
    displaced by a further N(0, noise).  Returns (igrid, npt, 3) float64."""
    import numpy as np
    rng = np.random.default_rng(seed)
    base = np.asarray(base_coor, np.float64)
    far = base + rng.normal(0.0, end_sigma, base.shape)
    r = np.empty((igrid,) + base.shape)
    for j in range(igrid):
        lam = j / max(igrid - 1, 1)
        r[j] = (1.0 - lam) * base + lam * far
        if j > 0:
            r[j] += rng.normal(0.0, noise, base.shape)
    return r


def chain_pick(npt, every=5, skip=3):
    """A partial selection (MOIL's ipick): atom i (0-based) is left out
    when i % every == skip."""
    import numpy as np
    p = np.ones(npt, np.int32)
    p[skip::every] = 0
    return p


def maxwell_velocities(masses, temperature, seed):
    """Velocities (npt, 3) in A per MOIL time unit drawn from the
    Maxwell-Boltzmann distribution at `temperature` K: sqrt(kB T / m) N(0, 1)
    per component (MOIL's velinit draws the same distribution,
    dynamics/velinit.f:37)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    m = np.asarray(masses, np.float64)
    return rng.normal(size=(len(m), 3)) * np.sqrt(abi.KBOLTZMANN * temperature / m)[:, None]
