"""The chain functional's oracle (oracle/chain.py) pinned three ways:
against MOIL's own chain/sds.f run (fixture chain_water216.npz, written by
tests/golden/gen_golden.py through oracle/_ref/moil_sds), bit for bit;
against central finite differences of the chain energy, which the
restatement's gradient must match whatever MOIL's code says; and by
translation invariance (every chain term depends on replica differences
only, so its gradient sums to zero over the replicas, atom by atom)."""
import hashlib

import numpy as np
import pytest

from oracle.chain import chain_energy_only, chain_terms
from tests.golden.gen_golden import chain_case, topo_checksum
from tests._util import load_golden


def test_chain_oracle_matches_moil_sds_bitwise():
    g = load_golden('chain_water216')
    s, cut, ew, r, pick = chain_case()
    assert hashlib.sha256(np.ascontiguousarray(r).tobytes()).hexdigest().encode() == \
        bytes(g['r_sha'])
    assert topo_checksum(s.topo).encode() == bytes(g['topo_sha'])
    assert 0 < pick.sum() < len(pick)  # a partial selection is exercised
    gi = np.zeros_like(r)
    gi[1:-1] = g['dpots']
    o = chain_terms(r, pick, *g['params'], g['etots'], gi)
    assert o['sener'] == g['sener'][0]
    assert np.array_equal(o['d0'], g['d0'])
    assert np.array_equal(o['e1'], g['e1'])
    assert np.array_equal(o['e0'][1:-1], g['e0'][1:-1])
    assert np.array_equal(o['dv'], g['dv'])
    # unselected atoms carry their replica's own gradient only
    off = pick == 0
    assert np.array_equal(g['dv'][1:-1][:, off], g['dpots'][:, off])
    assert not g['dv'][0][off].any() and not g['dv'][-1][off].any()


def _small_chain(seed, igrid=7, npt=6):
    rng = np.random.default_rng(seed)
    base = rng.normal(0.0, 2.0, (npt, 3))
    step = rng.normal(0.0, 0.6, (npt, 3))
    r = np.array([base + j * step + rng.normal(0.0, 0.15, (npt, 3)) for j in range(igrid)])
    pick = np.ones(npt, np.int32)
    pick[1] = 0
    return r, pick


@pytest.mark.parametrize('params', [(100.0, 100.0, 2.0), (3.0, 40.0, 0.7)])
def test_chain_gradient_matches_finite_differences(params):
    gamma, rho, lam = params
    r, pick = _small_chain(7)
    igrid, npt = r.shape[:2]
    zero_e = np.zeros(igrid)
    o = chain_terms(r, pick, gamma, rho, lam, zero_e, np.zeros_like(r))
    e_ref = chain_energy_only(r, pick, gamma, rho, lam)
    assert abs(o['sener'] - e_ref) <= 1e-12 * abs(e_ref)
    h = 1e-5
    fd = np.zeros_like(r)
    for idx in np.ndindex(r.shape):
        rp = r.copy()
        rp[idx] += h
        rm = r.copy()
        rm[idx] -= h
        fd[idx] = (chain_energy_only(rp, pick, gamma, rho, lam) -
                   chain_energy_only(rm, pick, gamma, rho, lam)) / (2 * h)
    scale = np.abs(fd).max()
    assert np.abs(o['dv'] - fd).max() <= 1e-6 * scale
    # the unselected atom feels nothing
    assert not o['dv'][:, pick == 0].any()


def test_chain_gradient_is_translation_invariant():
    r, pick = _small_chain(11, igrid=9, npt=10)
    o = chain_terms(r, pick, 100.0, 100.0, 2.0, np.zeros(9), np.zeros_like(r))
    tot = o['dv'].sum(axis=0)
    assert np.abs(tot).max() <= 1e-12 * np.abs(o['dv']).max()
    # shifting every replica by one vector changes nothing
    o2 = chain_terms(r + np.array([0.3, -1.7, 2.25]), pick, 100.0, 100.0, 2.0, np.zeros(9),
                     np.zeros_like(r))
    assert np.allclose(o2['dv'], o['dv'], rtol=0, atol=1e-11 * np.abs(o['dv']).max())
    assert abs(o2['sener'] - o['sener']) <= 1e-12 * abs(o['sener'])
