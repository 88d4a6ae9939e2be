"""Large heteropolymer-in-water systems on the GPU (the ~1.0M-atom
configuration and a 30k-atom box of the same recipe).

* 30k atoms: the GPU FP64 path against the CPU oracle on the same input --
  water and atom pair sets bit-exact, energies <= 1e-12 relative, forces
  <= 1e-10 RMS-relative -- and mixed precision within 1e-6 (e_total) /
  1e-5 (forces) of FP64.
* ~1.0M atoms (20,000 chains, 215 A box, 216^3 grid): the FP64 path
  against the oracle's energies, list sizes and the gradient of 4,096
  seeded atoms (fixture tests/golden/hetero1m.npz, written by
  tests/golden/gen_large.py), and against the exactly rounded exclusion
  correction; mixed precision within 1e-6 / 1e-5 of FP64 and
  bit-reproducible run to run; list sizes equal in both precisions.
"""
import os
import sys
import time

import numpy as np
import pytest

from moil_fnrfd_amd import abi, configs
from tests._util import GOLDEN, energy_errors, rms_rel, sort_rows

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mixed_vs_fp64(e32, f32, e64, f64):
    rel_e = abs(e32['e_total'] - e64['e_total']) / abs(e64['e_total'])
    rel_f = rms_rel(f32, f64)
    _say('  mixed vs FP64: e_total rel %.3e  force rms-rel %.3e' % (rel_e, rel_f))
    assert rel_e < 1e-6
    assert rel_f < 1e-5


def test_hetero_30k_matches_oracle():
    from moil_fnrfd_amd.eforce import MoilGPU
    from oracle.oracle import Oracle
    _say('case hetero30k generate')
    s, cut, ew = configs.hetero_box(600, 66.9, seed=4100)
    p64 = configs.params_for(s, cut, ew, precision=abi.PREC_FP64)
    o = Oracle(s.topo, p64)
    x = o.squeeze(s.coor)
    o.build_lists(x)
    e_ref, f_ref = o.eforce(x)
    w_ref, a_ref = sort_rows(o.water_pairs()), sort_rows(o.atom_pairs())
    o.close()
    _say('case hetero30k FP64 npt=%d' % s.topo.npt)
    m = MoilGPU(s.topo, p64)
    xg = m.squeeze(s.coor)
    assert np.array_equal(xg, x)
    m.nbondm(xg)
    assert np.array_equal(sort_rows(m.water_pairs()), w_ref)
    assert np.array_equal(sort_rows(m.atom_pairs()), a_ref)
    e64, f64 = m.eforce(xg)
    err = energy_errors(e64, e_ref)
    _say('  FP64 vs oracle: max energy rel %.3e  force rms-rel %.3e'
         % (max(err.values()), rms_rel(f64, f_ref)))
    assert max(err.values()) < 1e-12, err
    assert rms_rel(f64, f_ref) < 1e-10
    m.close()
    _say('case hetero30k MIXED')
    m = MoilGPU(s.topo, configs.params_for(s, cut, ew, precision=abi.PREC_MIXED))
    m.nbondm(xg)
    e32, f32 = m.eforce(xg)
    m.close()
    _mixed_vs_fp64(e32, f32, e64, f64)


@pytest.fixture(scope='module')
def big():
    path = os.path.join(GOLDEN, 'hetero1m.npz')
    g = np.load(path)
    t0 = time.time()
    _say('case hetero1m generate')
    s, cut, ew = configs.hetero_1m()
    _say('  generated %d atoms in %.1f s' % (s.topo.npt, time.time() - t0))
    assert s.topo.npt == int(g['npt'][0])
    return s, cut, ew, g


def test_hetero_1m(big):
    import torch
    from moil_fnrfd_amd.eforce import MoilGPU
    s, cut, ew, g = big
    torch.cuda.set_device(0)
    out = {}
    for prec, name in ((abi.PREC_FP64, 'FP64'), (abi.PREC_MIXED, 'MIXED')):
        _say('case hetero1m %s' % name)
        m = MoilGPU(s.topo, configs.params_for(s, cut, ew, precision=prec))
        t0 = time.time()
        x = m.squeeze(s.coor)
        m.nbondm(x)
        t1 = time.time()
        e, f = m.eforce(x)
        t2 = time.time()
        cnt = m.list_counts()
        _say('  lists %.2f s (host-driven, incl. transfers), first eforce %.2f s, counts %s'
             % (t1 - t0, t2 - t1, list(cnt[:8])))
        if prec == abi.PREC_MIXED:
            e2, f2 = m.eforce(x)
            assert np.array_equal(f, f2) and e == e2
        out[name] = (x, e, f, cnt)
        m.close()
    x64, e64, f64, c64 = out['FP64']
    x32, e32, f32, c32 = out['MIXED']
    assert np.array_equal(x64, x32)
    assert np.array_equal(c64, c32)
    # against the oracle
    assert int(c64[0] + c64[1] + c64[2] + c64[3]) == int(g['n_water_rows'][0])
    assert int(c64[4] + c64[5] + c64[6]) == int(g['n_atom_rows'][0])
    ref = dict(zip(abi.ENERGY_NAMES, g['energies']))
    err = energy_errors(e64, ref)
    idx = g['dpot_idx']
    _say('  FP64 vs oracle: energy rel %s  force rms-rel (subset) %.3e'
         % ({k: '%.1e' % v for k, v in err.items()}, rms_rel(f64[idx], g['dpot'])))
    # the oracle sums sequentially (as MOIL); over ~1.5M exclusion terms and
    # 10M grid points that carries ~1e-11 of rounding, so those sums (and
    # the totals that contain them) are held to 5e-11, the exclusion
    # correction itself to 1e-13 of its exactly rounded sum
    seq = ('e_corr', 'e_ew_receip', 'e_total', 'e_dir')
    assert max(v for k, v in err.items() if k not in seq) < 1e-12, err
    assert max(err[k] for k in seq) < 5e-11, err
    exact = float(g['e_corr_exact'][0])
    _say('  e_corr vs exactly rounded sum: %.2e' % (abs(e64['e_corr'] - exact) / abs(exact)))
    assert abs(e64['e_corr'] - exact) <= 1e-13 * abs(exact)
    assert rms_rel(f64[idx], g['dpot']) < 1e-10
    _mixed_vs_fp64(e32, f32, e64, f64)
