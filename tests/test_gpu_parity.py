"""GPU parity: the HIP library against MOIL's own outputs (golden
fixtures written by oracle/_ref) and against the CPU oracle.

Bars (per the energy-path contract):
  * pair sets: bit-exact (canonical (i, j, symop, list) rows);
  * FP64 mode: energies <= 1e-12 relative (e_total relative to the sum
    of the magnitudes of its terms), forces <= 1e-10 RMS-relative;
  * mixed mode (FP32 pairs, fixed-point forces): e_total <= 1e-6
    relative, forces <= 1e-5 RMS-relative against the FP64 run.
"""
import hashlib
import sys

import numpy as np
import pytest

from moil_fnrfd_amd import abi, configs
from moil_fnrfd_amd.eforce import MoilGPUError
from tests._util import (energy_errors, golden_energies, load_golden, rms_rel, rows_sha,
                         sort_rows)
from tests._cases import CASES, make_case

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope='module')
def gpu():
    import torch
    assert torch.cuda.is_available()
    from moil_fnrfd_amd.eforce import MoilGPU, load_library
    load_library()
    return MoilGPU


SMALL = CASES


@pytest.mark.parametrize('name', sorted(SMALL))
def test_fp64_matches_moil(gpu, name):
    _say('case %s FP64' % name)
    s, cut, ew, sqz = make_case(name)
    g = load_golden(name)
    m = gpu(s.topo, configs.params_for(s, cut, ew, precision=abi.PREC_FP64))
    x = m.squeeze(s.coor) if sqz else s.coor
    if 'coor_after' in g:
        assert np.array_equal(x, g['coor_after'])
    m.nbondm(x)
    rows = sort_rows(m.water_pairs())
    if 'pairs' in g:
        assert np.array_equal(rows, g['pairs'])
    assert rows_sha(rows) == bytes(g['pairs_sha']).decode()
    arows = sort_rows(m.atom_pairs())
    if 'apairs' in g:
        assert np.array_equal(arows, g['apairs'])
    assert rows_sha(arows) == bytes(g['apairs_sha']).decode()
    cnt = m.list_counts()
    assert int(cnt[4] + cnt[5] + cnt[6]) == int(g['n_apairs'][0])
    assert int(cnt[7]) == int(g['n_apairs'][4])
    e, dp = m.eforce(x)
    ref = golden_energies(g)
    info = m.ewald_info()
    assert info['beta'] == ref['ewaldcof']
    # the device's erfc table is MOIL's, byte for byte
    assert len(info['erf_arr']) == int(g['erftab_rows'][0])
    assert hashlib.sha256(np.ascontiguousarray(info['erf_arr']).tobytes()).hexdigest() == \
        bytes(g['erftab_sha']).decode()
    err = energy_errors(e, ref)
    own = abs(e['e_total'] - ref['e_total']) / abs(ref['e_total'])
    _say('  %s e_total %.12f (MOIL %.12f): %.2e of sum|terms|, %.2e of e_total; max term %.2e; '
         'force rms-rel %.2e' % (name, e['e_total'], ref['e_total'], err['e_total'], own,
                                 max(err[k] for k in err if k not in ('e_total', 'e_dir')),
                                 rms_rel(dp, g['dpot'])))
    assert max(err.values()) < 1e-12, err
    assert rms_rel(dp, g['dpot']) < 1e-10
    m.close()


def test_fp64_matches_moil_32k(gpu):
    _say('case water32k FP64')
    s, cut, ew = configs.water32k()
    g = load_golden('water32k')
    m = gpu(s.topo, configs.params_for(s, cut, ew, precision=abi.PREC_FP64))
    m.nbondm(s.coor)
    rows = m.water_pairs()
    assert len(rows) == int(g['n_pairs'][0])
    assert rows_sha(rows) == bytes(g['pairs_sha']).decode()
    e, dp = m.eforce(s.coor)
    err = energy_errors(e, golden_energies(g))
    assert max(err.values()) < 1e-12, err
    assert rms_rel(dp[g['dpot_idx']], g['dpot']) < 1e-10
    m.close()


def test_mixed_precision_32k(gpu):
    """FP32 pair math with fixed-point forces vs the FP64 path, and
    bit-reproducibility run to run."""
    _say('case water32k MIXED vs FP64')
    s, cut, ew = configs.water32k()
    m64 = gpu(s.topo, configs.params_for(s, cut, ew, precision=abi.PREC_FP64))
    m64.nbondm(s.coor)
    e64, f64 = m64.eforce(s.coor)
    m64.close()
    m32 = gpu(s.topo, configs.params_for(s, cut, ew, precision=abi.PREC_MIXED))
    m32.nbondm(s.coor)
    e32, f32 = m32.eforce(s.coor)
    e32b, f32b = m32.eforce(s.coor)
    m32.close()
    rel_e = abs(e32['e_total'] - e64['e_total']) / abs(e64['e_total'])
    rel_f = rms_rel(f32, f64)
    _say('  e_total rel %.3e  force rms-rel %.3e' % (rel_e, rel_f))
    assert rel_e < 1e-6
    assert rel_f < 1e-5
    assert np.array_equal(f32, f32b) and e32 == e32b


def test_short_grid_water_kernel_matches_plain(gpu, monkeypatch):
    """The mixed-precision water kernel on a short grid (a fixed number of
    workgroups per CU walking the chunks in alternating rounds) against the
    one-workgroup-per-chunk launch: every chunk writes the same partials
    either way, so forces and energies are identical bit for bit."""
    s, cut, ew = configs.water32k()
    p = configs.params_for(s, cut, ew, precision=abi.PREC_MIXED)
    out = {}
    for wg in ('0', '1', '5'):
        _say('case water32k MIXED water workgroups per CU %s' % wg)
        monkeypatch.setenv('MOIL_GPU_WATER_WG_PER_CU', wg)
        m = gpu(s.topo, p)
        m.nbondm(s.coor)
        out[wg] = m.eforce(s.coor)
        m.close()
    e0, f0 = out['0']
    for wg in ('1', '5'):
        e, f = out[wg]
        assert np.array_equal(f, f0), wg
        assert e == e0, wg


def test_lists_follow_perturbed_coordinates(gpu):
    """List rebuilds on perturbed / drifting coordinates (incl. squeeze)
    reproduce the oracle's pair sets exactly."""
    from oracle.oracle import Oracle
    s, cut, ew = configs.water_scaled(8, seed=1200)
    p = configs.params_for(s, cut, ew, precision=abi.PREC_MIXED)
    m = gpu(s.topo, p)
    o = Oracle(s.topo, p)
    rng = np.random.default_rng(9)
    x = s.coor.copy()
    for step in range(4):
        _say('case water_scaled8 perturbation %d' % step)
        x = x + rng.normal(0.0, 0.3, x.shape)
        xg = m.squeeze(x)
        xo = o.squeeze(x)
        assert np.array_equal(xg, xo)
        m.nbondm(xg)
        o.build_lists(xo)
        assert np.array_equal(sort_rows(m.water_pairs()), sort_rows(o.water_pairs()))
        x = xg
    m.close()


def test_energy_conservation_properties_32k(gpu):
    """Size-independent properties at full size: the real-space plus
    bonded force (everything except PME) sums to zero; total force sum
    is ~0 (PME conserves momentum only approximately)."""
    _say('case water32k momentum')
    s, cut, ew = configs.water32k()
    p = configs.params_for(s, cut, ew, precision=abi.PREC_FP64,
                           terms=abi.TERM_DEFAULT)
    m = gpu(s.topo, p)
    m.nbondm(s.coor)
    e, dp = m.eforce(s.coor)
    net = dp.sum(0)
    assert np.all(np.abs(net) < 1e-3 * np.sqrt((dp ** 2).sum(1)).max())
    m.close()


@pytest.mark.parametrize('name', ['hetero_small', 'hetero_small_out', 'hetero_small_arith'])
def test_mixed_precision_hetero(gpu, name):
    """Mixed mode on the atom lists (list1/list2/list3, images both ways)
    against the FP64 path, and run-to-run bit reproducibility."""
    _say('case %s MIXED vs FP64' % name)
    s, cut, ew, sqz = make_case(name)
    m64 = gpu(s.topo, configs.params_for(s, cut, ew, precision=abi.PREC_FP64))
    x = m64.squeeze(s.coor) if sqz else s.coor
    m64.nbondm(x)
    e64, f64 = m64.eforce(x)
    m64.close()
    m32 = gpu(s.topo, configs.params_for(s, cut, ew, precision=abi.PREC_MIXED))
    m32.nbondm(x)
    e32, f32 = m32.eforce(x)
    e32b, f32b = m32.eforce(x)
    m32.close()
    scale = sum(abs(e64[k]) for k in ('e_vdw', 'e_el', 'e_vsym', 'e_lsym', 'e_corr', 'e_self',
                                      'e_ew_receip', 'e_bond', 'e_theta', 'e_tors', 'e_imp',
                                      'e_el14', 'e_vdw14'))
    rel_e = abs(e32['e_total'] - e64['e_total']) / scale
    rel_f = rms_rel(f32, f64)
    _say('  e_total rel %.3e  force rms-rel %.3e' % (rel_e, rel_f))
    assert rel_e < 1e-6
    assert rel_f < 1e-5
    assert np.array_equal(f32, f32b) and e32 == e32b


def test_hetero_lists_follow_perturbed_coordinates(gpu):
    """Atom and water lists rebuilt on drifting heteropolymer coordinates
    match the oracle's exactly."""
    from oracle.oracle import Oracle
    s, cut, ew, _ = make_case('hetero_small')
    p = configs.params_for(s, cut, ew, precision=abi.PREC_MIXED)
    m = gpu(s.topo, p)
    o = Oracle(s.topo, p)
    rng = np.random.default_rng(19)
    x = s.coor.copy()
    for step in range(3):
        _say('case hetero_small perturbation %d' % step)
        x = x + rng.normal(0.0, 0.2, x.shape)
        xg = m.squeeze(x)
        xo = o.squeeze(x)
        assert np.array_equal(xg, xo)
        m.nbondm(xg)
        o.build_lists(xo)
        assert np.array_equal(sort_rows(m.water_pairs()), sort_rows(o.water_pairs()))
        assert np.array_equal(sort_rows(m.atom_pairs()), sort_rows(o.atom_pairs()))
        x = xg
    m.close()


@pytest.mark.parametrize('name', ['water216', 'hetero_small'])
def test_atoms_moved_beyond_spreading_margin(gpu, name):
    """The reciprocal sum keeps the spreading tiles of the last list build;
    atoms that moved further than the tiles' margin since then take the
    exact fallback (fixed-point overflow grid for the charges, direct grid
    reads in the gather).  With lists built at x and the energy evaluated
    at x' (40 atoms displaced by 1.5-3 A, five waters by a box length), the
    GPU's FP64 result equals the oracle's on the same stale lists (its
    reciprocal sum has no tiles)."""
    from oracle.oracle import Oracle
    _say('case %s stale spreading tiles' % name)
    s, cut, ew, sqz = make_case(name)
    p = configs.params_for(s, cut, ew, precision=abi.PREC_FP64)
    m = gpu(s.topo, p)
    x = m.squeeze(s.coor) if sqz else s.coor.copy()
    m.nbondm(x)
    rng = np.random.default_rng(77)
    xp = x.copy()
    idx = rng.choice(len(x), 40, replace=False)
    L = np.asarray(s.box, np.float64)
    for i in idx:
        # a displacement that keeps the atom 1.5 A from every other atom
        # (closer contacts give force terms beyond the FP64 fixed-point range)
        while True:
            d = rng.normal(size=3)
            d *= rng.uniform(1.5, 3.0) / np.linalg.norm(d)
            r = xp - (xp[i] + d)
            r -= L * np.round(r / L)
            r2 = (r * r).sum(1)
            r2[i] = np.inf
            if r2.min() >= 1.5 ** 2:
                break
        xp[i] += d
    # five whole waters one box length away (molecules stay whole, as
    # MOIL's exclusion correction assumes)
    for w in rng.choice(len(s.topo.idxtip3), 5, replace=False):
        a0 = int(s.topo.poipt[s.topo.idxtip3[w] - 1])
        xp[a0:a0 + 3, 0] += s.box[0]
    e, dp = m.eforce(xp)
    o = Oracle(s.topo, p)
    o.build_lists(x)
    e_ref, dp_ref = o.eforce(xp)
    err = energy_errors(e, e_ref)
    _say('  max term %.2e, e_ew_receip %.2e, force rms-rel %.2e'
         % (max(err.values()), err['e_ew_receip'], rms_rel(dp, dp_ref)))
    assert max(err.values()) < 1e-12, err
    assert rms_rel(dp, dp_ref) < 1e-10
    # mixed precision on the same stale tiles against the FP64 result
    mm = gpu(s.topo, configs.params_for(s, cut, ew, precision=abi.PREC_MIXED))
    mm.nbondm(x)
    em, dpm = mm.eforce(xp)
    assert rms_rel(dpm, dp_ref) < 1e-5
    m.close()
    mm.close()


def test_split_water_beyond_erfc_table(gpu):
    """A water split across the box (one H one box length from its O, as
    per-atom wrapping leaves it): its O-H exclusion correction and its site
    pairs lie beyond MOIL's erfc table (1.5 relx), where MOIL reads past its
    array.  The library reads the table's last row at its right end there,
    as the oracle does: FP64 energies and gradient equal the oracle's, the
    mixed gradient is within its bar of them, and every value is finite."""
    from oracle.oracle import Oracle
    _say('case water216 split water beyond the erfc table')
    s, cut, ew, _ = make_case('water216')
    p = configs.params_for(s, cut, ew, precision=abi.PREC_FP64)
    x = s.coor.copy()
    m = gpu(s.topo, p)
    m.nbondm(x)
    a0 = int(s.topo.poipt[s.topo.idxtip3[7] - 1])
    xp = x.copy()
    xp[a0 + 1, 1] += s.box[1]
    e, dp = m.eforce(xp)
    o = Oracle(s.topo, p)
    o.build_lists(x)
    e_ref, dp_ref = o.eforce(xp)
    err = energy_errors(e, e_ref)
    _say('  max term %.2e, e_corr %.6f, force rms-rel %.2e'
         % (max(err.values()), e['e_corr'], rms_rel(dp, dp_ref)))
    assert np.isfinite(dp).all() and all(np.isfinite(v) for v in e.values())
    assert max(err.values()) < 1e-12, err
    assert rms_rel(dp, dp_ref) < 1e-10
    mm = gpu(s.topo, configs.params_for(s, cut, ew, precision=abi.PREC_MIXED))
    mm.nbondm(x)
    em, dpm = mm.eforce(xp)
    assert rms_rel(dpm, dp_ref) < 1e-5
    m.close()
    mm.close()


def test_arith_context_refuses_geometric_parameter_paths(gpu):
    """set_params derives 1-4 coefficients and param_grad differentiates
    per-atom factors by the geometric rule; an arith context refuses both
    and keeps evaluating (its eforce still matches MOIL's arith run)."""
    _say('case hetero_small_arith refusals')
    s, cut, ew, sqz = make_case('hetero_small_arith')
    g = load_golden('hetero_small_arith')
    m = gpu(s.topo, configs.params_for(s, cut, ew, precision=abi.PREC_FP64))
    x = m.squeeze(s.coor) if sqz else s.coor
    m.nbondm(x)
    t = s.topo
    with pytest.raises(MoilGPUError, match='arith'):
        m.set_params(t.ptchg, t.epsgm6, t.epsgm12, 0.5, 0.5)
    with pytest.raises(MoilGPUError, match='arith'):
        m.param_grad(x)
    e, dp = m.eforce(x)
    assert rms_rel(dp, g['dpot']) < 1e-10
    m.close()
