"""The chain functional on the GPU (moil_gpu_chain_*), MOIL's sds
(chain/sds.f): the interior replicas evaluated as one batch after their
lists were rebuilt, plus the chain's spring and repulsion terms.

Checked against MOIL's own sds run (fixture chain_water216.npz) to the
FP64 bars (energies 1e-12 of their scale, gradient 1e-10 RMS-relative)
and the mixed-precision bars, and, at the size of the batched workload
(32 replicas of a 12k-atom system), against the oracle's restatement of
the chain terms applied to the GPU's own per-replica energies and
gradients -- which must equal one plain evaluation per replica of a
squeezed copy, bit for bit."""
import numpy as np
import pytest

from moil_fnrfd_amd import abi, configs
from oracle.chain import chain_terms
from oracle import ref_runner
from tests._util import load_golden, rms_rel
from tests.golden.gen_golden import chain_case

pytestmark = pytest.mark.gpu


def _contexts(s, cut, ew, n, prec):
    from moil_fnrfd_amd.eforce import MoilGPU
    p = configs.params_for(s, cut, ew, precision=prec)
    return [MoilGPU(s.topo, p) for _ in range(n)]


def _plain_per_replica(s, cut, ew, r, prec):
    """Each interior replica alone: squeeze a copy, build lists, eforce."""
    from moil_fnrfd_amd.eforce import MoilGPU
    m = MoilGPU(s.topo, configs.params_for(s, cut, ew, precision=prec))
    igrid = r.shape[0]
    e = np.zeros(igrid)
    g = np.zeros_like(r)
    for j in range(1, igrid - 1):
        x = m.squeeze(r[j])
        m.nbondm(x)
        en, f = m.eforce(x)
        e[j] = en['e_total']
        g[j] = f
    m.close()
    return e, g


@pytest.mark.parametrize('prec', [abi.PREC_FP64, abi.PREC_MIXED], ids=['fp64', 'mixed'])
def test_chain_matches_moil_sds(prec):
    from moil_fnrfd_amd.eforce import MoilChain
    s, cut, ew, r, pick = chain_case()
    g = load_golden('chain_water216')
    gamma, rho, lam = g['params']
    igrid = r.shape[0]
    print('case chain_water216 igrid=%d npt=%d prec=%d' % (igrid, s.topo.npt, prec), flush=True)
    ch = MoilChain(_contexts(s, cut, ew, igrid - 2, prec), igrid)
    o = ch.sds(r, pick, gamma, rho, lam)
    # chain geometry: distances and factors do not depend on the precision
    assert np.allclose(o['d0'], g['d0'], rtol=1e-13, atol=0)
    assert np.allclose(o['e1'], g['e1'], rtol=1e-12, atol=1e-13 * np.abs(g['e1']).max())
    scale = float(np.abs(g['etots']).sum())
    tol_e, tol_g = (1e-12, 1e-10) if prec == abi.PREC_FP64 else (1e-6, 1e-5)
    assert abs(o['sener'] - g['sener'][0]) <= tol_e * 10 * scale
    assert np.allclose(o['e0'][1:-1], g['e0'][1:-1], rtol=0, atol=tol_e * 10 * scale)
    assert o['e0'][0] == 0.0 and o['e0'][-1] == 0.0
    assert rms_rel(o['dv'], g['dv']) < tol_g
    # the chain part of the gradient alone (replica gradients removed)
    chain_ref = g['dv'].copy()
    chain_ref[1:-1] -= g['dpots']
    e_own, g_own = _plain_per_replica(s, cut, ew, r, prec)
    chain_gpu = o['dv'] - g_own
    assert rms_rel(chain_gpu, chain_ref) < 1e-10
    # a second call with moved replicas matches a fresh evaluation
    r2 = r + np.random.default_rng(5).normal(0.0, 0.01, r.shape)
    o2 = ch.sds(r2, pick, gamma, rho, lam)
    e2, g2 = _plain_per_replica(s, cut, ew, r2, prec)
    ref2 = chain_terms(r2, pick, gamma, rho, lam, e2, g2)
    assert np.array_equal(o2['e0'][1:-1], e2[1:-1])
    assert np.abs(o2['dv'] - ref2['dv']).max() <= 1e-12 * np.abs(ref2['dv']).max()
    assert abs(o2['sener'] - ref2['sener']) <= 1e-13 * scale
    ch.close()


def test_chain_32_replicas_of_12k_atoms():
    """The chain of the batched workload's size, mixed precision: GPU chain
    == oracle chain terms over the GPU's own per-replica results, which are
    bit-identical to one plain evaluation per squeezed replica."""
    import torch
    from moil_fnrfd_amd.eforce import MoilChain, MoilGPU
    torch.cuda.set_device(0)
    base, cut, ew = configs.batch_system(12000, 4000, 8)
    igrid = 32
    r = configs.chain_replicas(base.coor, igrid, seed=4001)
    pick = configs.chain_pick(base.topo.npt, every=7, skip=2)
    gamma, rho, lam = configs.CHAIN_GAMMA, configs.CHAIN_RHO, configs.CHAIN_LAMBDA
    print('case chain 32x%d mixed' % base.topo.npt, flush=True)
    ch = MoilChain(_contexts(base, cut, ew, igrid - 2, abi.PREC_MIXED), igrid)
    dev = torch.device('cuda:0')
    d_r = torch.tensor(r, dtype=torch.float64, device=dev)
    d_pick = torch.tensor(pick, device=dev)
    d_dv = torch.empty_like(d_r)
    d_d0 = torch.empty(2 * igrid - 3, dtype=torch.float64, device=dev)
    d_e1 = torch.empty_like(d_d0)
    d_e0 = torch.zeros(igrid, dtype=torch.float64, device=dev)
    d_s = torch.zeros(1, dtype=torch.float64, device=dev)
    sh = torch.cuda.current_stream().cuda_stream
    ch.sds_dev(d_r.data_ptr(), d_pick.data_ptr(), gamma, rho, lam, d_dv.data_ptr(),
               d_d0.data_ptr(), d_e0.data_ptr(), d_e1.data_ptr(), d_s.data_ptr(), sh)
    torch.cuda.synchronize()
    assert torch.equal(d_r.cpu(), torch.tensor(r))  # replicas not squeezed in place
    dv = d_dv.cpu().numpy()
    e0 = d_e0.cpu().numpy()
    # per-replica parts: one plain evaluation per squeezed copy
    m = MoilGPU(base.topo, configs.params_for(base, cut, ew, precision=abi.PREC_MIXED))
    e_own = np.zeros(igrid)
    g_own = np.zeros_like(r)
    for j in range(1, igrid - 1):
        x = m.squeeze(r[j])
        m.nbondm(x)
        en, f = m.eforce(x)
        e_own[j] = en['e_total']
        g_own[j] = f
    m.close()
    assert np.array_equal(e0[1:-1], e_own[1:-1])
    ref = chain_terms(r, pick, gamma, rho, lam, e_own, g_own)
    assert np.allclose(d_d0.cpu().numpy(), ref['d0'], rtol=1e-13, atol=0)
    assert np.allclose(d_e1.cpu().numpy(), ref['e1'], rtol=1e-11,
                       atol=1e-12 * np.abs(ref['e1']).max())
    assert np.abs(dv - ref['dv']).max() <= 1e-12 * np.abs(ref['dv']).max()
    assert abs(float(d_s.cpu()[0]) - ref['sener']) <= 1e-13 * np.abs(e_own).sum()
    # the chain terms' gradient sums to zero atom by atom, up to the rounding
    # of adding them to the replica's own gradient (close contacts in the
    # displaced replicas reach |dpot| ~ 1e9)
    chain = dv - g_own
    bound = 8 * np.finfo(np.float64).eps * (np.abs(dv) + np.abs(g_own)).sum(axis=0)
    assert (np.abs(chain.sum(axis=0)) <= bound + 1e-13 * np.abs(chain).max()).all()
    ch.close()


def test_chain_rejects_bad_arguments():
    from moil_fnrfd_amd.eforce import MoilChain, MoilGPUError, load_library
    s, cut, ew, r, pick = chain_case()
    ctx = _contexts(s, cut, ew, 3, abi.PREC_FP64)
    with pytest.raises(MoilGPUError):
        MoilChain([], 2)  # fewer than 3 replicas
    small, cut2, ew2 = configs.water_scaled(4, seed=1302)
    other = _contexts(small, cut2, ew2, 1, abi.PREC_FP64)
    with pytest.raises(MoilGPUError):
        MoilChain([ctx[0], other[0]], 4)  # replicas of different sizes
    ch = MoilChain(ctx, 5)
    with pytest.raises(MoilGPUError):
        ch.sds(r[:4], pick, 100.0, 100.0, 2.0)
    L = load_library()
    rc = L.moil_gpu_chain_sds_dev(ch._h, None, None, 1.0, 1.0, 1.0, None, None, None, None,
                                  None, None)
    assert rc == abi.EINVAL
    ch.close()


@pytest.mark.parametrize('exe', ['sds_over_shim', 'chain_shim'])
@pytest.mark.parametrize('fp64', [True, False])
def test_moil_sds_through_shim(exe, fp64):
    """MOIL's chain driver calling the GPU library through the Fortran
    shims: MOIL's unmodified sds with each replica's lists and eforce on the
    GPU one at a time (sds_over_shim), or sds itself replaced by
    moil_gpu_chain_sds (chain_shim)."""
    path = ref_runner.SDS_SHIM_EXE if exe == 'sds_over_shim' else ref_runner.SDS_CHAIN_EXE
    if not ref_runner.available(path):
        pytest.skip('%s not built' % path)
    s, cut, ew, r, pick = chain_case()
    g = load_golden('chain_water216')
    print('case chain_water216 through %s fp64=%s' % (exe, fp64), flush=True)
    d = ref_runner.run_moil_sds(s, cut, ew, r, pick, *g['params'], timeout=300, exe=path,
                                env={'MOIL_GPU_FP64': '1' if fp64 else '0'})
    scale = float(np.abs(g['etots']).sum())
    tol_e, tol_g = (1e-12, 1e-10) if fp64 else (1e-6, 1e-5)
    assert abs(d['SENER'][0] - g['sener'][0]) <= tol_e * 10 * scale
    assert rms_rel(d['DV'], g['dv'].reshape(-1)) < tol_g
    assert np.allclose(d['D0'], g['d0'], rtol=1e-13, atol=0)
    assert np.allclose(d['E0'][1:-1], g['e0'][1:-1], rtol=0, atol=tol_e * 10 * scale)
