"""The molecular virial on the GPU (moil_gpu_eforce_virial[_dev], MOIL's
eforce with pressON) against MOIL's own run (golden fixtures
tests/golden/press_*.npz): the real-space terms are evaluated in FP64 with
MOIL's erfc table in either precision mode, so they meet the FP64 bar
(<= 1e-12 per term) in both; the reciprocal terms follow the context's
precision (FP64: <= 1e-11, since rocFFT's transforms round differently
from MOIL's own FFT and the k-space sums cancel in part; mixed mode: FP32
grid and transforms, <= 2e-5).  The evaluation
that carries the virial returns the same energies and gradient as plain
eforce, bit for bit."""
import numpy as np
import pytest
import torch

from moil_fnrfd_amd import abi, configs
from moil_fnrfd_amd.eforce import MoilGPU, MoilGPUError, molecules
from oracle import ref_runner
from tests._cases import make_case
from tests._util import load_golden
from tests.golden.gen_golden import PRESS_CASES
from tests.test_virial_oracle import virial_errors

pytestmark = pytest.mark.gpu

PRECS = {'fp64': abi.PREC_FP64, 'mixed': abi.PREC_MIXED}


@pytest.mark.parametrize('prec', sorted(PRECS))
@pytest.mark.parametrize('name', PRESS_CASES)
def test_gpu_virial_matches_moil(name, prec):
    print('case', name, prec, flush=True)
    g = load_golden('press_' + name)
    s, cut, ew, sqz = make_case(name)
    eng = MoilGPU(s.topo, configs.params_for(s, cut, ew, precision=PRECS[prec]))
    x = eng.squeeze(s.coor) if sqz else s.coor
    eng.nbondm(x)
    eng.set_molecules(g['patom'], g['ordlist'], g['pmol'], g['ptms'])
    e, dp, v = eng.eforce_virial(x)
    err = virial_errors(v, g['virial'])
    real = [n for n in abi.VIRIAL_NAMES if not n.startswith(('V_PIrec', 'V_PIrel'))]
    recip = [n for n in abi.VIRIAL_NAMES if n.startswith(('V_PIrec', 'V_PIrel'))]
    assert max(err[n] for n in real) < 1e-12, err
    assert max(err[n] for n in recip) < (1e-11 if prec == 'fp64' else 2e-5), err
    e0, dp0 = eng.eforce(x)
    assert np.array_equal(dp, dp0)
    assert e == e0
    eng.close()


def test_gpu_virial_device_pointers_and_repeats():
    """The device-pointer form on the 1,000-water box: two calls give the
    same virial bit for bit, equal to the host form's, and the mixed-mode
    reciprocal terms stay within 2e-5 of the FP64 context's."""
    s, cut, ew = configs.water_buffered()
    patom, ordlist, pmol = molecules(s.topo)
    out = {}
    for prec in ('fp64', 'mixed'):
        eng = MoilGPU(s.topo, configs.params_for(s, cut, ew, precision=PRECS[prec]))
        eng.nbondm(s.coor)
        eng.set_molecules(patom, ordlist, pmol, s.topo.ptms)
        dev = torch.device('cuda:0')
        xc = torch.tensor(s.coor, dtype=torch.float64, device=dev)
        dp = torch.empty_like(xc)
        en = torch.zeros(abi.N_ENERGIES, dtype=torch.float64, device=dev)
        vs = []
        for _ in range(2):
            vd = torch.zeros(abi.N_VIRIAL, dtype=torch.float64, device=dev)
            eng.eforce_virial_dev(xc.data_ptr(), dp.data_ptr(), en.data_ptr(), vd.data_ptr(),
                                  torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            vs.append(vd.cpu().numpy())
        assert np.array_equal(vs[0], vs[1])
        _, _, vh = eng.eforce_virial(s.coor)
        assert np.array_equal(vh, vs[0])
        out[prec] = vs[0]
        eng.close()
    err = virial_errors(out['mixed'], out['fp64'])
    assert max(err.values()) < 2e-5, err


def test_gpu_virial_refusals():
    s, cut, ew = configs.water216()
    eng = MoilGPU(s.topo, configs.params_for(s, cut, ew))
    eng.nbondm(s.coor)
    with pytest.raises(MoilGPUError, match='set_molecules'):
        eng.eforce_virial(s.coor)
    patom, ordlist, pmol = molecules(s.topo)
    bad = pmol.copy()
    bad[5] = bad[4] + 1
    with pytest.raises(MoilGPUError, match='molecule'):
        eng.set_molecules(patom, ordlist, bad, s.topo.ptms)
    with pytest.raises(MoilGPUError, match='no molecule'):
        eng.set_molecules(patom[:-1], ordlist, pmol, s.topo.ptms)
    eng.set_molecules(patom, ordlist, pmol, s.topo.ptms)
    e, dp, v = eng.eforce_virial(s.coor)
    assert np.isfinite(v).all()
    eng.set_molecules([], [], [], [])  # removed again
    with pytest.raises(MoilGPUError, match='set_molecules'):
        eng.eforce_virial(s.coor)
    eng.close()


@pytest.mark.skipif(not ref_runner.available(ref_runner.SHIM_EXE),
                    reason='oracle/_ref/moil_dump_gpu not built')
@pytest.mark.parametrize('name', ['water216_out_sqz', 'hetero_small'])
def test_moil_pressure_path_through_shim(name):
    """MOIL's driver in its pressure mode (init_pressure's molecules, their
    centres of mass before eforce) linked with the Fortran shim: the shim
    hands the molecules to the library and fills PRESS.BLOCK's virial
    terms, which match MOIL's own run to the FP64 bars."""
    print('case %s shim pressure' % name, flush=True)
    s, cut, ew, sqz = make_case(name)
    g = load_golden('press_' + name)
    d = ref_runner.run_moil(s, cut, ew, squeeze=sqz, dump_lists=False, timeout=300,
                            exe=ref_runner.SHIM_EXE, env={'MOIL_GPU_FP64': '1'}, mode=5)
    err = virial_errors(d['VIRIAL'], g['virial'])
    recip = [n for n in abi.VIRIAL_NAMES if n.startswith(('V_PIrec', 'V_PIrel'))]
    assert max(v for n, v in err.items() if n not in recip) < 1e-12, err
    assert max(err[n] for n in recip) < 1e-11, err
    assert np.array_equal(d['PATOM'], g['patom'])
