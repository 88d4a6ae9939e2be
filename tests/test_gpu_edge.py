"""Edge cases of the GPU path: MOIL's error conditions turned into status
codes (ewald_init's neutrality STOP, symm/ewald_init.f:62-72; calls out of
order), and the batched forms over systems without water (whose water
lists are empty).  The FP64 parity of those systems against MOIL itself is
covered by test_gpu_parity.py (cases chains_only and chains_one_water)."""
import copy

import numpy as np
import pytest

from moil_fnrfd_amd import abi, configs
from tests._cases import make_case
from tests._util import load_golden, rms_rel

pytestmark = pytest.mark.gpu


def test_init_refuses_a_charged_system():
    from moil_fnrfd_amd.eforce import MoilGPU, MoilGPUError
    s, cut, ew = configs.chains_only()
    topo = copy.copy(s.topo)
    q = topo.ptchg.copy()
    q[0] += 0.01
    topo.ptchg = q
    print('case charged system', flush=True)
    with pytest.raises(MoilGPUError, match='error %d' % abi.ECHARGE):
        MoilGPU(topo, configs.params_for(s, cut, ew))


def test_calls_out_of_order():
    from moil_fnrfd_amd.eforce import MoilGPU, MoilGPUError
    s, cut, ew = configs.chains_only()
    m = MoilGPU(s.topo, configs.params_for(s, cut, ew))
    print('case eforce before lists', flush=True)
    with pytest.raises(MoilGPUError, match='error %d' % abi.ESTATE):
        m.eforce(s.coor)
    with pytest.raises(MoilGPUError, match='error %d' % abi.ESTATE):
        m.param_grad(s.coor)
    m.nbondm(s.coor)
    e, _ = m.eforce(s.coor)
    assert np.isfinite(e['e_total'])
    m.close()


@pytest.mark.parametrize('prec', [abi.PREC_FP64, abi.PREC_MIXED], ids=['fp64', 'mixed'])
def test_batch_of_systems_without_water(prec):
    """A batch mixing a water-free system, a one-water system and a water
    box: batched lists and evaluation equal MOIL's outputs."""
    import torch
    from moil_fnrfd_amd.eforce import MoilBatch, MoilGPU
    torch.cuda.set_device(0)
    dev = torch.device('cuda:0')
    sh = torch.cuda.current_stream().cuda_stream
    names = ['chains_only', 'chains_one_water', 'water216']
    ms, xs, dps, ens, golden = [], [], [], [], []
    for name in names:
        s, cut, ew, sqz = make_case(name)
        ms.append(MoilGPU(s.topo, configs.params_for(s, cut, ew, precision=prec)))
        xs.append(torch.tensor(s.coor, dtype=torch.float64, device=dev))
        dps.append(torch.empty_like(xs[-1]))
        ens.append(torch.zeros(abi.N_ENERGIES, dtype=torch.float64, device=dev))
        golden.append((load_golden(name), sqz))
    print('case batch without water prec=%d' % prec, flush=True)
    b = MoilBatch(ms)
    # squeeze where the fixture was squeezed, lists for all in one call
    for i, (g, sqz) in enumerate(golden):
        if not sqz:
            continue
        ms[i].squeeze_dev(xs[i].data_ptr(), sh)
    b.build_lists_dev([x.data_ptr() for x in xs], squeeze=False, stream=sh)
    b.eforce_dev([x.data_ptr() for x in xs], [d.data_ptr() for d in dps],
                 [e.data_ptr() for e in ens], sh)
    torch.cuda.synchronize()
    tol_e, tol_g = (1e-12, 1e-10) if prec == abi.PREC_FP64 else (1e-6, 1e-5)
    for i, (g, _) in enumerate(golden):
        en = ens[i].cpu().numpy()
        ref_total = g['energies'][14]
        scale = float(np.abs(g['energies'][[1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13]]).sum())
        assert abs(en[0] - ref_total) <= tol_e * scale, names[i]
        assert rms_rel(dps[i].cpu().numpy(), g['dpot']) < tol_g, names[i]
    b.close()
    for m in ms:
        m.close()


@pytest.mark.parametrize('name', ['hetero_small', 'water216'])
def test_context_from_a_con_file(name, tmp_path):
    """A context initialised from a .con file read by the library's reader
    evaluates exactly as one built from the topology the file was written
    from, and to the FP64 bars against MOIL."""
    import os
    from moil_fnrfd_amd.eforce import MoilCon, MoilGPU
    from moil_fnrfd_amd.systems import write_con
    s, cut, ew, sqz = make_case(name)
    path = os.path.join(tmp_path, name + '.con')
    write_con(path, s.topo)
    p = configs.params_for(s, cut, ew, precision=abi.PREC_FP64)
    print('case %s from .con' % name, flush=True)
    out = []
    for topo in (s.topo, MoilCon(path)):
        m = MoilGPU(topo, p)
        x = m.squeeze(s.coor) if sqz else s.coor
        m.nbondm(x)
        out.append(m.eforce(x))
        m.close()
    (e0, f0), (e1, f1) = out
    assert e0 == e1
    assert np.array_equal(f0, f1)
    g = load_golden(name)
    assert rms_rel(f1, g['dpot']) < 1e-10
