"""Captured evaluations (MOIL_GPU_GRAPH=1): moil_gpu_eforce[_dev] replays a
hipGraph of the evaluation (at timing level 1 except the first evaluation after a list
change, which runs in order and is timed) and runs in order at level 2.  Every
replay must reproduce the in-order result bit for bit, with unrelated
device work between replays, after a list rebuild (a new capture), with
new device pointers, after new parameters, in both precisions and on a
system with water groups, atom lists, bonded terms and PME."""
import numpy as np
import pytest

from moil_fnrfd_amd import abi, configs
from tests._cases import make_case

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason='captured-graph replay is off by default: a process that '
                     'recaptured ~40 times in a parameter finite-difference loop aborted '
                     'with the cause not identified; kept out of GPU runs until it is'),
]


def _eager(m, x):
    m.set_timing(2)
    e, f = m.eforce(x)
    m.stage_times()
    m.set_timing(0)
    return e, f


@pytest.mark.parametrize('prec', [abi.PREC_MIXED, abi.PREC_FP64], ids=['mixed', 'fp64'])
@pytest.mark.parametrize('name', ['hetero_small', 'water1000_buffered'])
def test_graph_replays_match_in_order(name, prec):
    import torch
    from moil_fnrfd_amd.eforce import MoilGPU
    torch.cuda.set_device(0)
    print('case graph %s prec=%d' % (name, prec), flush=True)
    s, cut, ew, sqz = make_case(name)
    m = MoilGPU(s.topo, configs.params_for(s, cut, ew, precision=prec))
    x = m.squeeze(s.coor)
    m.nbondm(x)
    e_ref, f_ref = _eager(m, x)
    junk = torch.randn(1 << 20, device='cuda:0')
    for rep in range(4):
        float((junk * rep).sum())  # device work between replays
        e, f = m.eforce(x)
        assert np.array_equal(f, f_ref) and e == e_ref, rep
    # new coordinates and lists: a new capture
    rng = np.random.default_rng(3)
    x2 = x + rng.normal(0.0, 0.02, x.shape)
    m.nbondm(x2)
    e_ref2, f_ref2 = _eager(m, x2)
    for rep in range(3):
        float(junk.sum())
        e, f = m.eforce(x2)
        assert np.array_equal(f, f_ref2) and e == e_ref2, rep
    # device pointers: two sets alternating
    dev = torch.device('cuda:0')
    sh = torch.cuda.current_stream().cuda_stream
    xs = [torch.tensor(x2, device=dev) for _ in range(2)]
    ds = [torch.empty_like(xs[0]) for _ in range(2)]
    es = [torch.zeros(abi.N_ENERGIES, dtype=torch.float64, device=dev) for _ in range(2)]
    for rep in range(4):
        k = rep % 2
        m.eforce_dev(xs[k].data_ptr(), ds[k].data_ptr(), es[k].data_ptr(), sh)
        torch.cuda.synchronize()
        assert np.array_equal(ds[k].cpu().numpy(), f_ref2), rep
        assert float(es[k][0]) == e_ref2['e_total']
    # timing level 1: the first evaluation after a rebuild runs in order with
    # the real-space kernel's dispatch events, the following ones replay
    m.set_timing(1)
    m.nbondm(x2)
    for rep in range(3):
        e, f = m.eforce(x2)
        assert np.array_equal(f, f_ref2)
    t = m.stage_times()
    assert t[5] == 1 and t[0] > 0.0
    m.set_timing(0)
    # new parameters: charges scaled, captured values must follow
    q = s.topo.ptchg * 1.01
    m.set_params(q, s.topo.epsgm6, s.topo.epsgm12, 0.5, 0.5)
    e_ref3, f_ref3 = _eager(m, x2)
    assert not np.array_equal(f_ref3, f_ref2)
    for rep in range(2):
        e, f = m.eforce(x2)
        assert np.array_equal(f, f_ref3) and e == e_ref3, rep
    m.close()
