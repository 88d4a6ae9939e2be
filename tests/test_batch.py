"""Batched evaluation (moil_gpu_batch_*): several systems of different
sizes and kinds evaluated by one call that replays a captured hipGraph of
multi-system launches (one grid per kernel over all systems, one batched
transform per grid size).  Every replay must reproduce one
moil_gpu_eforce_dev per system bit for bit (the kernel bodies are the same
and their sums are order independent), in both precisions, also after
eager work between replays and after a list rebuild forces a new capture.
Two pairs of systems share a grid size, so batched transforms of more than
one member are covered."""
import numpy as np
import pytest

from moil_fnrfd_amd import abi, configs
from tests._cases import make_case

pytestmark = pytest.mark.gpu


def _systems():
    out = []
    for name in ('water216', 'water1000_buffered', 'hetero_small', 'hetero_small_out',
                 'hetero_small_arith'):
        s, cut, ew, sqz = make_case(name)
        out.append((name, s, cut, ew))
    s, cut, ew = configs.water_scaled(8, seed=1300)
    out.append(('water512', s, cut, ew))
    s, cut, ew = configs.water_scaled(8, seed=1301)
    out.append(('water512b', s, cut, ew))
    return out


@pytest.mark.parametrize('prec', [abi.PREC_MIXED, abi.PREC_FP64], ids=['mixed', 'fp64'])
def test_batch_matches_plain_calls(prec):
    import torch
    from moil_fnrfd_amd.eforce import MoilBatch, MoilGPU
    torch.cuda.set_device(0)
    dev = torch.device('cuda:0')
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    systems, coor, dpot_b, en_b, dpot_p, en_p = [], [], [], [], [], []
    for name, s, cut, ew in _systems():
        print('case batch system %s npt=%d prec=%d' % (name, s.topo.npt, prec), flush=True)
        m = MoilGPU(s.topo, configs.params_for(s, cut, ew, precision=prec))
        x = torch.tensor(s.coor, dtype=torch.float64, device=dev)
        m.squeeze_dev(x.data_ptr(), sh)
        m.nbondm_dev(x.data_ptr(), sh)
        systems.append(m)
        coor.append(x)
        for lst in (dpot_b, dpot_p):
            lst.append(torch.empty_like(x))
        for lst in (en_b, en_p):
            lst.append(torch.zeros(abi.N_ENERGIES, dtype=torch.float64, device=dev))
    batch = MoilBatch(systems, nstreams=3)

    def plain():
        for i, m in enumerate(systems):
            m.eforce_dev(coor[i].data_ptr(), dpot_p[i].data_ptr(), en_p[i].data_ptr(), sh)

    def batched():
        batch.eforce_dev([c.data_ptr() for c in coor], [d.data_ptr() for d in dpot_b],
                         [e.data_ptr() for e in en_b], sh)

    def same():
        torch.cuda.synchronize()
        for i in range(len(systems)):
            assert torch.equal(dpot_b[i], dpot_p[i]), i
            assert torch.equal(en_b[i], en_p[i]), i

    for rep in range(3):
        print('case batch replay %d' % rep, flush=True)
        for d in dpot_b:
            d.fill_(0.0)
        plain()
        batched()
        same()
    # coordinates move and one system rebuilds its lists: new capture
    gen = torch.Generator(device=dev)
    gen.manual_seed(7)
    for rep in range(2):
        print('case batch after rebuild %d' % rep, flush=True)
        for x in coor:
            x.add_(torch.randn(x.shape, generator=gen, device=dev, dtype=torch.float64), alpha=0.02)
        systems[2].nbondm_dev(coor[2].data_ptr(), sh)
        plain()
        batched()
        same()
    batch.close()
    for m in systems:
        m.close()


def test_batch_refuses_mixed_precisions():
    import torch
    from moil_fnrfd_amd.eforce import MoilBatch, MoilGPU
    torch.cuda.set_device(0)
    ms = []
    for prec in (abi.PREC_MIXED, abi.PREC_FP64):
        s, cut, ew = configs.water216()
        ms.append(MoilGPU(s.topo, configs.params_for(s, cut, ew, precision=prec)))
    with pytest.raises(RuntimeError, match='same precision'):
        MoilBatch(ms)
    for m in ms:
        m.close()


@pytest.mark.parametrize('prec', [abi.PREC_MIXED, abi.PREC_FP64], ids=['mixed', 'fp64'])
def test_batch_list_builds_and_direct_launches(prec):
    """moil_gpu_batch_build_lists_dev (all systems' squeeze + list builds
    side by side) and graph replay switched off: after every move, the
    batch's results equal those of separate contexts that squeezed, built
    and evaluated one system at a time, bit for bit; switching the graph
    back on changes nothing."""
    import torch
    from moil_fnrfd_amd.eforce import MoilBatch, MoilGPU
    torch.cuda.set_device(0)
    dev = torch.device('cuda:0')
    sh = torch.cuda.current_stream().cuda_stream
    mine, ref, x_b, x_p = [], [], [], []
    for name, s, cut, ew in _systems():
        p = configs.params_for(s, cut, ew, precision=prec)
        mine.append(MoilGPU(s.topo, p))
        ref.append(MoilGPU(s.topo, p))
        x = torch.tensor(s.coor, dtype=torch.float64, device=dev)
        x_b.append(x)
        x_p.append(x.clone())
    n = len(mine)
    dp_b = [torch.empty_like(x) for x in x_b]
    dp_p = [torch.empty_like(x) for x in x_b]
    en_b = [torch.zeros(abi.N_ENERGIES, dtype=torch.float64, device=dev) for _ in range(n)]
    en_p = [torch.zeros(abi.N_ENERGIES, dtype=torch.float64, device=dev) for _ in range(n)]
    batch = MoilBatch(mine)
    batch.set_graph(False)
    gen = torch.Generator(device=dev)
    gen.manual_seed(11)
    for rep in range(3):
        print('case batch lists + direct launch %d prec=%d' % (rep, prec), flush=True)
        if rep:
            for xb, xp in zip(x_b, x_p):
                d = torch.randn(xb.shape, generator=gen, device=dev, dtype=torch.float64)
                xb.add_(d, alpha=0.05)
                xp.add_(d, alpha=0.05)
        batch.build_lists_dev([x.data_ptr() for x in x_b], squeeze=True, stream=sh)
        batch.eforce_dev([x.data_ptr() for x in x_b], [d.data_ptr() for d in dp_b],
                         [e.data_ptr() for e in en_b], sh)
        for i, m in enumerate(ref):
            m.squeeze_dev(x_p[i].data_ptr(), sh)
            m.nbondm_dev(x_p[i].data_ptr(), sh)
            m.eforce_dev(x_p[i].data_ptr(), dp_p[i].data_ptr(), en_p[i].data_ptr(), sh)
        torch.cuda.synchronize()
        for i in range(n):
            assert torch.equal(x_b[i], x_p[i]), i  # same squeeze
            assert torch.equal(dp_b[i], dp_p[i]), i
            assert torch.equal(en_b[i], en_p[i]), i
    batch.set_graph(True)
    for d in dp_b:
        d.fill_(0.0)
    batch.eforce_dev([x.data_ptr() for x in x_b], [d.data_ptr() for d in dp_b],
                     [e.data_ptr() for e in en_b], sh)
    torch.cuda.synchronize()
    for i in range(n):
        assert torch.equal(dp_b[i], dp_p[i]), i
        assert torch.equal(en_b[i], en_p[i]), i
    batch.close()
    for m in mine + ref:
        m.close()
