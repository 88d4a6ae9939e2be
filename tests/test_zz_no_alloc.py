"""moil_gpu_eforce_noalloc_dev (the evaluation with its reduction scratch in
the library's static device array) against moil_gpu_eforce_dev, both
called through the C ABI on device pointers.

Two calls in a row on two contexts, the second with a smaller system
(1386-atom solute + water system, then the 648-atom water box): the
gradient and the energy record of every call are bit-identical to
moil_gpu_eforce_dev's on the same inputs (the same kernels sum the same
fixed-point terms in the same order; only the scratch's address differs).
Then, in FP64 mode, one water moved onto its nearest neighbour (O-O 0.5 A:
a Lennard-Jones force beyond the fixed-point range) gives NaN energies
from both entry points, and the next call on the unchanged coordinates is
again identical and finite (the overflow word and the reduction's ticket
start at zero on every call).  The calls leave the device's free memory
unchanged.
"""
import numpy as np
import pytest

from moil_fnrfd_amd import abi, configs
from tests._cases import make_case

pytestmark = pytest.mark.gpu


def _context(name, prec, dev, sh):
    from moil_fnrfd_amd.eforce import MoilGPU
    import torch
    s, cut, ew, sqz = make_case(name)
    m = MoilGPU(s.topo, configs.params_for(s, cut, ew, precision=prec))
    x = torch.tensor(s.coor, dtype=torch.float64, device=dev)
    if sqz:
        m.squeeze_dev(x.data_ptr(), sh)
    m.nbondm_dev(x.data_ptr(), sh)
    return s, m, x


def _both(m, x, sh):
    import torch
    dp_a, dp_b = torch.full_like(x, 7.0), torch.full_like(x, -7.0)
    en_a = torch.full((abi.N_ENERGIES,), 3.0, dtype=torch.float64, device=x.device)
    en_b = torch.full((abi.N_ENERGIES,), -3.0, dtype=torch.float64, device=x.device)
    m.eforce_dev(x.data_ptr(), dp_a.data_ptr(), en_a.data_ptr(), sh)
    m.eforce_noalloc_dev(x.data_ptr(), dp_b.data_ptr(), en_b.data_ptr(), sh)
    torch.cuda.synchronize()
    return dp_a.cpu().numpy(), en_a.cpu().numpy(), dp_b.cpu().numpy(), en_b.cpu().numpy()


@pytest.mark.parametrize('prec', [abi.PREC_MIXED, abi.PREC_FP64], ids=['mixed', 'fp64'])
def test_noalloc_matches_eforce_dev(prec):
    import torch
    torch.cuda.set_device(0)
    dev = torch.device('cuda:0')
    sh = torch.cuda.current_stream().cuda_stream
    ctxs = []
    for name in ('hetero_small', 'water216'):  # the second call is the smaller system
        print('case noalloc %s prec=%d' % (name, prec), flush=True)
        s, m, x = _context(name, prec, dev, sh)
        ctxs.append((name, s, m, x))
        dp_a, en_a, dp_b, en_b = _both(m, x, sh)
        assert np.isfinite(en_a).all()
        assert np.array_equal(dp_a, dp_b), name
        assert np.array_equal(en_a, en_b), name
    # no allocation: the device's free memory is the same after more calls
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    dps = [torch.empty_like(x) for _, _, _, x in ctxs]
    ens = [torch.empty(abi.N_ENERGIES, dtype=torch.float64, device=dev) for _ in ctxs]
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    for rep in range(3):
        for k, (_, _, m, x) in enumerate(ctxs):
            m.eforce_noalloc_dev(x.data_ptr(), dps[k].data_ptr(), ens[k].data_ptr(), sh)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free1, (free0, free1)
    for _, _, m, _ in ctxs:
        m.close()


def test_noalloc_overflow_word_cleared():
    import torch
    torch.cuda.set_device(0)
    dev = torch.device('cuda:0')
    sh = torch.cuda.current_stream().cuda_stream
    print('case noalloc overflow water216 fp64', flush=True)
    s, m, x = _context('water216', abi.PREC_FP64, dev, sh)
    x0 = x.clone()
    c = x0.cpu().numpy()
    o = np.asarray(s.topo.poipt)[np.asarray(s.topo.idxtip3) - 1]  # O atom of each water
    d = c[o[1:]] - c[o[0]]
    L = np.asarray(s.box)
    d -= L * np.round(d / L)
    r = np.sqrt((d * d).sum(1))
    k = int(np.argmin(r))
    shift = d[k] * (0.5 / r[k] - 1.0)  # partner's O to 0.5 A from water 0's O
    c2 = c.copy()
    a0 = int(o[k + 1])
    c2[a0:a0 + 3] += shift
    x.copy_(torch.tensor(c2, dtype=torch.float64, device=dev))
    dp_a, en_a, dp_b, en_b = _both(m, x, sh)
    assert np.isnan(en_a).all() and np.isnan(en_b).all()
    x.copy_(x0)
    dp_a, en_a, dp_b, en_b = _both(m, x, sh)
    assert np.isfinite(en_a).all() and np.isfinite(en_b).all()
    assert np.array_equal(dp_a, dp_b) and np.array_equal(en_a, en_b)
    m.close()
