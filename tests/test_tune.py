"""Launch shapes (moil_gpu_set_tuning, moil_gpu_autotune) decide how an
evaluation maps onto the GPU and nothing else: the chunks write their own
partial rows and the forces are summed in fixed point, so the gradient is
bit-identical for every water/atom chunk budget, water-kernel grid and PME
spreading tile shape; the energies are FP64 sums per workgroup, and a
chunk budget regroups their terms, so they agree to summation order
(1e-13 relative).  The autotuner leaves the fastest
shapes applied (never slower than the shapes it started from, which it
times first), and a second context of the same system takes them from the
process's cache without timing again."""
import numpy as np
import pytest

from moil_fnrfd_amd import abi, configs
from tests._cases import make_case

pytestmark = pytest.mark.gpu

SHAPES = [
    dict(wchunk_pairs=1024, water_wg_per_cu=0, pme_tile=5),
    dict(wchunk_pairs=8192, achunk_pairs=2048, water_wg_per_cu=2, pme_tile=6),
    dict(wchunk_pairs=4096, achunk_pairs=16384, water_wg_per_cu=5, pme_tile=7),
    dict(wchunk_pairs=2048, achunk_pairs=8192, water_wg_per_cu=4, pme_tile=8),
]


def _same_energies(e, e0, what):
    for k in e0:
        assert abs(e[k] - e0[k]) <= 1e-13 * max(abs(e0[k]), 1.0), (what, k, e[k], e0[k])


@pytest.mark.parametrize('name,prec', [('water216', abi.PREC_MIXED),
                                       ('hetero_small', abi.PREC_MIXED),
                                       ('hetero_small', abi.PREC_FP64)],
                         ids=['water216-mixed', 'hetero_small-mixed', 'hetero_small-fp64'])
def test_shapes_leave_results_identical(name, prec):
    import torch
    from moil_fnrfd_amd.eforce import MoilGPU
    print('case tuning %s prec=%d' % (name, prec), flush=True)
    s, cut, ew, sqz = make_case(name)
    p = configs.params_for(s, cut, ew, precision=prec)
    m = MoilGPU(s.topo, p)
    x = m.squeeze(s.coor) if sqz else s.coor
    m.nbondm(x)
    e0, f0 = m.eforce(x)
    assert m.get_tuning()['pme_tile'] == 8
    for shapes in SHAPES:
        m.set_tuning(**shapes)
        m.nbondm(x)
        e, f = m.eforce(x)
        assert all(m.get_tuning()[k] == v for k, v in shapes.items())
        assert np.array_equal(f, f0), shapes
        _same_energies(e, e0, shapes)
    with pytest.raises(Exception):
        m.set_tuning(pme_tile=9)
    d_x = torch.tensor(x, dtype=torch.float64, device='cuda:0')
    t = m.autotune_dev(d_x.data_ptr(), reps=3)
    print('  autotuned: %s' % t, flush=True)
    assert t['us_tuned'] <= t['us_default']
    e, f = m.eforce(x)
    assert np.array_equal(f, f0)
    _same_energies(e, e0, t)
    # the same system again: shapes from the cache, no timing
    m2 = MoilGPU(s.topo, p)
    m2.nbondm(x)
    t2 = m2.autotune_dev(d_x.data_ptr(), reps=3)
    assert t2 == t
    e2, f2 = m2.eforce(x)
    assert np.array_equal(f2, f0)
    _same_energies(e2, e0, t2)
    m.close()
    m2.close()
