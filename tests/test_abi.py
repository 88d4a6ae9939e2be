"""The C ABI library loads on a CPU-only host and exports every entry
point declared in include/moil_gpu.h (no compute calls here)."""
import ctypes

import numpy as np
import os
import re

from moil_fnrfd_amd import abi, eforce

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, 'include', 'moil_gpu.h')).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(moil_\w+)\s*\(', txt)))


def test_library_exports_all_declared_symbols():
    L = eforce.load_library()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s


def test_abi_version_and_struct_layout():
    L = eforce.load_library()
    assert L.moil_gpu_abi_version() == abi.ABI_VERSION
    # struct sizes as the C compiler lays them out (x86-64 SysV)
    assert ctypes.sizeof(abi.MoilEnergies) == 15 * 8
    assert ctypes.sizeof(abi.MoilParams) == 10 * 8 + 6 * 4
    assert abi.MoilTopology.npt.offset == 0
    assert abi.MoilTopology.poipt.offset == 8


def test_init_rejects_bad_input_without_gpu():
    """Argument validation happens before any device call."""
    from moil_fnrfd_amd import configs
    s, cut, ew = configs.water216()
    p = configs.params_for(s, cut, ew)
    p.order = 6
    try:
        eforce.MoilGPU(s.topo, p)
        raise AssertionError('expected MoilGPUError')
    except eforce.MoilGPUError as e:
        assert 'order' in str(e)
    p = configs.params_for(s, cut, ew)
    s.topo.ptchg[0] += 0.01  # non-neutral: MOIL STOPs in ewald_init
    try:
        eforce.MoilGPU(s.topo, p)
        raise AssertionError('expected MoilGPUError')
    except eforce.MoilGPUError as e:
        assert 'charge' in str(e)


def _expect_einval(topo, p, needle):
    try:
        eforce.MoilGPU(topo, p)
        raise AssertionError('expected MoilGPUError')
    except eforce.MoilGPUError as e:
        assert needle in str(e), str(e)


def test_init_rejects_charges_beyond_the_grid_range_without_gpu():
    """The reciprocal grid's fixed-point conversion holds stencil weights
    |q| sqrt(kofdie/eps) < 2^11: a neutral system holding charges of
    +-120 e (weight 2187) is refused before any device call."""
    from moil_fnrfd_amd import configs
    s, cut, ew = configs.water216()
    p = configs.params_for(s, cut, ew)
    o1, o2 = 0, 3  # two water O atoms
    q0 = (s.topo.ptchg[o1], s.topo.ptchg[o2])
    s.topo.ptchg[o1], s.topo.ptchg[o2] = 120.0, -120.0 + q0[0] + q0[1]
    _expect_einval(s.topo, p, 'fixed-point range')


def test_init_rejects_arith_inputs_it_cannot_evaluate_without_gpu():
    """Under the arith rule a torsion of period outside 1..4 has no MOIL
    energy (pot/ephi.f:113-138 sets none), and unknown term bits are not
    guessed at: both refused before any device call."""
    import dataclasses
    from moil_fnrfd_amd import configs
    s, cut, ew = configs.hetero_small_arith()
    p = configs.params_for(s, cut, ew)
    assert p.terms & abi.TERM_ARITH
    per = s.topo.period.copy()
    per[3] = 0
    _expect_einval(dataclasses.replace(s.topo, period=per), p, 'period 0')
    per[3] = 5
    _expect_einval(dataclasses.replace(s.topo, period=per), p, 'period 5')
    p.terms |= 1 << 12
    _expect_einval(s.topo, p, 'unknown term bits')


def test_init_rejects_out_of_range_indices_without_gpu():
    """Every index array is checked against 1..npt (1..totmon for the water
    list) and the CSR pointers for monotonicity before anything reads them
    (no device call is reached)."""
    import copy
    from moil_fnrfd_amd import configs
    s, cut, ew = configs.hetero_small()
    p = configs.params_for(s, cut, ew)
    for field, k, bad, needle in [('ib1', 3, 0, 'ib1(4) = 0'),
                                  ('iangl3', 0, s.topo.npt + 1, 'iangl3(1)'),
                                  ('itor2', 5, -7, 'itor2(6)'),
                                  ('iimp4', 1, 0, 'iimp4(2)'),
                                  ('spec2', 2, s.topo.npt + 5, 'spec2(3)'),
                                  ('exc2', 10, 0, 'exclusion partner 0'),
                                  ('idxtip3', 0, s.topo.totmon + 1, 'idxtip3(1)')]:
        t = copy.deepcopy(s.topo)
        a = getattr(t, field).copy()
        a[k] = bad
        setattr(t, field, a)
        _expect_einval(t, p, needle)
    t = copy.deepcopy(s.topo)
    a = t.exc1.copy()
    a[7] = a[8] + 1
    t.exc1 = a
    _expect_einval(t, p, 'exc1 decreases')
    t = copy.deepcopy(s.topo)
    a = t.poipt.copy()
    a[3], a[4] = a[4], a[3]
    t.poipt = a
    _expect_einval(t, p, 'poipt decreases')


def test_con_file_with_blank_bond_atom_is_refused_at_init(tmp_path):
    """A .con record whose bond atom field is blank reads as atom 0 (MOIL's
    i7 edit descriptor reads blanks as zero); moil_gpu_init refuses it."""
    from moil_fnrfd_amd import configs
    from moil_fnrfd_amd.systems import write_con
    s, cut, ew = configs.hetero_small()
    path = os.path.join(str(tmp_path), 'h.con')
    write_con(path, s.topo)
    lines = open(path).read().splitlines()
    k = next(i for i, l in enumerate(lines) if l.startswith('~ ib1 ib2'))
    rec = lines[k + 1]
    lines[k + 1] = ' ' + ' ' * 7 + rec[8:]
    open(path, 'w').write('\n'.join(lines) + '\n')
    c = eforce.MoilCon(path)
    assert c.array('ib1', 1)[0] == 0
    _expect_einval(c, configs.params_for(s, cut, ew), 'ib1(1) = 0')
    c.close()


def test_init_reads_moils_trailing_exclusion_pointers():
    """MOIL's rconn leaves exc1 at 0 after the last atom with an exclusion
    record (generic/rconn.f:313-338); moil_gpu_init reads those entries as
    empty ranges (it then proceeds to the device: no GPU here)."""
    import copy
    from moil_fnrfd_amd import configs
    s, cut, ew = configs.water216()
    t = copy.deepcopy(s.topo)
    a = t.exc1.copy()
    last = int(np.nonzero(np.diff(a))[0].max()) + 1  # last atom owning exclusions
    a[last + 1:] = 0
    assert last + 1 < len(a)
    t.exc1 = a
    try:
        g = eforce.MoilGPU(t, configs.params_for(s, cut, ew))
        g.close()
    except eforce.MoilGPUError as e:
        assert 'no HIP device' in str(e), str(e)
