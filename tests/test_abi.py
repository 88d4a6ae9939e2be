"""The C ABI library loads on a CPU-only host and exports every entry
point declared in include/moil_gpu.h (no compute calls here)."""
import ctypes
import os
import re

from moil_fnrfd_amd import abi, eforce

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, 'include', 'moil_gpu.h')).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(moil_gpu_\w+)\s*\(', txt)))


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
