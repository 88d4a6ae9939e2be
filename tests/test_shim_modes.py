"""The Fortran shim refuses MOIL modes the library does not evaluate.

MOIL's dump driver linked with the shim (oracle/_ref/moil_dump_gpu) switches
one MOIL mode on after reading its input; the shim's ewald_init must stop
the program through MOIL's alert(...,1) (intrprtr/alert.f:10-24: RED ALERT,
then STOP) before any GPU call, instead of returning energies that differ
from MOIL's.  Runs without a GPU: nothing reaches the library."""
import pytest

from oracle import ref_runner
from tests._cases import make_case

pytestmark = pytest.mark.skipif(not ref_runner.available(ref_runner.SHIM_EXE),
                                reason='oracle/_ref/moil_dump_gpu not built')


@pytest.mark.parametrize('mode,why', [(1, 'distance-dependent dielectric'),
                                      (2, 'pressure / virial'),
                                      (3, 'frozen particles')])
def test_shim_stops_on_unsupported_mode(mode, why):
    s, cut, ew, sqz = make_case('hetero_small')
    with pytest.raises(RuntimeError) as ei:
        ref_runner.run_moil(s, cut, ew, squeeze=sqz, dump_lists=False, timeout=120,
                            exe=ref_runner.SHIM_EXE, mode=mode)
    msg = str(ei.value)
    assert 'RED ALERT' in msg, msg
    assert 'ROUTINE ewald_init' in msg, msg
    assert 'not supported by the GPU library: ' + why in msg, msg
