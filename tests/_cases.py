"""The parity systems of the golden fixtures (tests/golden/gen_golden.py):
name -> (system maker, coordinate modifier, squeeze before the lists)."""
from moil_fnrfd_amd import configs
from tests.golden.gen_golden import perturb, shift_out

CASES = {
    'water216': (configs.water216, None, False),
    'water216_out': (configs.water216, shift_out, False),
    'water216_out_sqz': (configs.water216, shift_out, True),
    'water1000_buffered': (configs.water_buffered, None, False),
    'water1000_dense': (lambda: configs.water_scaled(10, seed=1100), None, False),
    'hetero_small': (configs.hetero_small, lambda s: perturb(s, 0.05, 2001), True),
    'hetero_small_out': (lambda: configs.hetero_small(seed=2100, n_chains=6, L=20.0),
                         lambda s: perturb(shift_out(s, n_out=20, seed=78), 0.05, 2101), False),
    'hetero_small_nohvdw': (lambda: _no_hvdw(configs.hetero_small()),
                            lambda s: perturb(s, 0.05, 2001), True),
    'chains_only': (configs.chains_only, lambda s: perturb(s, 0.05, 2201), True),
    'hetero_small_phi0': (configs.hetero_small_phi0, lambda s: perturb(s, 0.05, 2401), True),
    'hetero_small_arith': (configs.hetero_small_arith, lambda s: perturb(s, 0.05, 2501), True),
    'chains_one_water': (lambda: configs.chains_only(seed=2300, n_chains=4, L=20.0, n_water=1),
                         lambda s: perturb(s, 0.05, 2301), False),
}


def _no_hvdw(made):
    """The same workload with MOIL's hydrogen rule off."""
    s, cut, ew = made
    return s, dict(cut, hvdw0=False), ew


def make_case(name):
    mk, mod, sqz = CASES[name]
    s, cut, ew = mk()
    if mod:
        s = mod(s)
    return s, cut, ew, sqz
