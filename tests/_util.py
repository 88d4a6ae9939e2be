"""Shared helpers of the parity tests."""
import hashlib
import os

import numpy as np

from moil_fnrfd_amd import abi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')

# order of the energies in a MOIL dump (oracle/ref/moil_dump.f, dump_all)
DUMP_ENERGIES = ['ewaldcof', 'e_self', 'e_corr', 'e_ew_receip', 'e_vdw', 'e_el', 'e_vsym',
                 'e_lsym', 'e_bond', 'e_theta', 'e_tors', 'e_imp', 'e_el14', 'e_vdw14',
                 'e_total', 'e_dir', 'erftbdns', 'fpme']
TERMS = ['e_bond', 'e_theta', 'e_tors', 'e_imp', 'e_vdw', 'e_el', 'e_vsym', 'e_lsym',
         'e_el14', 'e_vdw14', 'e_corr', 'e_self', 'e_ew_receip']


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name + '.npz'))


def golden_energies(g):
    return dict(zip(DUMP_ENERGIES, g['energies']))


def sort_rows(r):
    r = np.asarray(r, np.int32).reshape(-1, 4)
    if len(r) == 0:
        return r
    return r[np.lexsort((r[:, 3], r[:, 2], r[:, 1], r[:, 0]))]


def rows_sha(rows):
    return hashlib.sha256(np.ascontiguousarray(sort_rows(rows), np.int32).tobytes()).hexdigest()


def energy_errors(e, ref):
    """Relative error of every energy term against `ref`.  Each term is
    compared relative to its own magnitude; e_total and e_dir, which are
    cancellations of terms up to ~1e6 kcal/mol (e_self vs e_corr), are
    compared relative to the sum of the magnitudes of their terms (the
    conditioning of the sum)."""
    out = {}
    scale = sum(abs(ref[k]) for k in TERMS)
    for k in TERMS:
        # terms that are zero up to rounding (e.g. rigid-water bonds) are
        # compared on the scale of the whole energy
        out[k] = abs(e[k] - ref[k]) / max(abs(ref[k]), 1e-9 * scale)
    out['e_total'] = abs(e['e_total'] - ref['e_total']) / scale
    out['e_dir'] = abs(e['e_dir'] - ref['e_dir']) / max(abs(ref['e_el']) + abs(ref['e_lsym']), 1e-300)
    return out


def rms_rel(f, ref):
    return float(np.sqrt(((np.asarray(f) - ref) ** 2).sum() / (np.asarray(ref) ** 2).sum()))
