"""Resident molecular dynamics (moil_gpu_dyna_*) against MOIL's own dynamics
loop: oracle/_ref/moil_dyna runs dynamics/dyna.f:105-301's step sequence on
MOIL's eforce, mshakpt and mshakvl (matrix SHAKE of the TIP3 waters, with
MOIL's mshakinit constants), from the same coordinates and velocities.

* FP64 mode, 40 steps of the 216-water box, lists every 10 steps: every
  step's e_total, the final coordinates and velocities agree with MOIL's.
  The force evaluations differ from MOIL's only in their summation order
  (~1e-13 relative), so the trajectories separate only as fast as that
  difference grows;
* the constraints hold: every water's O-H and H-H distances and the bond
  components of its velocities meet MOIL's tolerance after every run;
* mixed precision conserves the total energy as well as FP64 does.
"""
import numpy as np
import pytest

from moil_fnrfd_amd import abi, configs
from oracle import ref_runner

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not ref_runner.available(ref_runner.DYNA_EXE),
                                 reason='oracle/_ref/moil_dyna not built')]


def _setup(nstep=40, nlist=10, dt=0.001, seed=31):
    s, cut, ew = configs.water216()
    v0 = configs.maxwell_velocities(s.topo.ptms, 300.0, seed)
    d = ref_runner.run_moil_dyna(s, cut, ew, v0, nstep, nlist, dt, mshake=True, tolcons=1e-7,
                                 timeout=300)
    msh = d['MSHAKE']
    consts = (msh[:9].reshape(3, 3, order='F'), msh[9], msh[10], msh[11], msh[12], msh[13])
    return s, cut, ew, v0, d, consts


def _kinetic(m, v):
    return 0.5 * float((m[:, None] * v * v).sum())


def _constraint_errors(x, v, consts):
    _, tol, roh2, rhh2, _, _ = consts
    o = np.arange(0, len(x), 3)
    r1, r2, rh = x[o] - x[o + 1], x[o] - x[o + 2], x[o + 1] - x[o + 2]
    dist = np.abs(np.concatenate([(r1 * r1).sum(1) - roh2, (r2 * r2).sum(1) - roh2,
                                  (rh * rh).sum(1) - rhh2]))
    vel = np.concatenate([((v[o] - v[o + 1]) * r1).sum(1), ((v[o] - v[o + 2]) * r2).sum(1),
                          ((v[o + 1] - v[o + 2]) * rh).sum(1)])
    return dist.max(), vel.max(), tol


def test_fp64_dynamics_matches_moil():
    from moil_fnrfd_amd.eforce import MoilDyna, MoilGPU
    print('case dyna water216 FP64 40 steps', flush=True)
    nstep = 40
    s, cut, ew, v0, d, consts = _setup(nstep)
    g = MoilGPU(s.topo, configs.params_for(s, cut, ew, precision=abi.PREC_FP64))
    dyn = MoilDyna(g, s.topo.ptms, 0.001, 10, consts)
    x, v, dp, en = dyn.run(s.coor, v0, nstep, energies=True)
    et = en[:, 0]
    ref_et = d['ETOT']
    ref_x, ref_v = d['COOR'].reshape(-1, 3), d['VELO'].reshape(-1, 3)
    scale = np.abs(ref_et).max()
    e_err = np.abs(et - ref_et).max() / scale
    x_err = np.abs(x - ref_x).max()
    v_err = np.sqrt(((v - ref_v) ** 2).sum() / (ref_v ** 2).sum())
    print('  e_total per step max %.2e of max|e_total|; coordinates max |dx| %.2e A; '
          'velocities rms-rel %.2e' % (e_err, x_err, v_err), flush=True)
    assert e_err < 1e-10
    assert x_err < 1e-9
    assert v_err < 1e-9
    dist, vel, tol = _constraint_errors(x, v, consts)
    assert dist <= 1.01 * tol and vel <= 1.01 * tol, (dist, vel, tol)
    dyn.close()
    g.close()


def test_mixed_dynamics_conserves_energy():
    """200 steps of 1 fs from the same start in both precisions: the total
    energy (e_total + kinetic) drifts no more in mixed precision than a
    small multiple of the FP64 run's drift, and the constraints hold."""
    from moil_fnrfd_amd.eforce import MoilDyna, MoilGPU
    print('case dyna water216 mixed vs FP64 energy conservation', flush=True)
    s, cut, ew, v0, d, consts = _setup(nstep=1)
    m = np.asarray(s.topo.ptms, np.float64)
    drift = {}
    for prec in (abi.PREC_FP64, abi.PREC_MIXED):
        g = MoilGPU(s.topo, configs.params_for(s, cut, ew, precision=prec))
        dyn = MoilDyna(g, s.topo.ptms, 0.001, 10, consts)
        x, v = s.coor, v0
        tot = []
        for _ in range(10):
            x, v, dp, en = dyn.run(x, v, 20, energies=True)
            tot.append(en[-1, 0] + _kinetic(m, v))
        dist, vel, tol = _constraint_errors(x, v, consts)
        assert dist <= 1.01 * tol and vel <= 1.01 * tol, (prec, dist, vel, tol)
        drift[prec] = np.abs(np.asarray(tot) - tot[0]).max()
        print('  precision %d: total energy drift %.4f kcal/mol over 200 steps (e_total %.2f)'
              % (prec, drift[prec], tot[-1]), flush=True)
        dyn.close()
        g.close()
    assert drift[abi.PREC_MIXED] < 2.0 * drift[abi.PREC_FP64] + 0.05
