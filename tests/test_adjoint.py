"""Parameter gradients (moil_gpu_param_grad) and the differentiable binding
(moil_fnrfd_amd.adjoint) on the 10,048-atom heteropolymer-in-water system
(200 chains, 46.4 A box, 48^3 grid).

* Euler identities, exact for this energy (homogeneous of degree 2 in the
  charges, and in the LJ factors): sum_i q_i dE/dq_i = 2 E_coulomb and
  sum_i (e12_i dE/de12_i + e6_i dE/de6_i) = 2 E_LJ, with E_coulomb =
  e_el + e_lsym + e_corr + e_self + e_ew_receip + e_el14 and E_LJ = e_vdw
  + e_vsym + e_vdw14 -- 1e-10 of the summed term magnitudes in FP64.
* Central finite differences of the FP64 energy on fixed lists
  (h = 1e-4): charges of 16 seeded chain atoms, the water O and H charges
  (tied: every water), eps and Rmin/2 of every LJ type -- 1e-6 relative
  (floor 1e-4 absolute); 8 atoms' coordinates against dpot with h = 1e-6
  (MOIL's cutoffs are unshifted, so larger steps can carry a pair across
  one).
* Mixed precision (FP32 pairs) against FP64: RMS-relative <= 1e-5.
* Autograd: torch.autograd.grad through moil_energy, with device
  coordinates, equals dpot, dq and the per-type chain rule.
"""
import sys

import numpy as np
import pytest

from moil_fnrfd_amd import abi, configs
from moil_fnrfd_amd.adjoint import lj_factors, moil_energy, type_gradients

pytestmark = pytest.mark.gpu

COUL = ['e_el', 'e_lsym', 'e_corr', 'e_self', 'e_ew_receip', 'e_el14']
LJ = ['e_vdw', 'e_vsym', 'e_vdw14']


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope='module')
def system():
    import torch
    s, cut, ew = configs.hetero_10k()
    m = s.meta
    e12, e6 = lj_factors(torch.tensor(m['lj_eps']), torch.tensor(m['lj_rmin_half']),
                         m['lj_type'])
    return s, cut, ew, e12.numpy(), e6.numpy()


def _engine(system, prec):
    from moil_fnrfd_amd.eforce import MoilGPU
    s, cut, ew, e12, e6 = system
    g = MoilGPU(s.topo, configs.params_for(s, cut, ew, precision=prec))
    x = g.squeeze(s.coor)
    g.nbondm(x)
    g.set_params(s.topo.ptchg, e6, e12, 0.5, 0.5)
    return g, x


def test_euler_identities(system):
    s, cut, ew, e12, e6 = system
    for prec, tol in ((abi.PREC_FP64, 1e-10), (abi.PREC_MIXED, 1e-6)):
        _say('case adjoint euler prec=%d' % prec)
        g, x = _engine(system, prec)
        e, _ = g.eforce(x)
        dq, de12, de6 = g.param_grad(x)
        g.close()
        q = s.topo.ptchg
        ec = sum(e[k] for k in COUL)
        elj = sum(e[k] for k in LJ)
        scale_c = sum(abs(e[k]) for k in COUL)
        scale_l = sum(abs(e[k]) for k in LJ)
        rc = abs(float(q @ dq) - 2 * ec) / scale_c
        rl = abs(float(e12 @ de12 + e6 @ de6) - 2 * elj) / scale_l
        _say('  sum q dE/dq vs 2 E_coul: %.2e   LJ: %.2e' % (rc, rl))
        assert rc < tol and rl < tol


def test_finite_differences_fp64(system):
    import torch
    s, cut, ew, e12, e6 = system
    m = s.meta
    g, x = _engine(system, abi.PREC_FP64)
    e0, dpot = g.eforce(x)
    dq, de12, de6 = g.param_grad(x)
    q0 = s.topo.ptchg.copy()
    h = 1e-4

    def energy_q(q):
        g.set_params(q, e6, e12, 0.5, 0.5)
        return g.eforce(x)[0]['e_total']

    def close(a, b):
        return abs(a - b) <= 1e-6 * max(abs(a), abs(b)) + 1e-4

    rng = np.random.default_rng(11)
    chain = np.arange(len(q0) - 3 * m['n_water'])  # chain atoms come first
    worst = 0.0
    for i in rng.choice(chain, 16, replace=False):
        fd = (energy_q(q0 + h * (np.arange(len(q0)) == i)) -
              energy_q(q0 - h * (np.arange(len(q0)) == i))) / (2 * h)
        worst = max(worst, abs(fd - dq[i]) / max(abs(fd), 1.0))
        assert close(fd, dq[i]), (i, fd, dq[i])
    # tied water charges (every water O, every water H)
    nw = m['n_water']
    first = len(q0) - 3 * nw
    for site, sel in (('O', np.arange(first, len(q0), 3)),
                      ('H', np.concatenate([np.arange(first + 1, len(q0), 3),
                                            np.arange(first + 2, len(q0), 3)]))):
        mask = np.zeros(len(q0))
        mask[sel] = 1.0
        fd = (energy_q(q0 + h * mask) - energy_q(q0 - h * mask)) / (2 * h)
        ad = float(dq[sel].sum())
        _say('  tied water %s charge: adjoint %.10g  fd %.10g' % (site, ad, fd))
        worst = max(worst, abs(fd - ad) / max(abs(fd), 1.0))
        assert close(fd, ad)
    g.set_params(q0, e6, e12, 0.5, 0.5)
    # LJ types
    eps0 = torch.tensor(m['lj_eps'])
    rmh0 = torch.tensor(m['lj_rmin_half'])
    geps, grmh = type_gradients(de12, de6, e12, e6, m['lj_eps'], m['lj_rmin_half'],
                                m['lj_type'])

    def energy_t(eps, rmh):
        a12, a6 = lj_factors(eps, rmh, m['lj_type'])
        g.set_params(q0, a6.numpy(), a12.numpy(), 0.5, 0.5)
        return g.eforce(x)[0]['e_total']

    for t in range(len(eps0)):
        d = torch.zeros_like(eps0)
        d[t] = h * eps0[t]
        fd_e = (energy_t(eps0 + d, rmh0) - energy_t(eps0 - d, rmh0)) / (2 * float(d[t]))
        d = torch.zeros_like(rmh0)
        d[t] = h * rmh0[t]
        fd_r = (energy_t(eps0, rmh0 + d) - energy_t(eps0, rmh0 - d)) / (2 * float(d[t]))
        worst = max(worst, abs(fd_e - geps[t]) / max(abs(fd_e), 1.0),
                    abs(fd_r - grmh[t]) / max(abs(fd_r), 1.0))
        assert close(fd_e, geps[t]), (t, fd_e, geps[t])
        assert close(fd_r, grmh[t]), (t, fd_r, grmh[t])
    g.set_params(q0, e6, e12, 0.5, 0.5)
    # coordinates against dpot; MOIL's cutoffs are unshifted (the energy
    # steps where a listed pair crosses rvmx / relx), so a small step keeps
    # the pairs on their side
    hx = 1e-6
    for a in rng.choice(len(q0), 8, replace=False):
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, k] += hx
            xm[a, k] -= hx
            fd = (g.eforce(xp)[0]['e_total'] - g.eforce(xm)[0]['e_total']) / (2 * hx)
            assert abs(fd - dpot[a, k]) <= 1e-5 * max(abs(fd), 1.0) + 1e-3, (a, k, fd, dpot[a, k])
    _say('  worst relative adjoint-vs-FD error %.2e' % worst)
    g.close()


def test_mixed_vs_fp64(system):
    out = {}
    for prec in (abi.PREC_FP64, abi.PREC_MIXED):
        g, x = _engine(system, prec)
        out[prec] = g.param_grad(x)
        g.close()
    for name, a, b in zip(('dq', 'de12', 'de6'), out[abi.PREC_MIXED], out[abi.PREC_FP64]):
        rr = float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))
        mr = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        _say('  mixed vs FP64 %s: rms-rel %.2e  max-rel %.2e' % (name, rr, mr))
        assert rr < 1e-5


def test_autograd_device(system):
    import torch
    s, cut, ew, e12, e6 = system
    m = s.meta
    g, x = _engine(system, abi.PREC_FP64)
    _, dpot = g.eforce(x)
    dq, de12, de6 = g.param_grad(x)
    geps, grmh = type_gradients(de12, de6, e12, e6, m['lj_eps'], m['lj_rmin_half'],
                                m['lj_type'])
    xt = torch.tensor(x, device='cuda:0', requires_grad=True)
    qt = torch.tensor(s.topo.ptchg, requires_grad=True)
    et = torch.tensor(m['lj_eps'], requires_grad=True)
    rt = torch.tensor(m['lj_rmin_half'], requires_grad=True)
    E = moil_energy(g, xt, qt, et, rt, m['lj_type'])
    gx, gq, ge, gr = torch.autograd.grad(E, [xt, qt, et, rt])
    assert np.allclose(gx.cpu().numpy(), dpot, rtol=0, atol=1e-12)
    assert np.allclose(gq.numpy(), dq, rtol=1e-12, atol=1e-12)
    assert np.allclose(ge.numpy(), geps, rtol=1e-12, atol=1e-12)
    assert np.allclose(gr.numpy(), grmh, rtol=1e-12, atol=1e-12)
    g.close()


def test_autograd_parameters_on_device_and_interleaved_calls(system):
    """Parameters as device tensors (per-type reduction on the device), a
    second forward with other charges on the same engine before the first
    backward: the first backward re-uploads its own parameters and gives
    the gradients of a lone call; a repeated forward with unchanged
    parameters uploads nothing and gives the same energy."""
    import torch
    s, cut, ew, e12, e6 = system
    m = s.meta
    g, x = _engine(system, abi.PREC_FP64)
    dev = torch.device('cuda:0')

    def call(q):
        xt = torch.tensor(x, device=dev, requires_grad=True)
        qt = torch.tensor(q, device=dev, requires_grad=True)
        et = torch.tensor(m['lj_eps'], device=dev, requires_grad=True)
        rt = torch.tensor(m['lj_rmin_half'], device=dev, requires_grad=True)
        return moil_energy(g, xt, qt, et, rt, m['lj_type']), (xt, qt, et, rt)

    q0 = np.array(s.topo.ptchg, np.float64)
    E0, a0 = call(q0)
    ref = [t.cpu().numpy() for t in torch.autograd.grad(E0, a0)]
    tok = g.param_token
    E0b, _ = call(q0)
    assert g.param_token == tok  # unchanged parameters: no upload
    assert float(E0b.detach()) == float(E0.detach())
    E1st, a1st = call(q0)
    q1 = q0 * 0.97  # still neutral
    E2, a2 = call(q1)
    assert float(E2.detach()) != float(E1st.detach())
    got = [t.cpu().numpy() for t in torch.autograd.grad(E1st, a1st)]
    for a, b in zip(got, ref):
        assert np.allclose(a, b, rtol=1e-12, atol=1e-12)
    g.close()
