"""Differentiable MOIL energy: a torch.autograd.Function over the C ABI.

``moil_energy(engine, x, q, eps, rmin_half, lj_type)`` returns MOIL's
e_total for coordinates ``x`` (npt, 3), per-atom charges ``q`` and LJ
types (``eps``, ``rmin_half`` per type, ``lj_type`` per atom, -1 = no LJ);
its backward gives

* dE/dx   -- MOIL's gradient dpot from the forward evaluation,
* dE/dq   -- per atom,
* dE/deps, dE/drmin_half -- per LJ type,

the last three from the library's hand-written parameter-gradient kernels
(``moil_gpu_param_grad``), which recompute every pair distance on the
engine's current lists instead of storing anything from the forward
pass.  The per-atom LJ factors follow MOIL's geometric rule
(connect/fillpt.f:150-179): sigma = 2 rmin_half / 2^(1/6),
epsgm12 = 2 sqrt(eps) sigma^6, epsgm6 = 2 sqrt(eps) sigma^3; the type
gradients are the chain rule over the atoms of each type:
  dE/deps_t  = sum_i (e12_i dE/de12_i + e6_i dE/de6_i) / (2 eps_t),
  dE/drmh_t  = sum_i (6 e12_i dE/de12_i + 3 e6_i dE/de6_i) / rmh_t.
(Rmin = 2 rmin_half, so dE/dRmin = dE/drmh / 2.)

The lists are the engine's (built by ``engine.nbondm`` for the
conformation, exactly as MOIL's buffered lists); backward raises if the
engine's lists were rebuilt after the forward evaluation (the parameter
gradients would be those of another pair set).  Every call sets the
parameters on the engine (``moil_gpu_set_params``), so e_self, the
last-water parameters and the 1-4 pair parameters follow them; a call
whose parameters equal the engine's current ones (compared on their own
device) uploads nothing, and backward uploads nothing when the engine
still holds its forward's parameters.  Coordinates may be host (CPU) or
device (GPU) tensors; a parameter upload goes through host memory (the
library derives e_self and the 1-4 coefficients on the host), the
per-type reduction of the backward pass stays on the parameters' device.
"""
from __future__ import annotations

import numpy as np
import torch

from . import abi


def lj_factors(eps, rmin_half, lj_type):
    """Per-atom MOIL LJ factors (epsgm12, epsgm6) from per-type eps and
    Rmin/2 (float64 tensors, on their own device); atoms of type -1 get
    zeros."""
    t = torch.as_tensor(lj_type, dtype=torch.long, device=eps.device)
    has = t >= 0
    tt = torch.where(has, t, torch.zeros_like(t))
    sig = 2.0 * rmin_half / 2.0 ** (1.0 / 6.0)
    e12_t = 2.0 * torch.sqrt(eps) * sig ** 6
    e6_t = 2.0 * torch.sqrt(eps) * sig ** 3
    zero = torch.zeros((), dtype=eps.dtype, device=eps.device)
    return torch.where(has, e12_t[tt], zero), torch.where(has, e6_t[tt], zero)


def _upload(engine, q, e12, e6, scale14):
    """Set the parameters on the engine unless it holds these already."""
    last = getattr(engine, '_adj_params', None)
    if (last is not None and last[3] == scale14 and
            getattr(engine, 'param_token', None) == last[4] and
            all(a.shape == b.shape and torch.equal(a, b) for a, b in zip(last[:3], (q, e12, e6)))):
        return
    engine.set_params(q.cpu().numpy(), e6.cpu().numpy(), e12.cpu().numpy(), *scale14)
    engine._adj_params = (q.clone(), e12.clone(), e6.clone(), scale14, engine.param_token)


class _MoilEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, q, eps, rmin_half, engine, lj_type, v14, el14):
        epsd = eps.detach().to(torch.float64)
        rmhd = rmin_half.detach().to(torch.float64)
        e12, e6 = lj_factors(epsd, rmhd, lj_type)
        qd = q.detach().to(torch.float64)
        _upload(engine, qd, e12, e6, (v14, el14))
        xd = x.detach()
        if xd.is_cuda:
            xc = xd.to(torch.float64).contiguous()
            dpot = torch.empty_like(xc)
            en = torch.zeros(abi.N_ENERGIES, dtype=torch.float64, device=xc.device)
            sh = torch.cuda.current_stream(xc.device).cuda_stream
            engine.eforce_dev(xc.data_ptr(), dpot.data_ptr(), en.data_ptr(), sh)
            e_total = en[0].clone()
        else:
            energies, dp = engine.eforce(xd.numpy())
            dpot = torch.from_numpy(dp)
            e_total = torch.tensor(energies['e_total'], dtype=torch.float64)
            xc = xd
        ctx.engine = engine
        ctx.list_gen = engine.list_generation()
        ctx.lj_type = torch.as_tensor(lj_type, dtype=torch.long, device=eps.device)
        ctx.save_for_backward(xc, dpot, e12, e6, epsd, rmhd, qd)
        ctx.q_device, ctx.t_device = q.device, eps.device
        ctx.scale14 = (v14, el14)
        return e_total.to(x.device)

    @staticmethod
    def backward(ctx, g):
        xc, dpot, e12, e6, eps, rmh, qh = ctx.saved_tensors
        engine = ctx.engine
        if any(ctx.needs_input_grad[1:4]) and engine.list_generation() != ctx.list_gen:
            # the parameter gradients are computed on the engine's lists; the
            # forward's pair set (MOIL's buffered lists, built at earlier
            # coordinates) is gone
            raise RuntimeError('moil_energy: the engine rebuilt its pair lists after this '
                               'forward evaluation; back-propagate before the next nbondm, '
                               'or give each conformation its own engine')
        gx = gq = geps = grmh = None
        if ctx.needs_input_grad[0]:
            gx = (g * dpot).to(dpot.dtype)
        if any(ctx.needs_input_grad[1:4]):
            # the parameters of this forward evaluation (another call may
            # have replaced them on the engine since)
            _upload(engine, qh, e12, e6, ctx.scale14)
            if xc.is_cuda:
                out = torch.zeros((3, xc.shape[0]), dtype=torch.float64, device=xc.device)
                sh = torch.cuda.current_stream(xc.device).cuda_stream
                engine.param_grad_dev(xc.data_ptr(), out[0].data_ptr(), out[1].data_ptr(),
                                      out[2].data_ptr(), sh)
                dq, de12, de6 = out[0], out[1], out[2]
            else:
                dq, de12, de6 = (torch.from_numpy(a) for a in engine.param_grad(xc.numpy()))
            gd = g.detach().to(torch.float64)
            if ctx.needs_input_grad[1]:
                gq = (gd.to(dq.device) * dq).to(ctx.q_device)
            # per-type chain rule on the parameters' device
            dv = eps.device
            t = ctx.lj_type
            has = t >= 0
            nt = eps.shape[0]
            a12 = de12.to(dv) * e12
            a6 = de6.to(dv) * e6
            s1 = torch.zeros(nt, dtype=torch.float64, device=dv).index_add_(
                0, t[has], (a12 + a6)[has])
            s2 = torch.zeros(nt, dtype=torch.float64, device=dv).index_add_(
                0, t[has], (6.0 * a12 + 3.0 * a6)[has])
            gt = gd.to(dv)
            geps = (gt * s1 / (2.0 * eps)).to(ctx.t_device)
            grmh = (gt * s2 / rmh).to(ctx.t_device)
        return gx, gq, geps, grmh, None, None, None, None


def moil_energy(engine, x, q, eps, rmin_half, lj_type, v14=0.5, el14=0.5):
    """Differentiable e_total of the engine's system (see module doc)."""
    return _MoilEnergy.apply(x, q, eps, rmin_half, engine, lj_type, v14, el14)


def type_gradients(de12, de6, e12, e6, eps, rmin_half, lj_type):
    """dE/deps and dE/drmin_half per type from the per-atom gradients
    (numpy), the chain rule used by the autograd backward."""
    t = np.asarray(lj_type)
    has = t >= 0
    nt = len(eps)
    a12, a6 = de12 * e12, de6 * e6
    s1 = np.bincount(t[has], (a12 + a6)[has], minlength=nt)
    s2 = np.bincount(t[has], (6.0 * a12 + 3.0 * a6)[has], minlength=nt)
    return s1 / (2.0 * np.asarray(eps)), s2 / np.asarray(rmin_half)


__all__ = ['lj_factors', 'moil_energy', 'type_gradients']
