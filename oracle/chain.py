"""CPU restatement of MOIL's chain-of-replicas functional (chain/sds.f).

Test infrastructure only: imported by tests/ as the checker of the GPU
chain entry point, never by the product.

MOIL evaluates a chain of igrid replicas r_1..r_igrid of one system
(chain/sds.f:1-16):

    S = sum_{j=2}^{igrid-1} V(r_j)                      interior replicas
      + gamma * sum_{j=1}^{igrid-1} (d_j - <d>)^2        nearest springs
      + rho/lambda * sum_{j=1}^{igrid-2} exp(-lambda D_j / <d>^2)

with d_j = |r_{j+1} - r_j| and D_j = |r_{j+2} - r_j|^2 over the selected
atoms (ipick > 0), <d> the mean of the d_j.  The gradient dv gets each
interior replica's own gradient plus the chain terms, <d>'s dependence on
the coordinates included (chain/sds.f:195-244).  The end replicas get
chain terms only.

chain_terms() follows sds.f's arithmetic step by step, sums included
(sequential over atoms, x then y then z; np.cumsum is a strictly
sequential sum), so that with MOIL's own interior energies and gradients
it reproduces MOIL's outputs exactly.

Parity pins (tests/test_chain_oracle.py): MOIL's own sds run
(oracle/_ref/moil_sds, fixture tests/golden/chain_water216.npz), central
finite differences of S in FP64, and translation invariance (the chain
terms' gradient sums to zero over all replicas).
"""
from __future__ import annotations

import math

import numpy as np


def _seq_sum(terms: np.ndarray) -> float:
    """Strictly left-to-right FP64 sum (Fortran's d = d + t loop)."""
    if terms.size == 0:
        return 0.0
    return float(np.cumsum(terms, dtype=np.float64)[-1])


def chain_terms(r, ipick, gamma, rho, lam, e_int, g_int):
    """MOIL's sds (chain/sds.f:46-257) with the interior replicas' energies
    and gradients given.

    r      (igrid, npt, 3) float64 replica coordinates
    ipick  (npt,) selection, > 0 selected (chain/chmin.f:389-407)
    e_int  (igrid,) e_total of each replica (entries 2..igrid-1 used)
    g_int  (igrid, npt, 3) dpot of each replica (interior used)
    Returns dict(sener, dv (igrid, npt, 3), d0 (2 igrid - 3), e0 (igrid),
    e1 (2 igrid - 3)) with MOIL's meanings: d0 holds the nearest distances
    d_j, then the squared next-nearest distances D_j; e1 the spring
    factors 2 gamma (1 - <d>/d_j), then exp(-lambda D_j/<d>^2)/<d>^2.
    """
    r = np.asarray(r, np.float64)
    igrid, npt, _ = r.shape
    sel = np.asarray(ipick) > 0
    gam2 = 2.0 * gamma
    rho2 = 2.0 * rho
    # sds.f:53-61, 71-110: interior energies summed in replica order, their
    # gradients copied, the end replicas zero
    sener = 0.0
    e0 = np.zeros(igrid)
    dv = np.zeros_like(r)
    for j in range(1, igrid - 1):
        e0[j] = e_int[j]
        sener = sener + float(e_int[j])
        dv[j] = g_int[j]
    nd = 2 * igrid - 3
    d0 = np.zeros(nd)

    def sqsum(a, b):
        t = (r[a] - r[b])[sel]
        return _seq_sum((t * t).reshape(-1))

    # sds.f:115-129 nearest, 134-148 next-nearest squared distances
    for j in range(igrid - 1):
        d0[j] = sqsum(j + 1, j)
    for j in range(igrid - 2):
        d0[j + igrid - 1] = sqsum(j + 2, j)
    # sds.f:149-154
    dave = 0.0
    for j in range(igrid - 1):
        d0[j] = math.sqrt(d0[j])
        dave = dave + d0[j]
    dave = dave / (igrid - 1)
    dave2 = 1.0 / (dave * dave)
    # sds.f:169-174 springs
    tmp = 0.0
    for j in range(igrid - 1):
        t1 = d0[j] - dave
        tmp = tmp + t1 * t1
    sener = sener + gamma * tmp
    # sds.f:180-188 repulsion; exponent over the squared distance
    e1 = np.zeros(nd)
    tmp = 0.0
    d2w = 0.0
    for j in range(igrid - 1, nd):
        e1[j] = math.exp(-lam * d0[j] / (dave * dave))
        tmp = tmp + e1[j]
        d2w = d2w + e1[j] * d0[j]
        e1[j] = e1[j] * dave2
    sener = sener + rho / lam * tmp
    d2w = rho2 * dave2 / ((igrid - 1) * dave) * d2w
    # sds.f:193-195
    for j in range(igrid - 1):
        e1[j] = gam2 * (1.0 - dave / d0[j])
    # sds.f:199-214 spring gradient
    for j in range(igrid - 1):
        t = e1[j] * (r[j] - r[j + 1])
        dv[j][sel] = dv[j][sel] + t[sel]
        dv[j + 1][sel] = dv[j + 1][sel] - t[sel]
    # sds.f:218-235 repulsion gradient
    for j in range(igrid - 2):
        t1 = rho2 * e1[j + igrid - 1]
        t = (r[j] - r[j + 2]) * t1
        dv[j][sel] = dv[j][sel] - t[sel]
        dv[j + 2][sel] = dv[j + 2][sel] + t[sel]
    # sds.f:236-250 the mean distance's dependence on the coordinates
    for j in range(igrid - 1):
        t1 = d2w / d0[j]
        t = (r[j] - r[j + 1]) * t1
        dv[j][sel] = dv[j][sel] + t[sel]
        dv[j + 1][sel] = dv[j + 1][sel] - t[sel]
    return dict(sener=sener, dv=dv, d0=d0, e0=e0, e1=e1)


def chain_energy_only(r, ipick, gamma, rho, lam):
    """The chain terms of S alone (no replica energies), for finite
    differences: gamma sum (d_j - <d>)^2 + rho/lambda sum exp(-lambda
    D_j/<d>^2)."""
    r = np.asarray(r, np.float64)
    igrid = r.shape[0]
    sel = np.asarray(ipick) > 0
    d = np.array([math.sqrt(float(((r[j + 1] - r[j])[sel] ** 2).sum()))
                  for j in range(igrid - 1)])
    D = np.array([float(((r[j + 2] - r[j])[sel] ** 2).sum()) for j in range(igrid - 2)])
    dave = d.mean()
    return gamma * float(((d - dave) ** 2).sum()) + rho / lam * float(
        np.exp(-lam * D / dave ** 2).sum())
