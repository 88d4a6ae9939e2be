"""Runs MOIL's own Fortran energy path (built by oracle/ref/Makefile into
oracle/_ref/moil_dump) on a synthetic system and parses its full-precision
dump.  Test infrastructure only (golden-fixture generation and the CPU
baseline); never imported by the product.
"""
from __future__ import annotations

import os
import re
import resource
import subprocess
import tempfile

import numpy as np

from moil_fnrfd_amd.systems import System, write_con

HERE = os.path.dirname(os.path.abspath(__file__))
REF_EXE = os.path.join(HERE, '_ref', 'moil_dump')
# the same MOIL driver and objects with MOIL's energy path replaced by the
# product's Fortran shim over libmoil_gpu.so (oracle/ref/Makefile, target gpu)
SHIM_EXE = os.path.join(HERE, '_ref', 'moil_dump_gpu')


def available(exe=REF_EXE) -> bool:
    return os.path.exists(exe)


def parse_dump(path: str) -> dict:
    out = {}
    with open(path, 'rb') as f:
        buf = f.read()
    p = 0
    while p < len(buf):
        name = buf[p:p + 8].decode().strip()
        typ, n = np.frombuffer(buf, np.int32, 2, p + 8)
        p += 16
        dt = np.float64 if typ == 1 else np.int32
        sz = 8 if typ == 1 else 4
        out[name] = np.frombuffer(buf, dt, int(n), p).copy()
        p += sz * int(n)
    return out


def run_moil(sysm: System, cut: dict, ewald: dict, *, squeeze=False, neval=1,
             nlist=1, sigma=0.0, seed=12345, dump_lists=True, workdir=None,
             timeout=3600, exe=REF_EXE, env=None, crd_file=None, mode=0) -> dict:
    """cut: rvmx relx rvbg rebg cutm (Angstrom); ewald: dtol, nfft (3), order.
    exe: the MOIL driver to run (REF_EXE, or SHIM_EXE for MOIL calling the
    GPU library); env: extra environment variables; crd_file: a coordinate
    file (path ending in .crd) for MOIL's getcrd instead of the system's
    raw coordinates; mode: a MOIL mode the driver switches on after
    reading the input (1 distance-dependent dielectric, 2 pressure, 3 a
    frozen particle: the modes the GPU shim refuses; 4 the arith flag on
    whatever parameters the file holds; 5 MOIL's pressure calculation)."""
    if not available(exe):
        raise RuntimeError('%s is not built (make -f oracle/ref/Makefile)' % exe)
    tmp = workdir or tempfile.mkdtemp(prefix='moilref_')
    os.makedirs(tmp, exist_ok=True)
    con = os.path.join(tmp, 'sys.con')
    crd = os.path.join(tmp, 'sys.bin')
    par = os.path.join(tmp, 'sys.par')
    dmp = os.path.join(tmp, 'sys.dump')
    write_con(con, sysm.topo)
    if crd_file:
        crd = crd_file
    else:
        np.ascontiguousarray(sysm.coor, np.float64).tofile(crd)
    a, b, c = sysm.box
    n1, n2, n3 = ewald['nfft']
    with open(par, 'w') as f:
        f.write(con + '\n' + crd + '\n')
        f.write('%.17g %.17g %.17g\n' % (a, b, c))
        f.write('%.17g %.17g %.17g %.17g %.17g\n' % (cut['rvmx'], cut['relx'], cut['rvbg'],
                                                    cut['rebg'], cut['cutm']))
        f.write('%.17g %d %d %d %d\n' % (ewald['dtol'], n1, n2, n3, ewald.get('order', 4)))
        f.write('%.17g\n' % cut.get('eps', 1.0))
        f.write('%d %d %d %.17g %d %d\n' % (int(squeeze), neval, nlist, sigma, seed,
                                           int(dump_lists)))
        f.write(dmp + '\n')
        # MOIL's hydrogen rule (pot/nbond.f:115), on unless cut['hvdw0'] is False
        f.write('%d\n' % int(cut.get('hvdw0', True)))
        f.write('%d\n' % int(mode))
    def _big_stack():
        # MOIL keeps large work arrays on the stack (e.g. symm/nbmsym.f:36-46)
        resource.setrlimit(resource.RLIMIT_STACK,
                           (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
    full_env = dict(os.environ)
    full_env.update(env or {})
    r = subprocess.run([exe, par], cwd=tmp, capture_output=True, text=True,
                       timeout=timeout, preexec_fn=_big_stack, env=full_env)
    if r.returncode != 0 or not os.path.exists(dmp):
        raise RuntimeError('moil_dump failed (%d):\n%s\n%s' % (r.returncode, r.stdout[-3000:],
                                                                r.stderr[-3000:]))
    d = parse_dump(dmp)
    d['stdout'] = r.stdout
    for line in r.stdout.splitlines():
        if line.startswith('MOILTIME'):
            kv = dict(re.findall(r'(\w+)=\s*([-+0-9.Ee]+)', line))
            d['tlist_s'] = float(kv['tlist_s'])
            d['tforce_s'] = float(kv['tforce_s'])
    return d


SDS_EXE = os.path.join(HERE, '_ref', 'moil_sds')
# MOIL's unmodified chain functional over the product's Fortran shim
SDS_SHIM_EXE = os.path.join(HERE, '_ref', 'moil_sds_gpu')
# the same driver with sds itself replaced by the chain shim
# (moil_fnrfd_amd/fortran/moil_gpu_shim_chain.f -> moil_gpu_chain_sds)
SDS_CHAIN_EXE = os.path.join(HERE, '_ref', 'moil_sds_gpuchain')


def run_moil_sds(sysm: System, cut: dict, ewald: dict, r, ipick, gamma, rho, lam, *,
                 workdir=None, timeout=3600, exe=SDS_EXE, env=None) -> dict:
    """MOIL's chain functional (chain/sds.f) over replicas r (igrid, npt, 3)
    of sysm's topology, via oracle/_ref/moil_sds (or SDS_SHIM_EXE).
    Returns the parsed dump: SENER, DV, D0, E0, E1 as sds returns them,
    plus DPOTS/ETOTS (each interior replica's own gradient and energy)
    and BETA."""
    if not available(exe):
        raise RuntimeError('%s is not built (make -f oracle/ref/Makefile sds)' % exe)
    r = np.ascontiguousarray(r, np.float64)
    igrid = r.shape[0]
    tmp = workdir or tempfile.mkdtemp(prefix='moilsds_')
    os.makedirs(tmp, exist_ok=True)
    con = os.path.join(tmp, 'sys.con')
    rf = os.path.join(tmp, 'chain.bin')
    pk = os.path.join(tmp, 'pick.bin')
    par = os.path.join(tmp, 'sds.par')
    dmp = os.path.join(tmp, 'sds.dump')
    write_con(con, sysm.topo)
    r.tofile(rf)
    np.ascontiguousarray(ipick, np.int32).tofile(pk)
    a, b, c = sysm.box
    n1, n2, n3 = ewald['nfft']
    with open(par, 'w') as f:
        f.write(con + '\n' + rf + '\n' + pk + '\n')
        f.write('%.17g %.17g %.17g\n' % (a, b, c))
        f.write('%.17g %.17g %.17g %.17g %.17g\n' % (cut['rvmx'], cut['relx'], cut['rvbg'],
                                                    cut['rebg'], cut['cutm']))
        f.write('%.17g %d %d %d %d\n' % (ewald['dtol'], n1, n2, n3, ewald.get('order', 4)))
        f.write('%.17g\n' % cut.get('eps', 1.0))
        f.write('%d %.17g %.17g %.17g\n' % (igrid, gamma, rho, lam))
        f.write(dmp + '\n')
        f.write('%d\n' % int(cut.get('hvdw0', True)))

    def _big_stack():
        resource.setrlimit(resource.RLIMIT_STACK,
                           (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
    full_env = dict(os.environ)
    full_env.update(env or {})
    p = subprocess.run([exe, par], cwd=tmp, capture_output=True, text=True, timeout=timeout,
                       preexec_fn=_big_stack, env=full_env)
    if p.returncode != 0 or not os.path.exists(dmp):
        raise RuntimeError('moil_sds failed (%d):\n%s\n%s' % (p.returncode, p.stdout[-3000:],
                                                               p.stderr[-3000:]))
    d = parse_dump(dmp)
    d['stdout'] = p.stdout
    for line in p.stdout.splitlines():
        if line.startswith('MOILTIME'):
            d['sds_s'] = float(line.split('=')[1])
    return d


DYNA_EXE = os.path.join(HERE, '_ref', 'moil_dyna')


def run_moil_dyna(sysm: System, cut: dict, ewald: dict, velo, nstep, nlist, dt_ps, *,
                  mshake=True, tolcons=1e-7, workdir=None, timeout=3600) -> dict:
    """MOIL's dynamics step loop (dynamics/dyna.f:105-301 on MOIL's eforce and
    matrix SHAKE) via oracle/_ref/moil_dyna from sysm's coordinates and the
    velocities velo (npt, 3) in A per MOIL time unit, for nstep steps of
    dt_ps picoseconds with the lists updated every nlist steps.  Returns
    the parsed dump: COOR, VELO, DPOT (final), ETOT (e_total after each
    step) and MSHAKE (cmat(3,3) column-major, tolcons, reqoh2, reqhh2,
    invmh, invmo as mshakinit set them)."""
    if not available(DYNA_EXE):
        raise RuntimeError('%s is not built (make -f oracle/ref/Makefile dyna)' % DYNA_EXE)
    tmp = workdir or tempfile.mkdtemp(prefix='moildyn_')
    os.makedirs(tmp, exist_ok=True)
    con, crd, vel = (os.path.join(tmp, n) for n in ('sys.con', 'sys.bin', 'vel.bin'))
    par, dmp = os.path.join(tmp, 'dyna.par'), os.path.join(tmp, 'dyna.dump')
    write_con(con, sysm.topo)
    np.ascontiguousarray(sysm.coor, np.float64).tofile(crd)
    np.ascontiguousarray(velo, np.float64).tofile(vel)
    a, b, c = sysm.box
    n1, n2, n3 = ewald['nfft']
    with open(par, 'w') as f:
        f.write(con + '\n' + crd + '\n')
        f.write('%.17g %.17g %.17g\n' % (a, b, c))
        f.write('%.17g %.17g %.17g %.17g %.17g\n' % (cut['rvmx'], cut['relx'], cut['rvbg'],
                                                    cut['rebg'], cut['cutm']))
        f.write('%.17g %d %d %d %d\n' % (ewald['dtol'], n1, n2, n3, ewald.get('order', 4)))
        f.write('%.17g\n' % cut.get('eps', 1.0))
        f.write('%d %d %.17g %d %.17g\n' % (nstep, nlist, dt_ps, int(mshake), tolcons))
        f.write(vel + '\n' + dmp + '\n')

    def _big_stack():
        resource.setrlimit(resource.RLIMIT_STACK,
                           (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
    r = subprocess.run([DYNA_EXE, par], cwd=tmp, capture_output=True, text=True,
                       timeout=timeout, preexec_fn=_big_stack)
    if r.returncode != 0 or not os.path.exists(dmp):
        raise RuntimeError('moil_dyna failed (%d):\n%s\n%s' % (r.returncode, r.stdout[-3000:],
                                                                r.stderr[-3000:]))
    d = parse_dump(dmp)
    d['stdout'] = r.stdout
    return d
