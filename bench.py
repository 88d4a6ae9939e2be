"""Benchmark: MOIL energy/force evaluations per second on the 98,304-atom
PME water box (32,768 rigid 3-site waters, 68.9 A box, 10 A cutoffs with a
1.5 A list skin, 72^3 grid, beta 0.31), FP32 pair arithmetic with 64-bit
fixed-point forces and FP64 energies, on one MI355X.

One step = one eforce call on device-resident coordinates; between steps
the coordinates get N(0, 0.01 A) noise (cumulative, drawn on the device),
and every 20 steps the list pipeline (squeeze + nbondm + syminit) is
rebuilt on the GPU -- both inside the timed region.  An evaluation whose
forces left the fixed-point range (NaN energies) stops the benchmark.  Then the batched leg: 96 systems evaluated by one
batched call against one call per system (speedup and force error).
Prints one JSON line.

  python bench.py [--gpus 1] [--steps 200] [--warmup 20] [--no-batched]

The noise accumulates without dynamics to hold the molecules apart, so in
the 3x-liquid-density box some atoms drift into contact: after ~1000
steps a force term passes the mixed mode's fixed-point range (2^39
kcal/mol/A) and the evaluation reports the overflow (NaN energies); 200
steps stay well inside it.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NLIST = 20
SIGMA = 0.01


def algorithmic_bytes_realspace(n_water_pairs, npt):
    """Real-space gather model: 4 B per list entry, a 16 B float4 per
    gathered j-site (3 per water entry), 16 B per atom read and a 24 B
    int64x3 force write per atom."""
    E = n_water_pairs
    Sj = 3 * n_water_pairs
    return 4 * E + 16 * Sj + 16 * npt + 24 * npt


# the dominant kernel of the metric's loop: the mixed-precision water pair
# kernel on its short grid (4 workgroups per CU walking the 2,044 chunks of
# this workload in rounds)
RS_KERNEL = 'k_water_rounds'


def measured_traffic():
    """HBM bytes per launch of the water pair kernel from the committed
    counter summary (FETCH_SIZE + WRITE_SIZE, each scaled by a plain copy of
    known size); None if the summary is absent."""
    path = os.path.join(ROOT, 'profiles', '%s_water32k_counters.md' % RS_KERNEL)
    try:
        import re
        m = re.search(r'traffic_bytes_per_launch: (\d+)', open(path).read())
        return int(m.group(1)) if m else None
    except OSError:
        return None


def cpu_baseline(sysm, cut, ew):
    """MOIL's own Fortran (oracle/_ref) on the host cores: one list build
    plus a bounded number of eforce calls; rate with the list build
    amortised over NLIST evaluations as in the GPU metric."""
    try:
        from oracle import ref_runner
        if ref_runner.available():
            nev = 6
            d = ref_runner.run_moil(sysm, cut, ew, squeeze=True, neval=nev, nlist=1000,
                                    sigma=SIGMA, dump_lists=False, timeout=900)
            t_list = d['tlist_s']
            t_eval = d['tforce_s'] / nev
            rate = 1.0 / (t_eval + t_list / NLIST)
            out = dict(value=rate, unit='evals/s', cores=1, kind='reference',
                       sample='MOIL Fortran (flang -O2), 98,304-atom water box: 1 list build '
                              '(%.2f s) + %d eforce calls (%.2f s each), list cost amortised '
                              'over %d evals' % (t_list, nev, t_eval, NLIST))
            out['aggregate'] = cpu_aggregate(sysm, cut, ew)
            return out
    except Exception as ex:  # fall back to the C restatement
        print('reference CPU baseline unavailable: %s' % ex, file=sys.stderr)
    from moil_fnrfd_amd import configs
    from oracle.oracle import Oracle
    o = Oracle(sysm.topo, configs.params_for(sysm, cut, ew))
    t0 = time.time()
    o.build_lists(sysm.coor)
    t1 = time.time()
    nev = 4
    for _ in range(nev):
        o.eforce(sysm.coor)
    t2 = time.time()
    t_eval = (t2 - t1) / nev
    rate = 1.0 / (t_eval + (t1 - t0) / NLIST)
    return dict(value=rate, unit='evals/s', cores=1, kind='port',
                sample='C restatement (gcc -O2), 98,304-atom water box: 1 list build + %d evals'
                       % nev)


def cpu_aggregate(sysm, cut, ew, nproc=8, nev=3):
    """MOIL on nproc host cores at once: nproc independent MOIL processes
    on the same workload side by side (MOIL's serial build has no threads),
    each one list build plus nev eforce calls; the aggregate rate is the sum
    of the processes' own rates (list cost amortised as above)."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import ref_runner

    def one(k):
        d = ref_runner.run_moil(sysm, cut, ew, squeeze=True, neval=nev, nlist=1000,
                                sigma=SIGMA, seed=12345 + k, dump_lists=False, timeout=900)
        return 1.0 / (d['tforce_s'] / nev + d['tlist_s'] / NLIST)

    try:
        t0 = time.perf_counter()
        with ThreadPoolExecutor(nproc) as ex:
            rates = list(ex.map(one, range(nproc)))
        wall = time.perf_counter() - t0
    except Exception as e:
        print('CPU aggregate unavailable: %s' % e, file=sys.stderr)
        return None
    return dict(value=float(sum(rates)), unit='evals/s', cores=nproc, processes=nproc,
                per_process=[float(r) for r in rates], wall_s=wall,
                sample='%d MOIL processes side by side, each 1 list build + %d eforce calls '
                       'on the 98,304-atom water box' % (nproc, nev))


def copy_probe(mib):
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device='cuda:0')
    dst = torch.empty_like(src)
    for _ in range(10):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({'copy_probe_bytes_read': 8 * n, 'copy_probe_bytes_written': 8 * n,
                      'launches': 10}))


def batched_leg(steps=10, nstreams=8):
    """The batched workload (64 systems of 3k-50k atoms + 32 slices of a
    12k-atom system, mixed precision): K steps of one moil_gpu_eforce_dev
    per system in sequence (plain) against K steps of one
    moil_gpu_batch_eforce_dev over all systems (batched), both on resident
    coordinates with the lists built once; plus the force error of the
    batched mixed-precision path against FP64 on the 12k-atom slice."""
    import numpy as np
    import torch

    from moil_fnrfd_amd import abi, configs
    from moil_fnrfd_amd.eforce import MoilBatch, MoilGPU

    dev = torch.device('cuda:0')
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    work = configs.batch_workload()
    systems, coor, dpot, energ = [], [], [], []
    for s, cut, ew in work:
        m = MoilGPU(s.topo, configs.params_for(s, cut, ew, precision=abi.PREC_MIXED))
        x = torch.tensor(s.coor, dtype=torch.float64, device=dev)
        m.squeeze_dev(x.data_ptr(), sh)
        m.nbondm_dev(x.data_ptr(), sh)
        systems.append(m)
        coor.append(x)
        dpot.append(torch.empty_like(x))
        energ.append(torch.zeros(abi.N_ENERGIES, dtype=torch.float64, device=dev))
    cp = [c.data_ptr() for c in coor]
    dp = [d.data_ptr() for d in dpot]
    ep = [e.data_ptr() for e in energ]

    def plain():
        for i, m in enumerate(systems):
            m.eforce_dev(cp[i], dp[i], ep[i], sh)

    batch = MoilBatch(systems, nstreams=nstreams)

    def batched():
        batch.eforce_dev(cp, dp, ep, sh)

    res = {}
    for name, fn in (('plain', plain), ('batched', batched)):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) / steps
    # per-stage GPU time of each system run alone (events between stages)
    stages = np.zeros(5)
    for i, m in enumerate(systems):
        m.set_timing(2)
        m.eforce_dev(cp[i], dp[i], ep[i], sh)
        torch.cuda.synchronize()
        stages += np.asarray(m.stage_times()[:5])
        m.set_timing(0)
    # accuracy of the batched (mixed) forces against FP64 on the first slice
    k = len(work) - 32
    f_mixed = dpot[k].cpu().numpy()
    s, cut, ew = work[k]
    m64 = MoilGPU(s.topo, configs.params_for(s, cut, ew, precision=abi.PREC_FP64))
    x = m64.squeeze(s.coor)
    m64.nbondm(x)
    _, f64 = m64.eforce(x)
    m64.close()
    err = float(np.sqrt(((f_mixed - f64) ** 2).sum() / (f64 ** 2).sum()))
    batch.close()
    for m in systems:
        m.close()
    return {
        'systems': len(work),
        'atoms': int(sum(w[0].topo.npt for w in work)),
        'streams': nstreams,
        'steps': steps,
        'plain_ms_per_step': 1e3 * res['plain'],
        'batched_ms_per_step': 1e3 * res['batched'],
        'speedup': res['plain'] / res['batched'],
        'alone_stage_ms': {'realspace_water': stages[0], 'realspace_atoms': stages[1],
                           'pme': stages[2], 'bonded_excl': stages[3],
                           'eforce_total': stages[4]},
        'force_rms_rel_vs_fp64': err,
        'workload': '64 systems, atom counts log-uniform in [3k, 50k] (water + 1-8 '
                    'heteropolymer chains, 9 A cutoffs, 1.5 A skin, ~1 A PME grids) + 32 '
                    'slices of a 12k-atom system interpolated between two conformations',
    }


def chain_leg(steps=5, igrid=32):
    """MOIL's chain functional sds (chain/sds.f) on igrid replicas of the
    batched workload's 12k-atom system (mixed precision): one
    moil_gpu_chain_sds_dev call (list rebuild of every interior replica, one
    batched evaluation, the chain terms) against MOIL's own pattern run
    through the library one replica at a time (squeeze, lists, eforce per
    interior replica on one context, as sds does with the Fortran shim)."""
    import numpy as np
    import torch

    from moil_fnrfd_amd import abi, configs
    from moil_fnrfd_amd.eforce import MoilChain, MoilGPU

    dev = torch.device('cuda:0')
    sh = torch.cuda.current_stream().cuda_stream
    base, cut, ew = configs.batch_system(12000, 4000, 8)
    npt = base.topo.npt
    r = configs.chain_replicas(base.coor, igrid, seed=4001)
    par = configs.params_for(base, cut, ew, precision=abi.PREC_MIXED)
    ch = MoilChain([MoilGPU(base.topo, par) for _ in range(igrid - 2)], igrid)
    d_r = torch.tensor(r, dtype=torch.float64, device=dev)
    d_pick = torch.ones(npt, dtype=torch.int32, device=dev)
    d_dv = torch.empty_like(d_r)
    d_d0 = torch.empty(2 * igrid - 3, dtype=torch.float64, device=dev)
    d_e1 = torch.empty_like(d_d0)
    d_e0 = torch.zeros(igrid, dtype=torch.float64, device=dev)
    d_s = torch.zeros(1, dtype=torch.float64, device=dev)
    g, rho, lam = configs.CHAIN_GAMMA, configs.CHAIN_RHO, configs.CHAIN_LAMBDA

    def chain():
        ch.sds_dev(d_r.data_ptr(), d_pick.data_ptr(), g, rho, lam, d_dv.data_ptr(),
                   d_d0.data_ptr(), d_e0.data_ptr(), d_e1.data_ptr(), d_s.data_ptr(), sh)

    one = MoilGPU(base.topo, par)
    scr = torch.empty_like(d_r[0])
    dp = torch.empty_like(scr)
    en = torch.zeros(abi.N_ENERGIES, dtype=torch.float64, device=dev)

    def per_replica():
        for j in range(1, igrid - 1):
            scr.copy_(d_r[j])
            one.squeeze_dev(scr.data_ptr(), sh)
            one.nbondm_dev(scr.data_ptr(), sh)
            one.eforce_dev(scr.data_ptr(), dp.data_ptr(), en.data_ptr(), sh)

    res = {}
    for name, fn in (('chain', chain), ('per_replica', per_replica)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) / steps
    sener = float(d_s.cpu()[0])
    ch.close()
    one.close()
    return {
        'replicas': igrid,
        'atoms_per_replica': npt,
        'calls': steps,
        'ms_per_call': 1e3 * res['chain'],
        'per_replica_loop_ms': 1e3 * res['per_replica'],
        'speedup': res['per_replica'] / res['chain'],
        'sener_last': sener,
        'workload': '%d replicas of the batched workload\'s 12k-atom system (straight line to '
                    'a +N(0, 0.5 A) conformation, +N(0, 0.05 A) per replica), gamma=rho=100, '
                    'lambda=2, all atoms selected; every call rebuilds the interior replicas\' '
                    'lists as sds does' % igrid,
    }


def timed_loop(sysm, cut, ew, steps, warmup, seed, tune=False, shapes=None):
    """The metric's loop on one system (mixed precision, coordinates and
    outputs resident in HBM): per step N(0, 0.01 A) noise added to the
    coordinates (cumulative; drawn on the device NLIST steps at a time),
    every NLIST-th step a list rebuild (squeeze + build_lists), then
    eforce.  Times `steps`
    steps after `warmup` untimed ones with the real-space kernel's own
    dispatch events (no packets between kernels) on every 10th evaluation,
    then a short untimed pass with events between the stages for the
    breakdown."""
    import numpy as np
    import torch

    from moil_fnrfd_amd import abi, configs
    from moil_fnrfd_amd.eforce import MoilGPU

    params = configs.params_for(sysm, cut, ew, precision=abi.PREC_MIXED)
    eng = MoilGPU(sysm.topo, params)
    dev = torch.device('cuda:0')
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    coor = torch.tensor(sysm.coor, dtype=torch.float64, device=dev).contiguous()
    dpot = torch.empty_like(coor)
    energ = torch.zeros(abi.N_ENERGIES, dtype=torch.float64, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    noise = torch.empty((NLIST,) + tuple(coor.shape), dtype=torch.float64, device=dev)
    ev_list = []
    # launch shapes from the library's autotuner (before the timed region;
    # the shapes change no force, only how the evaluation maps onto the GPU)
    torch.cuda.synchronize()  # the tuner runs on the context's own stream
    if shapes:
        # fixed launch shapes (profiled runs): set, then the lists rebuilt
        eng.set_tuning(**shapes)
        eng.squeeze_dev(coor.data_ptr(), sh)
        eng.nbondm_dev(coor.data_ptr(), sh)
        torch.cuda.synchronize()
        tuning = eng.get_tuning()
    else:
        tuning = eng.autotune_dev(coor.data_ptr(), reps=10) if tune else eng.get_tuning()

    def step(k):
        if k % NLIST == 0:
            torch.randn(noise.shape, generator=gen, device=dev, dtype=torch.float64, out=noise)
        if k > 0:
            coor.add_(noise[k % NLIST], alpha=SIGMA)
        if k % NLIST == 0:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            eng.squeeze_dev(coor.data_ptr(), sh)
            eng.nbondm_dev(coor.data_ptr(), sh)
            e1.record(stream)
            ev_list.append((e0, e1))
        eng.eforce_dev(coor.data_ptr(), dpot.data_ptr(), energ.data_ptr(), sh)

    for k in range(warmup):
        step(k)
    torch.cuda.synchronize()
    n_pairs = int(eng.list_counts()[:4].sum())
    ev_list.clear()
    # the real-space kernels' dispatch events on every 10th evaluation of
    # the timed loop (on every one they cost ~2 % of the rate), on every
    # one in a short run
    eng.set_timing(1, stride=1 if steps <= 50 else 10)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        step(k)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = eng.stage_times()
    t_list_ms = [a.elapsed_time(b) for a, b in ev_list]
    eng.set_timing(2)
    for k in range(NLIST):
        step(k + 1)
    torch.cuda.synchronize()
    st2 = eng.stage_times()
    eng.set_timing(0)
    n2 = max(st2[5], 1.0)
    n_ev = max(st[5], 1.0)
    counts = eng.list_counts()
    e_last = float(energ.cpu().numpy()[0])
    eng.close()
    if not np.isfinite(e_last):
        raise RuntimeError('the last evaluation reported a fixed-point overflow (NaN energies)')
    return {
        'elapsed': t1 - t0,
        'n_pairs': n_pairs,
        'list_counts': [int(c) for c in counts],
        't_rs_ms': st[0] / n_ev,
        'stage_ms': {'realspace_water': st[0] / n_ev, 'realspace_atoms': st[1] / n_ev,
                     'pme': st2[2] / n2, 'bonded_excl': st2[3] / n2,
                     'eforce_total': st2[4] / n2,
                     'list_rebuild': (sum(t_list_ms) / len(t_list_ms)) if t_list_ms else None},
        'e_total_last': e_last,
        'tuning': tuning,
    }


def large_leg(steps=200, warmup=20):
    """The ~1.0M-atom solvated heteropolymer (20,000 chains, 215 A box,
    216^3 grid) through the metric's loop: evals/s with a list rebuild
    every NLIST evaluations, mixed precision."""
    from moil_fnrfd_amd import configs
    t0 = time.perf_counter()
    sysm, cut, ew = configs.hetero_1m()
    t_gen = time.perf_counter() - t0
    r = timed_loop(sysm, cut, ew, steps, warmup, seed=2102)
    return {
        'workload': 'hetero1m: %d atoms (20,000 chains of 10-20 atoms + %d waters), 215 A box, '
                    'rvmx=relx=10 A, rvbg=rebg=11.5 A, PME 216^3, list rebuild every %d evals'
                    % (sysm.topo.npt, sysm.meta['n_water'], NLIST),
        'atoms': sysm.topo.npt,
        'steps': steps,
        'evals_per_s': steps / r['elapsed'],
        'ms_per_step': 1e3 * r['elapsed'] / steps,
        'list_counts': r['list_counts'],
        'stage_ms': r['stage_ms'],
        'generate_s': t_gen,
    }


def adjoint_leg(n_conf=128, seed=4001, group=32):
    """Differentiation workload: the 10,048-atom heteropolymer-in-water
    system, n_conf conformations (base + N(0, 0.1 A) per coordinate).  Per
    conformation, timed: list build, forward energy + dE/dx, and the
    parameter adjoint (dE/dq per atom, dE/deps and dE/dRmin per LJ type)
    in mixed precision (FP32 pairs, FP64 reductions), `group` conformations
    at a time on as many contexts: one batched list build, one batched
    evaluation, the adjoints side by side on their own streams.  Untimed:
    the same one conformation at a time in FP64 for the error report (max
    and RMS relative over all conformations), and central finite
    differences (h = 1e-4, FP64) of a few parameters on the first
    conformation."""
    import numpy as np
    import torch

    from moil_fnrfd_amd import abi, configs
    from moil_fnrfd_amd.adjoint import lj_factors, type_gradients
    from moil_fnrfd_amd.eforce import MoilBatch, MoilGPU

    dev = torch.device('cuda:0')
    main = torch.cuda.current_stream()
    sh = main.cuda_stream
    s, cut, ew = configs.hetero_10k()
    m = s.meta
    e12, e6 = (a.numpy() for a in lj_factors(torch.tensor(m['lj_eps']),
                                             torch.tensor(m['lj_rmin_half']), m['lj_type']))
    eng = {}
    for prec in (abi.PREC_MIXED, abi.PREC_FP64):
        g = MoilGPU(s.topo, configs.params_for(s, cut, ew, precision=prec))
        g.set_params(s.topo.ptchg, e6, e12, 0.5, 0.5)
        eng[prec] = g
    npt = s.topo.npt
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    base = torch.tensor(s.coor, dtype=torch.float64, device=dev)
    X = base + 0.1 * torch.randn((n_conf, npt, 3), generator=gen, device=dev,
                                 dtype=torch.float64)
    out = {p: (torch.empty_like(X), torch.zeros((n_conf, 3, npt), dtype=torch.float64, device=dev),
               torch.zeros((n_conf, abi.N_ENERGIES), dtype=torch.float64, device=dev))
           for p in eng}

    # the mixed-precision group: `group` contexts, one batch over them
    mix = [eng[abi.PREC_MIXED]]
    for _ in range(group - 1):
        g = MoilGPU(s.topo, configs.params_for(s, cut, ew, precision=abi.PREC_MIXED))
        g.set_params(s.topo.ptchg, e6, e12, 0.5, 0.5)
        mix.append(g)
    batch = MoilBatch(mix)
    batch.set_graph(False)  # lists change before every evaluation
    side = [torch.cuda.Stream() for _ in range(group)]

    def run_grouped():
        dp, pg, en = out[abi.PREC_MIXED]
        for b0 in range(0, n_conf, group):
            idx = list(range(b0, b0 + group))
            xs = [X[b].data_ptr() for b in idx]
            batch.build_lists_dev(xs, squeeze=True, stream=sh)
            batch.eforce_dev(xs, [dp[b].data_ptr() for b in idx],
                             [en[b].data_ptr() for b in idx], sh)
            for k, b in enumerate(idx):
                side[k].wait_stream(main)
                mix[k].param_grad_dev(X[b].data_ptr(), pg[b, 0].data_ptr(), pg[b, 1].data_ptr(),
                                      pg[b, 2].data_ptr(), side[k].cuda_stream)
            for st in side:
                main.wait_stream(st)

    def run_one(prec):
        g = eng[prec]
        dp, pg, en = out[prec]
        for b in range(n_conf):
            x = X[b]
            g.squeeze_dev(x.data_ptr(), sh)
            g.nbondm_dev(x.data_ptr(), sh)
            g.eforce_dev(x.data_ptr(), dp[b].data_ptr(), en[b].data_ptr(), sh)
            g.param_grad_dev(x.data_ptr(), pg[b, 0].data_ptr(), pg[b, 1].data_ptr(),
                             pg[b, 2].data_ptr(), sh)

    assert n_conf % group == 0
    run_grouped()  # warm-up (and the squeezed coordinates)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run_grouped()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # one conformation at a time on one context, for comparison
    t0 = time.perf_counter()
    run_one(abi.PREC_MIXED)
    torch.cuda.synchronize()
    dt_one = time.perf_counter() - t0
    run_grouped()  # the grouped results are the ones checked below
    run_one(abi.PREC_FP64)
    torch.cuda.synchronize()
    batch.close()

    def errs(a, b):
        a, b = a.reshape(len(a), -1), b.reshape(len(b), -1)
        rms = float(torch.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))
        mx = float(((a - b).abs().amax(dim=1) / b.abs().amax(dim=1)).max())
        return {'rms_rel': rms, 'max_rel': mx}

    d32, p32, _ = out[abi.PREC_MIXED]
    d64, p64, _ = out[abi.PREC_FP64]
    t32 = np.stack([np.concatenate(type_gradients(p[1], p[2], e12, e6, m['lj_eps'],
                                                  m['lj_rmin_half'], m['lj_type']))
                    for p in p32.cpu().numpy()])
    t64 = np.stack([np.concatenate(type_gradients(p[1], p[2], e12, e6, m['lj_eps'],
                                                  m['lj_rmin_half'], m['lj_type']))
                    for p in p64.cpu().numpy()])
    rep = {
        'dE_dx': errs(d32, d64),
        'dE_dq': errs(p32[:, 0], p64[:, 0]),
        'dE_deps_dRmin_per_type': errs(torch.tensor(t32), torch.tensor(t64)),
    }
    # finite differences of the FP64 energy on the first conformation
    g = eng[abi.PREC_FP64]
    x0 = X[0].cpu().numpy()
    g.nbondm(x0)
    dq0, de12_0, de6_0 = g.param_grad(x0)
    geps, grmh = type_gradients(de12_0, de6_0, e12, e6, m['lj_eps'], m['lj_rmin_half'],
                                m['lj_type'])
    h = 1e-4
    q0 = s.topo.ptchg
    worst = 0.0
    for i in (0, 100, 1000, 2000):
        dqv = np.zeros(npt)
        dqv[i] = h
        ep, em = [], []
        for sgn, acc in ((1, ep), (-1, em)):
            g.set_params(q0 + sgn * dqv, e6, e12, 0.5, 0.5)
            acc.append(g.eforce(x0)[0]['e_total'])
        fd = (ep[0] - em[0]) / (2 * h)
        worst = max(worst, abs(fd - dq0[i]) / max(abs(fd), 1.0))
    eps0 = torch.tensor(m['lj_eps'])
    rmh0 = torch.tensor(m['lj_rmin_half'])
    for t in (0, 7, 15, len(eps0) - 1):
        for which, ad in ((0, geps[t]), (1, grmh[t])):
            vals = []
            for sgn in (1, -1):
                ee, rr = eps0.clone(), rmh0.clone()
                (ee if which == 0 else rr)[t] *= 1 + sgn * h
                a12, a6 = lj_factors(ee, rr, m['lj_type'])
                g.set_params(q0, a6.numpy(), a12.numpy(), 0.5, 0.5)
                vals.append(g.eforce(x0)[0]['e_total'])
            step = 2 * h * float((eps0 if which == 0 else rmh0)[t])
            fd = (vals[0] - vals[1]) / step
            worst = max(worst, abs(fd - ad) / max(abs(fd), 1.0))
    rep['fd_max_rel_fp64'] = worst
    for g in eng.values():
        g.close()
    for g in mix[1:]:
        g.close()
    return {
        'conformations': n_conf,
        'atoms': npt,
        'lj_types': len(m['lj_eps']),
        'conf_per_s': n_conf / dt,
        'ms_per_conf': 1e3 * dt / n_conf,
        'group': group,
        'one_at_a_time_conf_per_s': n_conf / dt_one,
        'errors_mixed_vs_fp64': rep,
        'workload': '10,048-atom heteropolymer in water (200 chains, 46.4 A box, 48^3 grid), '
                    'conformations = base + N(0, 0.1 A); per conformation list build + energy '
                    '+ dE/dx + dE/dq per atom + dE/deps, dE/dRmin per LJ type (mixed), '
                    '%d conformations at a time (batched lists and evaluation, adjoints on '
                    'their own streams)' % group,
    }


def _progress(msg):
    """One line per leg on stderr (stdout carries the JSON line only)."""
    print('bench: %s' % msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-tune', action='store_true',
                    help='keep the default launch shapes instead of autotuning them')
    ap.add_argument('--shapes', default='',
                    help='fixed launch shapes instead of the autotuner, e.g. '
                         'wchunk_pairs=1024,water_wg_per_cu=4 (profiled runs)')
    ap.add_argument('--no-batched', action='store_true',
                    help='skip the batched/plain leg (96 systems)')
    ap.add_argument('--batched-only', action='store_true',
                    help='run only the batched/plain leg and print its JSON')
    ap.add_argument('--batched-steps', type=int, default=10)
    ap.add_argument('--no-chain', action='store_true',
                    help='skip the chain-functional leg (32 replicas)')
    ap.add_argument('--chain-only', action='store_true',
                    help='run only the chain-functional leg and print its JSON')
    ap.add_argument('--no-adjoint', action='store_true',
                    help='skip the differentiation leg (128 conformations)')
    ap.add_argument('--large-only', action='store_true',
                    help='run only the ~1.0M-atom heteropolymer loop and print its JSON')
    ap.add_argument('--large-steps', type=int, default=200)
    ap.add_argument("--adjoint-group", type=int, default=32,
                    help='conformations per group in the differentiation leg')
    ap.add_argument('--adjoint-only', action='store_true',
                    help='run only the differentiation leg and print its JSON')
    ap.add_argument('--copy-probe', type=int, default=0, metavar='MIB',
                    help='only run a plain device copy of MIB MiB ten times (the reference '
                         'kernel for scaling the FETCH_SIZE / WRITE_SIZE counters)')
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print('bench.py: this build runs on one GPU only (--gpus %d requested); the multi-GPU '
              'decomposition is a design note, not an implemented path.' % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch

    from moil_fnrfd_amd import configs

    torch.cuda.set_device(0)
    if args.batched_only:
        print(json.dumps(batched_leg(steps=args.batched_steps)))
        return
    if args.adjoint_only:
        print(json.dumps(adjoint_leg(group=args.adjoint_group)))
        return
    if args.chain_only:
        print(json.dumps(chain_leg()))
        return
    if args.large_only:
        print(json.dumps(large_leg(steps=args.large_steps)))
        return
    sysm, cut, ew = configs.water32k()
    _progress('main loop: water32k, %d steps after %d warm-up' % (args.steps, args.warmup))
    shapes = dict((kv.split('=')[0], int(kv.split('=')[1])) for kv in args.shapes.split(',') if kv)
    r = timed_loop(sysm, cut, ew, args.steps, args.warmup, seed=1001, tune=not args.no_tune,
                   shapes=shapes)
    evals_per_s = args.steps / r['elapsed']
    n_pairs = r['n_pairs']
    t_rs_ms = r['t_rs_ms']
    B_rs = algorithmic_bytes_realspace(n_pairs, sysm.topo.npt)
    achieved = B_rs / (t_rs_ms * 1e-3) / 1e9
    peak = 8000.0
    traffic = measured_traffic()
    line = {
        'metric': 'force evals/s (96k-atom PME) + batched/plain speedup at force rel-err<=1e-5',
        'value': evals_per_s,
        'unit': 'evals/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': 1e3 * r['elapsed'] / args.steps,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'fp32 pair math, int64 fixed-point forces, fp64 energies',
        'data': 'synthetic: seeded lattice+jitter water box, N(0,0.01 A) noise per step',
        'config': {
            'workload': 'water32k: 32,768 rigid 3-site waters (98,304 atoms), 68.9 A box, '
                        'rvmx=relx=10 A, rvbg=rebg=11.5 A, PME 72^3 order 4 beta=0.31, '
                        'list rebuild every 20 evals',
            'atoms': sysm.topo.npt,
            'water_pairs': n_pairs,
            'list_rebuild_every': NLIST,
            'stage_ms': r['stage_ms'],
            'e_total_last': r['e_total_last'],
            # launch shapes (moil_gpu_autotune) and the mean evaluation time
            # with the default and the chosen shapes
            'tuning': r['tuning'],
        },
        'roofline': {
            # the figure is algorithmic bytes against HBM peak; the kernel is
            # limited by VALU issue (counters: profiles/)
            'bound': 'hbm',
            'limiter': 'valu',
            'achieved': achieved,
            'peak': peak,
            'unit': 'GB/s',
            'frac': achieved / peak,
            'traffic': traffic,
            'kernel': RS_KERNEL,
            'algorithmic_bytes_per_launch': B_rs,
            # the measured HBM bytes over the same live duration (below the
            # algorithmic figure: j-site data is re-read from L2 and LDS)
            'frac_measured_traffic': (traffic / (t_rs_ms * 1e-3) / 1e9 / peak) if traffic else None,
            'pair_interactions_per_s': 9.0 * n_pairs / (t_rs_ms * 1e-3),
        },
    }
    if not args.no_batched:
        _progress('batched leg: 96 systems')
        line['batched'] = batched_leg(steps=args.batched_steps)
    if not args.no_chain:
        _progress('chain leg: 32 replicas')
        line['chain'] = chain_leg()
    if not args.no_adjoint:
        _progress('differentiation leg: 128 conformations')
        line['adjoint'] = adjoint_leg(group=args.adjoint_group)
    if not args.no_cpu_baseline:
        _progress('CPU baseline: MOIL on one core')
        line['cpu_baseline'] = cpu_baseline(sysm, cut, ew)
    print(json.dumps(line))


if __name__ == '__main__':
    main()
