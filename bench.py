"""Benchmark: chi2Profile evaluations per second on one MI355X.

Workload: ring+envelope radial profile (Nr=1000, Rmin=1e-4", dR=0.002"),
nxy=4096, dxy=0.01", inc=40 deg, PA=120 deg, dRA=dDec=0.02", Nvis=1e6
log-uniform baselines in [5e3, 1e7] wavelengths, noisy observation; float64;
galario's central-pixel guard off (it rejects dR=0.002" at dxy=0.01").
A step is one chi2Profile evaluation with all inputs resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S]

Prints one JSON line.  `value` is timed with HIP events on the kernels' stream
around exactly K calls (no per-stage events in that loop); `steady_state`
repeats the measurement over 250 calls in 5 blocks and reports the median
block.  Every rate carries its HBM-roofline fraction `frac` = algorithmic
bytes per evaluation x evaluations/s / 8 TB/s.  `roofline` is measured
live: each forward kernel is launched 20 times back to back inside a pair of
HIP events (galario_stage_repeat), giving its average launch duration;
`achieved` = algorithmic bytes / duration for the kernel that moves the most
HBM bytes per evaluation (each forward kernel's figures are in `kernels`).
`traffic` is that kernel's HBM bytes per launch from the committed rocprofv3
FETCH_SIZE / WRITE_SIZE summary (calibrated on a copy kernel).  `cpu_baseline`
times the C/OpenMP restatement of galario's CPU loops (FFT by scipy.fft) on a
bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "chi2Profile evals/s fwd & fwd+grad, nxy=4096 Nvis=1e6 FP64; % HBM roofline"


def stage_bytes(nxy, nvis, nr, rmax, grad=False, s=8, kw=1):
    """Algorithmic HBM bytes per launch of each stage of the profile path (each
    needed array touched once), s bytes per stored real.  The swept image is
    even in x and y, so its spectrum is real and even: the row pass writes
    Rs'+1 real rows (Rs' = min(rmax+1, nxy/2)), the column pass reads them and
    writes the (nxy/2+1) x (nxy/2+1) real half spectrum H, the sample kernel
    streams u, v, Re, Im, w (once per kw parameter sets in the batched form)
    and gathers 4 real corners per point.  With the gradient it also scatters
    4 corner gradients (FP64 atomics) into the adjoint Hh (double in both
    precisions); the adjoint column pass reads Hh, clears it and writes Rs'+1
    adjoint rows; the adjoint row pass reads them (the profile gradient is a
    few KB)."""
    C = nxy // 2 + 1
    rs = min(rmax + 1, nxy // 2)
    rows = rs + 1
    half = C * (nxy // 2 + 1) * s
    half_adj = C * (nxy // 2 + 1) * 8
    b = {
        'row_pass': rows * C * s + nr * s,
        'column_pass': rows * C * s + half,
        'sample': int(5 * s * nvis / kw) + 4 * s * nvis,
    }
    if grad:
        b['sample'] += 4 * 8 * nvis
        b['adjoint_column_pass'] = 2 * half_adj + rows * C * s
        b['adjoint_row_pass'] = rows * C * s + 2 * nr * 8
    return b


def swept_rmax(nxy, nr, Rmin, dR, dxy):
    """Half side of the swept box: min(ceil((Rmin + nr dR) / dxy), nxy/2) (galario.cpp:1200)."""
    import math
    return min(int(math.ceil((Rmin + nr * dR) / dxy)), nxy // 2)


def eval_bytes(grids, nr, grad=False, s=8, kw=1):
    """Algorithmic bytes of one parameter set's evaluation over datasets grouped by grid: grids is a
    list of (nxy, rmax, [nvis of each dataset on that grid]); one spectrum (and, with grad, one adjoint
    spectrum) per grid, one sample per dataset."""
    tot = 0
    for nxy, rmax, nvis_list in grids:
        b = stage_bytes(nxy, 0, nr, rmax, grad=grad, s=s, kw=kw)
        tot += sum(v for k, v in b.items() if k != 'sample')
        for nv in nvis_list:
            tot += stage_bytes(nxy, nv, nr, rmax, grad=grad, s=s, kw=kw)['sample']
    return tot


def frac_of(nbytes, evals_per_s):
    """HBM-roofline fraction: algorithmic bytes per evaluation x evaluations/s over 8 TB/s."""
    return round(nbytes * evals_per_s / (HBM_PEAK_GBS * 1e9), 4)


KERNEL_NAMES = {'row_pass': 'k_rowpass_sym<double, 12>', 'column_pass': 'k_colpass_sub<double, 12>',
                'sample': 'k_sample_chi2_n<double, true, 2>'}
TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'hbm_traffic_profile_nxy4096_nvis1e6_f64.txt')


def measured_traffic(kernel, nxy, nvis):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (lines 'traffic <kernel> <bytes>'), for the default workload only; None otherwise."""
    if (nxy, nvis) != (4096, 1_000_000) or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        for line in f:
            parts = line.rstrip('\n').split('\t')
            if len(parts) == 3 and parts[0] == 'traffic' and parts[1] == kernel:
                return int(parts[2])
    return None


def host_cpu() -> str:
    """CPU model and logical CPU count of this host (for the CPU baseline)."""
    model = 'unknown CPU'
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    return f'{model}, {os.cpu_count()} logical CPUs'


def batched_rates(ta, ds, wl, I, chi2_plain, counts, counts_grad, steps, stream):
    """Batched form over the same dataset: P parameter sets (walker 0 = the plain call's parameters, the
    others jittered: profile x lognormal(0.1), inc, PA, dRA, dDec perturbed) per launch of the captured
    graph; evals/s = P x graph launches/s, timed like the plain form."""
    import numpy as np
    import torch
    counts = [int(c) for c in counts.split(',') if c.strip()]
    counts_grad = [int(c) for c in counts_grad.split(',') if c.strip()]
    pmax = max(counts + counts_grad)
    rng = np.random.default_rng(7)
    nr = len(wl.intensity)
    Ib = np.repeat(wl.intensity[None, :], pmax, 0) * np.exp(0.1 * rng.standard_normal((pmax, 1)))
    g0 = np.array([wl.inc, wl.PA, wl.dRA, wl.dDec])
    gb = g0[None, :] + np.array([0.05, 0.1, 0.002 * 4.84813681109536e-06, 0.002 * 4.84813681109536e-06])[None, :] \
        * rng.standard_normal((pmax, 4))
    Ib[0], gb[0] = wl.intensity, g0
    dev = I.device
    It = torch.from_numpy(np.ascontiguousarray(Ib)).to(dev)
    Gt = torch.from_numpy(np.ascontiguousarray(gb)).to(dev)
    plan = ta.BatchedChi2Profile([(ds, wl.nxy, wl.dxy)], nr, wl.Rmin, wl.dR, max_walkers=pmax,
                                 check_central_pixel=False)
    grids = [(wl.nxy, swept_rmax(wl.nxy, nr, wl.Rmin, wl.dR, wl.dxy), [len(wl.u)])]
    out = {'workload': f'batched form: P parameter sets on the same dataset (nxy={wl.nxy}, Nvis={len(wl.u)}), '
                       f'one captured HIP graph per call; frac = algorithmic bytes per set (the streams counted '
                       f'once per walkers_per_sample sets) x sets/s / 8 TB/s', 'forward': {}, 'forward_grad': {}}

    def rate(fn, P, reps):
        fn()                                                  # tuning + capture
        torch.cuda.synchronize()
        for _ in range(2):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        return ms

    for P in counts:
        Ip, Gp = It[:P].contiguous(), Gt[:P].contiguous()
        reps = max(5, min(steps, int(2000 // P)))
        ms = rate(lambda: plan(Ip, Gp), P, reps)
        c = plan(Ip, Gp)
        torch.cuda.synchronize()
        rel = abs(float(c[0]) - chi2_plain) / abs(chi2_plain)
        chunk, kw, _ = plan.info(P)
        ev = P * 1000.0 / ms
        nb = eval_bytes(grids, nr, kw=kw)
        out['forward'][str(P)] = {'evals_per_s': round(ev, 1), 'ms_per_batch': round(ms, 4),
                                  'chunk': chunk, 'walkers_per_sample': kw, 'point_blocks': plan.tile(P)[0],
                                  'chi2_rel_diff_vs_plain': rel,
                                  'alg_bytes_per_eval': nb, 'frac': frac_of(nb, ev)}
    for P in counts_grad:
        Ip, Gp = It[:P].contiguous(), Gt[:P].contiguous()
        reps = max(5, min(steps, int(1000 // P)))
        ms = rate(lambda: plan.grad(Ip, Gp), P, reps)
        chunk, _, _ = plan.info(P, grad=True)
        ev = P * 1000.0 / ms
        nb = eval_bytes(grids, nr, grad=True)
        out['forward_grad'][str(P)] = {'evals_per_s': round(ev, 1), 'ms_per_batch': round(ms, 4),
                                       'chunk': chunk, 'grad_window': plan.tile(P, grad=True)[1],
                                       'alg_bytes_per_eval': nb, 'frac': frac_of(nb, ev)}
    plan.close()
    return out


def extra_configs(ta, g, stream, steps):
    """The other configurations of the benchmark suite (each timed like the main one, fewer steps):
    differentiation at nxy=2048 (plain gradient, FP64 and FP32), the batched forward of 256 walkers over
    three datasets of uneven nxy and Nvis, and the batched forward+gradient of 64 sets over three nxy=4096
    datasets (FP64 and FP32, errors against the FP64 plain path), the multi-source chi2 (K = 1 and 4 sources)
    on the main shape, and vectorised HMC on galario's quickstart posterior (64 chains)."""
    import numpy as np
    import torch
    from galario_amd import synthetic as syn
    A = syn.ARCSEC
    dev = torch.device('cuda:0')
    f64 = torch.float64

    def model(wl):
        return g.sampleProfile(wl.intensity, wl.Rmin, wl.dR, wl.nxy, wl.dxy, wl.u, wl.v, check_central_pixel=False,
                               **wl.kwargs())

    def timeit(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    out = {}
    # differentiation: nxy=2048, Nvis=2e6, Nr=500, dR=0.004", 16 parameter points
    wl = syn.ring_profile_workload(model, nvis=2_000_000, nxy=2048, nr=500, seed=2, dR=0.004 * A)
    I16, G16 = syn.walker_ensemble(np.random.default_rng(2), 16, 500, wl.Rmin, wl.dR, wl.inc, wl.PA, wl.dRA,
                                   wl.dDec)
    res = {'workload': 'plain chi2Profile + gradient, nxy=2048, Nvis=2e6, Nr=500, 16 parameter points'}
    grads = {}
    for name, dt in (('f64', torch.float64), ('f32', torch.float32)):
        ds = ta.Dataset(wl.u, wl.v, wl.re, wl.im, wl.w, device=dev, dtype=dt)
        It = torch.from_numpy(I16).to(dev, dt)
        calls = [dict(inc=float(G16[k, 0]), PA=float(G16[k, 1]), dRA=float(G16[k, 2]), dDec=float(G16[k, 3]))
                 for k in range(16)]
        it = iter(range(10 ** 9))

        def one():
            k = next(it) % 16
            return ds.chi2_profile_grad(It[k], wl.Rmin, wl.dR, wl.nxy, wl.dxy, check_central_pixel=False, **calls[k])

        ms = timeit(one, max(16, min(steps, 64)))
        grads[name] = [one() for _ in range(16)]
        nb = eval_bytes([(wl.nxy, swept_rmax(wl.nxy, 500, wl.Rmin, wl.dR, wl.dxy), [len(wl.u)])], 500, grad=True,
                        s=8 if dt == torch.float64 else 4)
        res[name] = {'fwd_grad_evals_per_s': round(1000.0 / ms, 1), 'ms_per_eval': round(ms, 4),
                     'alg_bytes_per_eval': nb, 'frac': frac_of(nb, 1000.0 / ms)}
        del ds
    rel = max(float(torch.linalg.norm(grads['f32'][k][1] - grads['f64'][k][1]) / torch.linalg.norm(grads['f64'][k][1]))
              for k in range(16))
    res['f32_grad_I_rel_err_vs_f64'] = rel
    out['differentiation_nxy2048_nvis2e6'] = res

    # batched forward: 256 walkers x datasets (512, 1e4), (1024, 2e5), (2048, 1e6); Nr=300
    bw = syn.walkers_uneven_grids_workload(model)
    sets = [(ta.Dataset(d['u'], d['v'], d['re'], d['im'], d['w'], device=dev, dtype=f64), d['nxy'], d['dxy'])
            for d in bw.sets]
    plan = ta.BatchedChi2Profile(sets, bw.nr, bw.Rmin, bw.dR, max_walkers=256, check_central_pixel=False)
    It, Gt = torch.from_numpy(bw.intensity).to(dev), torch.from_numpy(bw.geom).to(dev)
    ms = timeit(lambda: plan(It, Gt), 10)
    chunk, kw, _ = plan.info(256)
    per = plan(It, Gt, return_sets=True)[1][0].cpu().numpy()
    G0 = bw.geom[0]
    ref = [ds.chi2_profile(It[0], bw.Rmin, bw.dR, nxy, dx, check_central_pixel=False, inc=float(G0[0]),
                           PA=float(G0[1]), dRA=float(G0[2]), dDec=float(G0[3])).item()
           for ds, nxy, dx in sets]
    grids = [(d['nxy'], swept_rmax(d['nxy'], bw.nr, bw.Rmin, bw.dR, d['dxy']), [len(d['u'])]) for d in bw.sets]
    nb = eval_bytes(grids, bw.nr, kw=kw)
    ev = 256 * 1000.0 / ms
    out['batched_256_walkers_3_datasets_fwd'] = {
        'workload': '256 walkers (Nr=300) x datasets (nxy, Nvis) = (512, 1e4), (1024, 2e5), (2048, 1e6), FP64, '
                    'per-walker chi2 summed over the datasets',
        'walker_evals_per_s': round(ev, 1), 'ms_per_batch': round(ms, 4), 'chunk': chunk,
        'walkers_per_sample': kw, 'point_blocks': plan.tile(256)[0], 'alg_bytes_per_walker_eval': nb,
        'frac': frac_of(nb, ev),
        'chi2_rel_diff_vs_plain': float(np.max(np.abs(per - ref) / np.abs(ref)))}
    plan.close()
    del sets

    # batched forward+gradient: 64 sets x 3 datasets at nxy=4096 (Nvis 5e5, 1.5e6, 3e6), Nr=500: FP64, FP32
    # bulk with FP64 geometry / phase / per-point arithmetic, and pure FP32, each against the FP64 plain path
    bw = syn.sets_three_datasets_4096_workload(model)
    res = {'workload': '64 parameter sets (Nr=500) x 3 datasets at nxy=4096 (Nvis 5e5, 1.5e6, 3e6; 5e6 total), '
                       'chi2 + gradient per set; errors of sets 0, 8, ..., 56 against the FP64 plain path'}
    grids = [(4096, swept_rmax(4096, bw.nr, bw.Rmin, bw.dR, bw.sets[0]['dxy']), [len(d['u']) for d in bw.sets])]
    sets64 = [(ta.Dataset(d['u'], d['v'], d['re'], d['im'], d['w'], device=dev, dtype=f64), d['nxy'], d['dxy'])
              for d in bw.sets]
    It64 = torch.from_numpy(bw.intensity).to(dev)
    check = list(range(0, 64, 8))
    refs = {}
    for w in check:
        G = bw.geom[w]
        r = [ds.chi2_profile_grad(It64[w], bw.Rmin, bw.dR, nxy, dxy, check_central_pixel=False, inc=float(G[0]),
                                  PA=float(G[1]), dRA=float(G[2]), dDec=float(G[3])) for ds, nxy, dxy in sets64]
        refs[w] = (sum(x[0].item() for x in r), sum(x[1] for x in r), sum(x[2] for x in r))
    for name, dt, mixed in (('f64', torch.float64, False), ('f32_bulk_f64_phase', torch.float32, True),
                            ('f32', torch.float32, False)):
        sets = sets64 if dt == f64 else [(ta.Dataset(d['u'], d['v'], d['re'], d['im'], d['w'], device=dev,
                                                     dtype=dt), d['nxy'], d['dxy']) for d in bw.sets]
        plan = ta.BatchedChi2Profile(sets, bw.nr, bw.Rmin, bw.dR, max_walkers=64, check_central_pixel=False,
                                     phase_f64=mixed)
        It = torch.from_numpy(bw.intensity).to(dev, dt)
        Gt = torch.from_numpy(bw.geom).to(dev, f64 if mixed else dt)
        ms = timeit(lambda: plan.grad(It, Gt), 5)
        c, _, gI, gg = plan.grad(It, Gt)
        chunk, _, _ = plan.info(64, grad=True)
        ev = 64 * 1000.0 / ms
        nb = eval_bytes(grids, bw.nr, grad=True, s=8 if dt == f64 else 4)
        e = [0.0, 0.0, 0.0]
        for w in check:
            c_ref, gI_ref, gg_ref = refs[w]
            e = [max(e[0], abs(float(c[w]) - c_ref) / abs(c_ref)),
                 max(e[1], float(torch.linalg.norm(gI[w] - gI_ref) / torch.linalg.norm(gI_ref))),
                 max(e[2], float(torch.max(torch.abs(gg[w] - gg_ref) / torch.abs(gg_ref))))]
        res[name] = {'fwd_grad_set_evals_per_s': round(ev, 1), 'ms_per_batch': round(ms, 4), 'chunk': chunk,
                     'grad_window': plan.tile(64, grad=True)[1],
                     'alg_bytes_per_set_eval': nb, 'frac': frac_of(nb, ev),
                     'chi2_rel_err_vs_f64_plain': e[0], 'grad_I_rel_err_vs_f64_plain': e[1],
                     'grad_geom_max_rel_err_vs_f64_plain': e[2]}
        plan.close()
        del plan
        if dt != f64:
            del sets
    del sets64
    out['batched_64_sets_3_datasets_nxy4096_grad'] = res

    # multi-source summation: K sources sharing the profile, inc and PA, on the main workload's shape
    wl = syn.ring_profile_workload(model, nvis=1_000_000, nxy=4096, seed=1)
    ds = ta.Dataset(wl.u, wl.v, wl.re, wl.im, wl.w, device=dev, dtype=f64)
    It = torch.from_numpy(np.ascontiguousarray(wl.intensity)).to(dev)
    args = (It, wl.Rmin, wl.dR, wl.nxy, wl.dxy)
    res = {'workload': 'chi2 of K sources (flux, dRA, dDec) sharing one profile, inc and PA, nxy=4096, '
                       'Nvis=1e6, FP64 (one spectrum, K phase-shifted samples per point)'}
    for K in (1, 4):
        src = torch.tensor([[1.0 / K, wl.dRA * (1 + 0.5 * k), wl.dDec * (1 - 0.3 * k)] for k in range(K)],
                           dtype=f64, device=dev)
        ms = timeit(lambda: ds.chi2_profile_multi(*args, src, PA=wl.PA, inc=wl.inc, check_central_pixel=False),
                    max(20, min(steps, 200)))
        nb = eval_bytes([(wl.nxy, swept_rmax(wl.nxy, len(wl.intensity), wl.Rmin, wl.dR, wl.dxy), [len(wl.u)])],
                        len(wl.intensity))
        res[f'K{K}'] = {'evals_per_s': round(1000.0 / ms, 1), 'ms_per_eval': round(ms, 4),
                        'alg_bytes_per_eval': nb, 'frac': frac_of(nb, 1000.0 / ms)}
    one = torch.tensor([[1.0, wl.dRA, wl.dDec]], dtype=f64, device=dev)
    c1 = ds.chi2_profile_multi(*args, one, PA=wl.PA, inc=wl.inc, check_central_pixel=False).item()
    c0 = ds.chi2_profile(*args, PA=wl.PA, inc=wl.inc, dRA=wl.dRA, dDec=wl.dDec, check_central_pixel=False).item()
    res['one_source_rel_diff_vs_plain'] = abs(c1 - c0) / abs(c0)
    out['multisource_nxy4096_nvis1e6'] = res
    del ds

    # gradient sampler: galario's quickstart posterior (6 parameters) for 64 chains, leapfrog HMC
    from galario_amd import sampler as S
    rng = np.random.default_rng(6)
    nxy, dxy, nR = 1024, 0.02 * A, 500
    Rmin, dR = 1e-4 * A, 0.01 * A
    truth = np.array([10.0, 0.4, 50.0, 30.0, 0.05, -0.05])
    qs = syn.ProfileWorkload(intensity=10 ** truth[0] * np.exp(-0.5 * ((Rmin + dR * np.arange(nR)) /
                                                                      (truth[1] * A)) ** 2),
                             Rmin=Rmin, dR=dR, nxy=nxy, dxy=dxy, inc=truth[2] * syn.DEG, PA=truth[3] * syn.DEG,
                             dRA=truth[4] * A, dDec=truth[5] * A, u=None, v=None)
    qs.u, qs.v = syn.log_uniform_uv(rng, 200_000, 2 / (nxy * dxy), 0.4 / dxy)
    re, im, w = syn.noisy_observation(rng, model(qs))
    ds = ta.Dataset(qs.u, qs.v, re, im, w, device=dev, dtype=f64)
    W, n_steps, n_leap = 64, 3, 10
    plan = ta.BatchedChi2Profile([(ds, nxy, dxy)], nR, Rmin, dR, max_walkers=W, check_central_pixel=False)
    post = S.QuickstartPosterior(plan, [[1, 20], [0, 8], [0, 90], [0, 180], [-2, 2], [-2, 2]], Rmin, dR, nR)
    x0 = torch.tensor(truth[None, :] + 0.01 * rng.standard_normal((W, 6)) * np.array([1, 1, 10, 10, 1, 1]),
                      device=dev)
    scale = torch.tensor([1e-3, 1e-3, 1e-1, 1e-1, 1e-3, 1e-3], dtype=f64, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    S.hmc(post, x0, n_steps=1, step_size=scale, n_leapfrog=2, generator=gen)       # plan set-up and autotune
    ms = timeit(lambda: S.hmc(post, x0, n_steps=n_steps, step_size=scale, n_leapfrog=n_leap, generator=gen), 2)
    grads = n_steps * n_leap + 1
    ev = W * grads * 1000.0 / ms
    nb = eval_bytes([(nxy, swept_rmax(nxy, nR, Rmin, dR, dxy), [len(qs.u)])], nR, grad=True)
    out['hmc_quickstart_posterior_64_chains'] = {
        'workload': "galario's quickstart posterior (Gaussian profile, 6 parameters) on 2e5 visibilities, "
                    f'nxy={nxy}, Nr={nR}, FP64: {W} HMC chains x {n_steps} steps x {n_leap} leapfrog steps',
        'chain_gradient_evals_per_s': round(ev, 1), 'ms_per_hmc_call': round(ms, 3),
        'alg_bytes_per_chain_gradient': nb, 'frac': frac_of(nb, ev)}
    plan.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--warmup', type=int, default=50)
    ap.add_argument('--cpu-seconds', type=float, default=12.0, help='CPU-baseline time budget (0: skip)')
    ap.add_argument('--nvis', type=int, default=1_000_000)
    ap.add_argument('--nxy', type=int, default=4096)
    ap.add_argument('--batch', type=str, default='1,8,64,256',
                    help='parameter-set counts of the batched form (comma list; empty: skip)')
    ap.add_argument('--batch-grad', type=str, default='1,8,64',
                    help='parameter-set counts of the batched forward+gradient')
    ap.add_argument('--extras', type=int, default=1, help='also time the other configurations (0: skip)')
    ap.add_argument('--steady-calls', type=int, default=250,
                    help='calls of the steady-state measurement of each plain form (0: skip)')
    ap.add_argument('--copy-calibration-mb', type=int, default=0,
                    help='first run 5 device-to-device copies of this many MiB (HBM byte-counter calibration)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): sharding walkers or uv points "
              f"over several GPUs is designed but not implemented.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import galario_amd
    from galario_amd import synthetic as syn
    from galario_amd import torch_api as ta

    g = galario_amd.double
    dev = torch.device('cuda:0')
    torch.cuda.set_device(dev)
    if args.copy_calibration_mb:
        n = args.copy_calibration_mb * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    def gpu_model(wl):
        return g.sampleProfile(wl.intensity, wl.Rmin, wl.dR, wl.nxy, wl.dxy, wl.u, wl.v,
                               check_central_pixel=False, **wl.kwargs())

    wl = syn.ring_profile_workload(gpu_model, nvis=args.nvis, nxy=args.nxy, seed=1)
    t_up = time.perf_counter()
    ds = ta.Dataset(wl.u, wl.v, wl.re, wl.im, wl.w, device=dev, dtype=torch.float64)   # uploaded once, sorted
    torch.cuda.synchronize()
    upload_ms = (time.perf_counter() - t_up) * 1e3
    I = torch.from_numpy(np.ascontiguousarray(wl.intensity)).to(dev)
    out = torch.empty(1, dtype=torch.float64, device=dev)
    kw = wl.kwargs()

    def step():
        ds.chi2_profile(I, wl.Rmin, wl.dR, wl.nxy, wl.dxy, check_central_pixel=False, out=out, **kw)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    chi2_dev = float(out.item())
    chi2_host = g.chi2Profile(*wl.args(), check_central_pixel=False, **kw)
    assert abs(chi2_dev - chi2_host) <= 1e-12 * abs(chi2_host), (chi2_dev, chi2_host)
    assert not ta.take_oob_flag()

    import math
    rmax = min(int(math.ceil((wl.Rmin + len(wl.intensity) * wl.dR) / wl.dxy)), wl.nxy // 2)  # swept box half side
    stream = torch.cuda.current_stream(dev)

    def run_timed(fn, steps, blocks=10):
        """Exactly `steps` calls between events on the kernels' stream, bracketed by synchronisations,
        in at most `blocks` blocks of (nearly) equal size and at least 50 calls each (a timing event
        between blocks costs the stream a few microseconds): (mean ms per call over all steps, std, min
        and median over the blocks' per-call means, steps)."""
        blocks = max(1, min(blocks, steps // 50))
        sizes = [steps // blocks + (1 if b < steps % blocks else 0) for b in range(blocks)]
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(blocks + 1)]
        torch.cuda.synchronize()
        ev[0].record(stream)
        for b in range(blocks):
            for _ in range(sizes[b]):
                fn()
            ev[b + 1].record(stream)
        torch.cuda.synchronize()
        blk = np.array([ev[b].elapsed_time(ev[b + 1]) / sizes[b] for b in range(blocks)])
        return (float(ev[0].elapsed_time(ev[blocks]) / steps), float(blk.std()), float(blk.min()),
                float(np.median(blk)), steps)

    def steady_state(fn, blocks=5):
        """A separate, longer run for a robust per-call figure: --steady-calls calls in `blocks` blocks,
        the median block's per-call time (and the spread of the blocks)."""
        if args.steady_calls <= 0:
            return None
        mean, std, mn, med, n = run_timed(fn, args.steady_calls, blocks)
        return {'calls': n, 'blocks': blocks, 'median_block_ms_per_call': round(med, 5),
                'min_block_ms_per_call': round(mn, 5), 'mean_ms_per_call': round(mean, 5),
                'evals_per_s_median_block': round(1000.0 / med, 1)}

    def stage_breakdown(fn, steps, grad):
        """Per-stage HIP events inside each call (a separate pass: the events add their own gaps)."""
        ta.stage_timer(True)
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        stage_ms, ncalls = ta.stage_times()
        ta.stage_timer(False)
        sb = stage_bytes(wl.nxy, len(wl.u), len(wl.intensity), rmax, grad=grad)
        stages = {}
        for k in ta.STAGES:
            avg = stage_ms[k] / max(ncalls, 1)
            if avg == 0.0:
                continue
            b = sb.get(k)
            stages[k] = {'ms': round(avg, 5), 'alg_bytes': b,
                         'GBps': round(b / (avg * 1e-3) / 1e9, 1) if b else None}
        return stages, sum(sb.values())

    def kernel_launch_ms(fn, stage, reps=20, calls=5):
        """Average launch duration of one forward kernel: `reps` back-to-back launches inside the
        stage's event interval, `calls` calls."""
        ta.stage_timer(True)
        ta.stage_repeat(stage, reps)
        for _ in range(calls):
            fn()
        torch.cuda.synchronize()
        stage_ms, ncalls = ta.stage_times()
        ta.stage_repeat(None, 1)
        ta.stage_timer(False)
        return stage_ms[stage] / (ncalls * reps)

    ms_per_step, ms_std, ms_min, _, nsteps = run_timed(step, args.steps)
    evals = 1000.0 / ms_per_step
    steady = steady_state(step)
    stages, b_alg = stage_breakdown(step, min(args.steps, 50), grad=False)
    kernels = {}
    sb = stage_bytes(wl.nxy, len(wl.u), len(wl.intensity), rmax)
    for k in ('row_pass', 'column_pass', 'sample'):
        kms = kernel_launch_ms(step, k)
        gbps = sb[k] / (kms * 1e-3) / 1e9
        kernels[k] = {'avg_launch_ms': round(kms, 5), 'alg_bytes': sb[k], 'GBps': round(gbps, 1),
                      'hbm_frac': round(gbps / HBM_PEAK_GBS, 4)}

    def step_grad():
        ds.chi2_profile_grad(I, wl.Rmin, wl.dR, wl.nxy, wl.dxy, check_central_pixel=False, **kw)

    for _ in range(args.warmup):
        step_grad()
    ms_grad, ms_grad_std, ms_grad_min, _, _ = run_timed(step_grad, args.steps)
    evals_grad = 1000.0 / ms_grad
    steady_grad = steady_state(step_grad)
    stages_grad, b_alg_grad = stage_breakdown(step_grad, min(args.steps, 50), grad=True)

    # the kernel that moves the most HBM bytes per evaluation (the sample kernel: 72 MB against the
    # column pass's 37 MB; the two take about the same time, and the column pass is bound by FP64
    # arithmetic and LDS, not HBM); every forward kernel's own fraction is in `kernels`
    dom = max(kernels, key=lambda k: kernels[k]['alg_bytes'])
    achieved = kernels[dom]['GBps']
    traffic = measured_traffic(KERNEL_NAMES[dom], wl.nxy, len(wl.u))
    C = wl.nxy // 2 + 1
    b_complex = 2 * 8 * wl.nxy * C + 13 * 8 * len(wl.u) + 8 * len(wl.intensity)   # complex-spectrum model

    batched = None
    if args.batch:
        batched = batched_rates(ta, ds, wl, I, chi2_dev, args.batch, args.batch_grad, args.steps, stream)

    # plain forward in FP32 on the same workload
    ds32 = ta.Dataset(wl.u, wl.v, wl.re, wl.im, wl.w, device=dev, dtype=torch.float32)
    I32 = I.to(torch.float32)

    def step32():
        ds32.chi2_profile(I32, wl.Rmin, wl.dR, wl.nxy, wl.dxy, check_central_pixel=False, out=out, **kw)

    for _ in range(args.warmup):
        step32()
    ms32, _, _, _, _ = run_timed(step32, args.steps)
    c32 = float(out.item())
    b32 = sum(stage_bytes(wl.nxy, len(wl.u), len(wl.intensity), rmax, s=4).values())
    plain_f32 = {'evals_per_s': round(1000.0 / ms32, 2), 'ms_per_step': round(ms32, 5),
                 'alg_bytes_per_eval': b32, 'frac': frac_of(b32, 1000.0 / ms32),
                 'steady_state': steady_state(step32),
                 'chi2_rel_err_vs_f64': abs(c32 - chi2_dev) / abs(chi2_dev)}
    del ds32

    extras = extra_configs(ta, g, stream, args.steps) if args.extras else None

    def read_rate_gbps(nbytes, grids=(1024, 2048, 4096), reps=20):
        """Best rate of a plain 16-byte-load stream over `nbytes` of device memory (the library's
        read-rate reference kernel), `reps` back-to-back launches per grid between two events."""
        buf = torch.zeros(nbytes // 16 * 2, dtype=torch.float64, device=dev)
        sums = torch.empty(max(grids), dtype=torch.float64, device=dev)
        best = 0.0
        for gr in grids:
            for _ in range(3):
                ta.read_probe(buf, gr, sums)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                ta.read_probe(buf, gr, sums)
            e1.record(stream)
            e1.synchronize()
            best = max(best, buf.numel() * 8 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9)
        del buf, sums
        return best

    # plain streams of the sample kernel's bytes and of 1 GiB: what reads of that size reach here
    read_same = read_rate_gbps(sb['sample'])
    read_large = read_rate_gbps(1 << 30, reps=5)

    # host-pointer (drop-in) form, PCIe uploads included: a few calls
    nh = 5
    t = time.perf_counter()
    for _ in range(nh):
        g.chi2Profile(*wl.args(), check_central_pixel=False, **kw)
    host_rate = nh / (time.perf_counter() - t)

    cpu = None
    if args.cpu_seconds > 0:
        from oracle import port
        ncores = min(16, os.cpu_count() or 1)
        port.threads(ncores)
        times = []
        tstart = time.perf_counter()
        while time.perf_counter() - tstart < args.cpu_seconds and len(times) < 400:
            t = time.perf_counter()
            c = port.chi2_profile(wl.intensity, wl.Rmin, wl.dR, wl.dxy, wl.nxy, wl.inc, wl.dRA, wl.dDec, wl.duv,
                                  wl.PA, wl.u, wl.v, wl.re, wl.im, wl.w, workers=ncores)
            times.append(time.perf_counter() - t)
        keep = times[1:] if len(times) > 1 else times
        # the same path on one thread, over a few evaluations (about 1 s each)
        port.threads(1)
        t1 = []
        tstart = time.perf_counter()
        while time.perf_counter() - tstart < min(4.0, args.cpu_seconds) and len(t1) < 4:
            t = time.perf_counter()
            port.chi2_profile(wl.intensity, wl.Rmin, wl.dR, wl.dxy, wl.nxy, wl.inc, wl.dRA, wl.dDec, wl.duv,
                              wl.PA, wl.u, wl.v, wl.re, wl.im, wl.w, workers=1)
            t1.append(time.perf_counter() - t)
        port.threads(ncores)
        cpu = {'value': round(1.0 / float(np.mean(keep)), 4), 'unit': 'evals/s', 'cores': ncores, 'kind': 'port',
               'sample': f'{len(times)} full-size evaluations (first dropped), C/OpenMP restatement of galario '
                         f'CPU loops + scipy.fft.rfft2(workers={ncores}); chi2 rel. diff vs GPU '
                         f'{abs(c - chi2_host) / abs(chi2_host):.1e}',
               'one_thread_evals_per_s': round(len(t1) / float(np.sum(t1)), 4) if t1 else None,
               'host': host_cpu()}

    res = {
        'metric': METRIC, 'value': round(evals, 2), 'unit': 'evals/s', 'n_gpus': 1, 'steps': nsteps,
        'warmup': args.warmup, 'ms_per_step': round(ms_per_step, 5), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f64', 'data': 'synthetic',
        'config': {'workload': f'chi2Profile forward, plain device-resident form (dataset uploaded once in '
                               f'locality order), nxy={wl.nxy}, Nvis={len(wl.u)}, Nr={len(wl.intensity)}, FP64',
                   'nxy': wl.nxy, 'nvis': len(wl.u), 'nr': len(wl.intensity), 'rmax': rmax,
                   'parallelism': 'single GPU'},
        'ms_per_step_std': round(ms_std, 5), 'ms_per_step_min': round(ms_min, 5),
        'steady_state': steady,
        'roofline': {'bound': 'hbm', 'kernel': dom, 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4), 'traffic': traffic,
                     'alg_bytes_per_launch': kernels[dom]['alg_bytes'],
                     'avg_launch_ms': kernels[dom]['avg_launch_ms'],
                     'stream_read_same_bytes_GBps': round(read_same, 1),
                     'stream_read_1GiB_GBps': round(read_large, 1),
                     'achieved_vs_stream_read_same_bytes': round(achieved / read_same, 4)},
        'kernels': kernels,
        'pipeline_roofline': {'alg_bytes_per_eval': b_alg,
                              'frac': round(b_alg * evals / (HBM_PEAK_GBS * 1e9), 4),
                              'complex_spectrum_model_bytes': b_complex,
                              'complex_model_frac': round(b_complex * evals / (HBM_PEAK_GBS * 1e9), 4)},
        'stages': stages,
        'forward_grad': {'evals_per_s': round(evals_grad, 2), 'ms_per_step': round(ms_grad, 5),
                         'ms_per_step_std': round(ms_grad_std, 5), 'ms_per_step_min': round(ms_grad_min, 5),
                         'steady_state': steady_grad,
                         'alg_bytes_per_eval': b_alg_grad,
                         'frac': round(b_alg_grad * evals_grad / (HBM_PEAK_GBS * 1e9), 4),
                         'stages': stages_grad},
        'plain_f32': plain_f32,
        'batched': batched,
        'other_configs': extras,
        'host_api_evals_per_s_incl_pcie': round(host_rate, 2),
        'dataset_upload_ms': round(upload_ms, 2),
        'cpu_baseline': cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == '__main__':
    main()
