"""Benchmark: chi2Profile evaluations per second on one MI355X.

Workload: ring+envelope radial profile (Nr=1000, Rmin=1e-4", dR=0.002"),
nxy=4096, dxy=0.01", inc=40 deg, PA=120 deg, dRA=dDec=0.02", Nvis=1e6
log-uniform baselines in [5e3, 1e7] wavelengths, noisy observation; float64;
galario's central-pixel guard off (it rejects dR=0.002" at dxy=0.01").
A step is one chi2Profile evaluation with all inputs resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S]

Prints one JSON line.  `value` is timed with HIP events on the kernels' stream
around K calls (no per-stage events in that loop).  `roofline` is measured
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


def stage_bytes(nxy, nvis, nr, rmax, grad=False, s=8):
    """Algorithmic HBM bytes per launch of each stage of the profile path (each
    needed array touched once).  The swept image is even in x and y, so its
    spectrum is real and even: the row pass writes Rs'+1 real rows
    (Rs' = min(rmax+1, nxy/2)), the column pass reads them and writes the
    (nxy/2+1) x (nxy/2+1) real half spectrum H, the sample kernel streams
    u, v, Re, Im, w and gathers 4 real corners per point.  With the gradient
    it also scatters 4 corner gradients (FP64 atomics) into the adjoint Hh;
    the adjoint column pass reads Hh, clears it and writes Rs'+1 adjoint rows;
    the adjoint row pass reads them (the profile gradient is a few KB)."""
    C = nxy // 2 + 1
    rs = min(rmax + 1, nxy // 2)
    rows = rs + 1
    half = C * (nxy // 2 + 1) * s
    b = {
        'row_pass': rows * C * s + nr * s,
        'column_pass': rows * C * s + half,
        'sample': 5 * s * nvis + 4 * s * nvis,
    }
    if grad:
        b['sample'] += 4 * 8 * nvis
        b['adjoint_column_pass'] = 2 * half + rows * C * s
        b['adjoint_row_pass'] = rows * C * s + 2 * nr * 8
    return b


KERNEL_NAMES = {'row_pass': 'k_rowpass_sym<double, 12>', 'column_pass': 'k_colpass_sym<double, 12>',
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
    out = {'workload': f'batched form: P parameter sets on the same dataset (nxy={wl.nxy}, Nvis={len(wl.u)}), '
                       f'one captured HIP graph per call', 'forward': {}, 'forward_grad': {}}

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
        out['forward'][str(P)] = {'evals_per_s': round(P * 1000.0 / ms, 1), 'ms_per_batch': round(ms, 4),
                                  'chunk': chunk, 'walkers_per_sample': kw, 'chi2_rel_diff_vs_plain': rel}
    for P in counts_grad:
        Ip, Gp = It[:P].contiguous(), Gt[:P].contiguous()
        reps = max(5, min(steps, int(1000 // P)))
        ms = rate(lambda: plan.grad(Ip, Gp), P, reps)
        chunk, _, _ = plan.info(P, grad=True)
        out['forward_grad'][str(P)] = {'evals_per_s': round(P * 1000.0 / ms, 1), 'ms_per_batch': round(ms, 4),
                                       'chunk': chunk}
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
        res[name] = {'fwd_grad_evals_per_s': round(1000.0 / ms, 1), 'ms_per_eval': round(ms, 4)}
        del ds
    rel = max(float(torch.linalg.norm(grads['f32'][k][1] - grads['f64'][k][1]) / torch.linalg.norm(grads['f64'][k][1]))
              for k in range(16))
    res['f32_grad_I_rel_err_vs_f64'] = rel
    out['differentiation_nxy2048_nvis2e6'] = res

    # batched forward: 256 walkers x datasets (512, 1e4), (1024, 2e5), (2048, 1e6); Nr=300
    rng = np.random.default_rng(3)
    Rmin, dR, dxy = 1e-4 * A, 2.0 * A / 300, 0.03 * A
    base = syn.ProfileWorkload(intensity=syn.ring_envelope_profile(300, Rmin, dR), Rmin=Rmin, dR=dR, nxy=0,
                               dxy=dxy, inc=40 * syn.DEG, PA=120 * syn.DEG, dRA=0.02 * A, dDec=0.02 * A, u=None, v=None)
    sets = []
    for nxy, nvis in ((512, 10_000), (1024, 200_000), (2048, 1_000_000)):
        base.nxy = nxy
        base.u, base.v = syn.log_uniform_uv(rng, nvis, 3e4, 3e6)
        re, im, w = syn.noisy_observation(rng, model(base))
        sets.append((ta.Dataset(base.u, base.v, re, im, w, device=dev, dtype=f64), nxy, dxy))
    I256, G256 = syn.walker_ensemble(rng, 256, 300, Rmin, dR, base.inc, base.PA, base.dRA, base.dDec)
    plan = ta.BatchedChi2Profile(sets, 300, Rmin, dR, max_walkers=256, check_central_pixel=False)
    It, Gt = torch.from_numpy(I256).to(dev), torch.from_numpy(G256).to(dev)
    ms = timeit(lambda: plan(It, Gt), 10)
    chunk, kw, _ = plan.info(256)
    per = plan(It, Gt, return_sets=True)[1][0].cpu().numpy()
    ref = [ds.chi2_profile(It[0], Rmin, dR, nxy, dx, check_central_pixel=False, inc=float(G256[0, 0]),
                           PA=float(G256[0, 1]), dRA=float(G256[0, 2]), dDec=float(G256[0, 3])).item()
           for ds, nxy, dx in sets]
    out['batched_256_walkers_3_datasets_fwd'] = {
        'workload': '256 walkers (Nr=300) x datasets (nxy, Nvis) = (512, 1e4), (1024, 2e5), (2048, 1e6), FP64, '
                    'per-walker chi2 summed over the datasets',
        'walker_evals_per_s': round(256 * 1000.0 / ms, 1), 'ms_per_batch': round(ms, 4), 'chunk': chunk,
        'walkers_per_sample': kw,
        'chi2_rel_diff_vs_plain': float(np.max(np.abs(per - ref) / np.abs(ref)))}
    plan.close()
    del sets

    # batched forward+gradient: 64 sets x 3 datasets at nxy=4096 (Nvis 5e5, 1.5e6, 3e6), Nr=500
    rng = np.random.default_rng(4)
    Rmin, dR, dxy, nxy = 1e-4 * A, 0.004 * A, 0.01 * A, 4096
    base = syn.ProfileWorkload(intensity=syn.ring_envelope_profile(500, Rmin, dR), Rmin=Rmin, dR=dR, nxy=nxy,
                               dxy=dxy, inc=40 * syn.DEG, PA=120 * syn.DEG, dRA=0.02 * A, dDec=0.02 * A, u=None, v=None)
    raw = []
    for nvis in (500_000, 1_500_000, 3_000_000):
        base.u, base.v = syn.log_uniform_uv(rng, nvis, 5e3, 1e7)
        raw.append((base.u, base.v) + tuple(syn.noisy_observation(rng, model(base))))
    I64, G64 = syn.walker_ensemble(rng, 64, 500, Rmin, dR, base.inc, base.PA, base.dRA, base.dDec)
    res = {'workload': '64 parameter sets (Nr=500) x 3 datasets at nxy=4096 (Nvis 5e5, 1.5e6, 3e6; 5e6 total), '
                       'chi2 + gradient per set'}
    ref = None
    for name, dt in (('f64', torch.float64), ('f32', torch.float32)):
        sets = [(ta.Dataset(*r, device=dev, dtype=dt), nxy, dxy) for r in raw]
        plan = ta.BatchedChi2Profile(sets, 500, Rmin, dR, max_walkers=64, check_central_pixel=False)
        It, Gt = torch.from_numpy(I64).to(dev, dt), torch.from_numpy(G64).to(dev, dt)
        ms = timeit(lambda: plan.grad(It, Gt), 5)
        c, _, gI, gg = plan.grad(It, Gt)
        chunk, _, _ = plan.info(64, grad=True)
        if ref is None:   # FP64 plain path, set 0
            sets64 = sets
            r = [ds.chi2_profile_grad(It[0], Rmin, dR, nxy, dxy, check_central_pixel=False, inc=float(G64[0, 0]),
                                      PA=float(G64[0, 1]), dRA=float(G64[0, 2]), dDec=float(G64[0, 3]))
                 for ds, _, _ in sets64]
            ref = (sum(x[0].item() for x in r), sum(x[1] for x in r), sum(x[2] for x in r))
        res[name] = {'fwd_grad_set_evals_per_s': round(64 * 1000.0 / ms, 1), 'ms_per_batch': round(ms, 4),
                     'chunk': chunk,
                     'chi2_rel_err_vs_f64_plain': abs(float(c[0]) - ref[0]) / abs(ref[0]),
                     'grad_I_rel_err_vs_f64_plain': float(torch.linalg.norm(gI[0] - ref[1]) / torch.linalg.norm(ref[1])),
                     'grad_geom_max_rel_err_vs_f64_plain': float(torch.max(torch.abs(gg[0] - ref[2]) / torch.abs(ref[2])))}
        plan.close()
        del sets, plan
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
        res[f'K{K}'] = {'evals_per_s': round(1000.0 / ms, 1), 'ms_per_eval': round(ms, 4)}
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
    out['hmc_quickstart_posterior_64_chains'] = {
        'workload': "galario's quickstart posterior (Gaussian profile, 6 parameters) on 2e5 visibilities, "
                    f'nxy={nxy}, Nr={nR}, FP64: {W} HMC chains x {n_steps} steps x {n_leap} leapfrog steps',
        'chain_gradient_evals_per_s': round(W * grads * 1000.0 / ms, 1), 'ms_per_hmc_call': round(ms, 3)}
    plan.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--cpu-seconds', type=float, default=12.0, help='CPU-baseline time budget (0: skip)')
    ap.add_argument('--nvis', type=int, default=1_000_000)
    ap.add_argument('--nxy', type=int, default=4096)
    ap.add_argument('--batch', type=str, default='1,8,64,256',
                    help='parameter-set counts of the batched form (comma list; empty: skip)')
    ap.add_argument('--batch-grad', type=str, default='1,8,64',
                    help='parameter-set counts of the batched forward+gradient')
    ap.add_argument('--extras', type=int, default=1, help='also time the other configurations (0: skip)')
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
        """`steps` calls between events on the kernels' stream, in `blocks` equal blocks:
        (mean ms per call over all steps, std and min over the blocks' per-call means)."""
        per = max(1, steps // blocks)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(blocks + 1)]
        torch.cuda.synchronize()
        ev[0].record(stream)
        for b in range(blocks):
            for _ in range(per):
                fn()
            ev[b + 1].record(stream)
        torch.cuda.synchronize()
        blk = np.array([ev[b].elapsed_time(ev[b + 1]) / per for b in range(blocks)])
        return float(ev[0].elapsed_time(ev[blocks]) / (per * blocks)), float(blk.std()), float(blk.min()), per * blocks

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

    ms_per_step, ms_std, ms_min, nsteps = run_timed(step, args.steps)
    evals = 1000.0 / ms_per_step
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
    ms_grad, ms_grad_std, ms_grad_min, _ = run_timed(step_grad, args.steps)
    evals_grad = 1000.0 / ms_grad
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
    ms32, _, _, _ = run_timed(step32, args.steps)
    c32 = float(out.item())
    plain_f32 = {'evals_per_s': round(1000.0 / ms32, 2), 'ms_per_step': round(ms32, 5),
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
