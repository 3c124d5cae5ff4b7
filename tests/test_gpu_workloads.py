"""GPU: the benchmark suite's workloads at their full sizes.

* Differentiation (nxy=2048, dxy=0.01", Nvis=2e6, Nr=500, dR=0.004", seed 2,
  16 parameter points of the jittered ring+envelope ensemble), double:
  - every point against the torch.fft composition of galario's forward math
    (tests/torch_reference.py, CPU float64 autograd on the complex
    spectrum): chi2 1e-11, the profile gradient 1e-7 norm-wise, each geometry
    gradient 1e-8 relative;
  - every point against central finite differences of the GPU forward,
    relative 1e-6, with galario's kinks handled as in test_gpu_grad.py:
    points whose interpolated spectrum changes sign between the two
    evaluations are dropped, for PA also points whose grid cell changes; the
    profile direction is smooth; the step is 1e-6 relative for the profile
    and PA, 1e-5 for dRA and dDec (smooth: they only move each point's
    phase, and the larger step keeps the two chi2 sums' rounding well below
    their difference), 1e-8 for inc.  inc moves every pixel radius: a pixel
    crossing a knot of the
    profile's linear interpolation (the jittered profiles have large slope
    jumps there) gives a central difference an O(h) error, and one crossing
    the last bin (galario's sweep drops pixels beyond it,
    src/galario.cpp:1063-1065) an O(1/h) error.  So the point's inclination
    is nudged (by at most 4e-7 relative) to a value where no pixel of the
    swept box changes its profile bin between inc - h and inc + h, and every
    difference and the analytic gradient are taken there;
  - the float path against the double path at every point: chi2 1e-5, the
    profile gradient 1e-4 norm-wise, the geometry gradients 1e-2.
* Batched forward: 256 walkers (Nr=300) x datasets (512, 1e4), (1024, 2e5),
  (2048, 1e6), double: every (walker, dataset) chi2 against the single call
  on the same dataset, 1e-12; two walkers on the (2048, 1e6) dataset against
  the oracle, 1e-12.
* Batched forward+gradient: 64 sets (Nr=500) x three datasets on one
  nxy=4096 grid (Nvis 5e5, 1.5e6, 3e6):
  - double: every set's chi2 per dataset (1e-12) and its gradients summed
    over the datasets (profile 1e-9 norm-wise, geometry 1e-9) against the
    single calls; one set's chi2 on each dataset against the oracle (1e-12);
  - FP32 bulk with FP64 geometry, phase and per-point arithmetic, and the
    pure float form, against the double single calls: the error of each is
    printed and bounded (mixed: chi2 1e-6, profile gradient 1e-4,
    geometry gradients 1e-3; pure float: chi2 1e-5, profile gradient 1e-3,
    geometry gradients 1e-2).
"""
import math

import numpy as np
import pytest

from galario_amd import synthetic as syn
from oracle import galario_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')
import torch_reference as TR  # noqa: E402

DEV = torch.device('cuda:0')
A = syn.ARCSEC


def oracle_model(wl):
    return O.sample_profile(wl.intensity, wl.Rmin, wl.dR, wl.dxy, wl.nxy, wl.inc, wl.dRA, wl.dDec, wl.duv, wl.PA,
                            wl.u, wl.v, check_central=False)


def gpu_model(wl):
    import galario_amd
    return galario_amd.double.sampleProfile(wl.intensity, wl.Rmin, wl.dR, wl.nxy, wl.dxy, wl.u, wl.v,
                                            check_central_pixel=False, **wl.kwargs())


def kw(g):
    return dict(inc=float(g[0]), PA=float(g[1]), dRA=float(g[2]), dDec=float(g[3]), check_central_pixel=False)


# ------------------------------------------------------------------ differentiation, nxy=2048, Nvis=2e6
@pytest.fixture(scope='module')
def diff_workload():
    from galario_amd import torch_api as ta
    wl = syn.ring_profile_workload(oracle_model, nvis=2_000_000, nxy=2048, nr=500, seed=2, dR=0.004 * A)
    I16, G16 = syn.walker_ensemble(np.random.default_rng(2), 16, 500, wl.Rmin, wl.dR, wl.inc, wl.PA, wl.dRA,
                                   wl.dDec)
    ds = ta.Dataset(wl.u, wl.v, wl.re, wl.im, wl.w, device=DEV)
    ds_raw = ta.Dataset(wl.u, wl.v, wl.re, wl.im, wl.w, device=DEV, sort=False)
    return wl, I16, G16, ds, ds_raw


def grad_at(ds, wl, I, g, dtype=torch.float64):
    c, gI, gg = ds.chi2_profile_grad(torch.tensor(I, dtype=dtype, device=DEV), wl.Rmin, wl.dR, wl.nxy, wl.dxy,
                                     **kw(g))
    return c.item(), gI.cpu().numpy(), gg.cpu().numpy()


def test_differentiation_workload_vs_torch_reference(diff_workload):
    wl, I16, G16, ds, _ = diff_workload
    T = lambda a: torch.tensor(a, dtype=torch.float64)  # noqa: E731
    u, v, re, im, w = (T(a) for a in (wl.u, wl.v, wl.re, wl.im, wl.w))
    worst = [0.0, 0.0, 0.0]
    for k in range(16):
        c, gI, gg = grad_at(ds, wl, I16[k], G16[k])
        I = T(I16[k]).requires_grad_(True)
        g = T(G16[k]).requires_grad_(True)
        cr = TR.chi2_profile(I, g, wl.Rmin, wl.dR, wl.nxy, wl.dxy, u, v, re, im, w)
        cr.backward()
        c_ref, gI_ref, gg_ref = float(cr.detach()), I.grad.numpy(), g.grad.numpy()
        e = [abs(c - c_ref) / abs(c_ref), np.linalg.norm(gI - gI_ref) / np.linalg.norm(gI_ref),
             float(np.max(np.abs(gg - gg_ref) / np.abs(gg_ref)))]
        worst = [max(a, b) for a, b in zip(worst, e)]
        assert e[0] <= 1e-11, (k, e)
        assert e[1] <= 1e-7, (k, e)
        for j in range(4):
            assert math.isclose(gg[j], gg_ref[j], rel_tol=1e-8), (k, j, gg[j], gg_ref[j])
    print(f"\n16 points vs torch reference: chi2 {worst[0]:.1e}, dI {worst[1]:.1e}, geometry {worst[2]:.1e}")


def _signs(ds_raw, wl, I, g):
    """sign of the real interpolated spectrum at each point (caller's order): V e^{-i theta} = sgn(Z) A."""
    vis = ds_raw.sample_profile(torch.tensor(I, device=DEV), wl.Rmin, wl.dR, wl.nxy, wl.dxy, **kw(g)).cpu().numpy()
    theta = 2 * np.pi * (wl.u * g[2] + wl.v * g[3])
    return np.sign((vis * np.exp(-1j * theta)).real)


def _cells(wl, PA):
    urot = wl.u * np.cos(PA) - wl.v * np.sin(PA)
    vrot = wl.u * np.sin(PA) + wl.v * np.cos(PA)
    s = np.copysign(1.0, urot)
    return np.stack([np.floor(np.abs(urot) / wl.duv), np.floor(wl.nxy // 2 + s * vrot / wl.duv)])


def _box_bins(wl, inc):
    """Profile bin of every pixel of the swept box at inclination inc (galario.cpp:1048-1071: r from
    x / cos(inc) and y, bin floor((r - Rmin) / dR) clamped at 0)."""
    rmax = O.sweep_rmax(len(wl.intensity), wl.Rmin, wl.dR, wl.nxy, wl.dxy)
    c = np.arange(-rmax + 1, rmax + 1) * wl.dxy
    r = np.sqrt((c[None, :] / np.cos(inc)) ** 2 + c[:, None] ** 2)
    return np.maximum(np.floor((r - wl.Rmin) / wl.dR), 0)


def _fd_cases(wl, I, g):
    """The difference cases at parameter point (I, g); g[0] (inc) is nudged in place first (see above)."""
    h0 = 1e-8 * abs(g[0])
    for k in range(40):
        inc = g[0] * (1 + 1e-8 * k)
        if np.array_equal(_box_bins(wl, inc + h0), _box_bins(wl, inc - h0)):
            break
    else:
        raise AssertionError('no inclination within 4e-7 relative without a pixel crossing a profile knot')
    g[0] = inc
    R = np.arange(len(I))
    dirI = I * (np.sin(R / 37.0) + 0.5 * np.cos(R / 91.0))
    h = 1e-6
    cases = [('profile', I + h * dirI, I - h * dirI, g, g, 2 * h)]
    # inc: 1e-8 (the nudge above keeps every pixel's bin); PA: 1e-6 (points changing cell are dropped);
    # dRA, dDec: 1e-5 -- they enter only through the phase 2 pi (u dRA + v dDec) of each point, so chi2 is
    # smooth in them (truncation error ~ (2 pi u h)^2 / 6 < 1e-8 relative) and the larger step keeps the
    # rounding of the two chi2 sums (~1e-16 chi2 each) well below the difference
    for j in range(4):
        hj = {0: 1e-8, 1: 1e-6}.get(j, 1e-5) * abs(g[j])
        gp, gm = g.copy(), g.copy()
        gp[j] += hj
        gm[j] -= hj
        cases.append((f'geom[{j}]', I, I, gp, gm, 2 * hj))
    return dirI, cases


@pytest.mark.parametrize('k', range(16))
def test_differentiation_workload_vs_finite_differences(diff_workload, k):
    from galario_amd import torch_api as ta
    wl, I16, G16, _, ds_raw = diff_workload
    I, g = I16[k].copy(), G16[k].copy()
    dirI, cases = _fd_cases(wl, I, g)
    keep = np.ones(len(wl.u), bool)
    for _, Ip, Im, gp, gm, _ in cases:
        keep &= _signs(ds_raw, wl, Ip, gp) == _signs(ds_raw, wl, Im, gm)
    nsign = int((~keep).sum())
    keep &= np.all(_cells(wl, cases[2][3][1]) == _cells(wl, cases[2][4][1]), axis=0)
    assert keep.mean() > 0.99
    ds = ta.Dataset(wl.u[keep], wl.v[keep], wl.re[keep], wl.im[keep], wl.w[keep], device=DEV)
    c, gI, gg = grad_at(ds, wl, I, g)
    analytic = [float(gI @ dirI)] + list(gg)
    errs = []
    for (name, Ip, Im, gp, gm, den), an in zip(cases, analytic):
        cp = ds.chi2_profile(torch.tensor(Ip, device=DEV), wl.Rmin, wl.dR, wl.nxy, wl.dxy, **kw(gp)).item()
        cm = ds.chi2_profile(torch.tensor(Im, device=DEV), wl.Rmin, wl.dR, wl.nxy, wl.dxy, **kw(gm)).item()
        fd = (cp - cm) / den
        errs.append((name, abs(fd - an) / abs(an), abs(cp - cm) / abs(c)))
    print(f"\npoint {k}: {nsign} sign changes, {int((~keep).sum()) - nsign} cell changes dropped; rel errors "
          + ", ".join(f"{n} {e:.1e} (|dchi2|/chi2 {d:.0e})" for n, e, d in errs))
    bad = [(n, e) for n, e, _ in errs if not e <= 1e-6]
    assert not bad, bad


def test_differentiation_workload_float_vs_double(diff_workload):
    from galario_amd import torch_api as ta
    wl, I16, G16, ds, _ = diff_workload
    ds32 = ta.Dataset(wl.u, wl.v, wl.re, wl.im, wl.w, device=DEV, dtype=torch.float32)
    worst = [0.0, 0.0, 0.0]
    for k in range(16):
        c64, gI64, gg64 = grad_at(ds, wl, I16[k], G16[k])
        c32, gI32, gg32 = grad_at(ds32, wl, I16[k], G16[k], dtype=torch.float32)
        e = [abs(c32 - c64) / abs(c64), np.linalg.norm(gI32 - gI64) / np.linalg.norm(gI64),
             float(np.max(np.abs(gg32 - gg64) / np.abs(gg64)))]
        worst = [max(a, b) for a, b in zip(worst, e)]
    print(f"\nfloat vs double over 16 points: chi2 {worst[0]:.1e}, dI {worst[1]:.1e}, geometry {worst[2]:.1e}")
    assert worst[0] <= 1e-5 and worst[1] <= 1e-4 and worst[2] <= 1e-2, worst


# ------------------------------------------------------------------ batched forward, 256 walkers x 3 grids
def test_batched_256_walkers_uneven_datasets():
    from galario_amd import torch_api as ta
    bw = syn.walkers_uneven_grids_workload(gpu_model)
    sets = [(ta.Dataset(d['u'], d['v'], d['re'], d['im'], d['w'], device=DEV), d['nxy'], d['dxy']) for d in bw.sets]
    It, Gt = torch.tensor(bw.intensity, device=DEV), torch.tensor(bw.geom, device=DEV)
    plan = ta.BatchedChi2Profile(sets, bw.nr, bw.Rmin, bw.dR, max_walkers=256, check_central_pixel=False)
    tot, per = plan(It, Gt, return_sets=True)
    per = per.cpu().numpy()
    assert not plan.take_oob_flag()
    ref = np.empty_like(per)
    for w in range(256):
        for s, (ds, nxy, dxy) in enumerate(sets):
            ref[w, s] = ds.chi2_profile(It[w], bw.Rmin, bw.dR, nxy, dxy, **kw(bw.geom[w])).item()
    rel = np.abs(per - ref) / np.abs(ref)
    print(f"\n256 walkers x 3 datasets: max rel diff vs single calls {rel.max():.1e}; chunk, walkers per "
          f"sample = {plan.info(256)[:2]}")
    assert rel.max() <= 1e-12
    np.testing.assert_allclose(tot.cpu().numpy(), per.sum(axis=1), rtol=1e-15)
    d = bw.sets[2]
    for w in (0, 137):
        g = bw.geom[w]
        c_or = O.chi2_profile(bw.intensity[w], bw.Rmin, bw.dR, d['dxy'], d['nxy'], g[0], g[2], g[3],
                              1 / (d['nxy'] * d['dxy']), g[1], d['u'], d['v'], d['re'], d['im'], d['w'],
                              check_central=False)
        assert math.isclose(per[w, 2], c_or, rel_tol=1e-12), (w, per[w, 2], c_or)
    plan.close()


# ------------------------------------------------------------------ batched forward+gradient, 64 sets x nxy=4096
@pytest.fixture(scope='module')
def grad_batch_workload():
    return syn.sets_three_datasets_4096_workload(gpu_model)


def _plain_f64(sets, bw, w):
    r = [ds.chi2_profile_grad(torch.tensor(bw.intensity[w], device=DEV), bw.Rmin, bw.dR, nxy, dxy,
                              **kw(bw.geom[w])) for ds, nxy, dxy in sets]
    return (np.array([x[0].item() for x in r]), sum(x[1] for x in r).cpu().numpy(),
            sum(x[2] for x in r).cpu().numpy())


def test_batched_64_sets_gradient_double(grad_batch_workload):
    from galario_amd import torch_api as ta
    bw = grad_batch_workload
    sets = [(ta.Dataset(d['u'], d['v'], d['re'], d['im'], d['w'], device=DEV), d['nxy'], d['dxy']) for d in bw.sets]
    It, Gt = torch.tensor(bw.intensity, device=DEV), torch.tensor(bw.geom, device=DEV)
    plan = ta.BatchedChi2Profile(sets, bw.nr, bw.Rmin, bw.dR, max_walkers=64, check_central_pixel=False)
    tot, per, gI, gg = (t.cpu().numpy() for t in plan.grad(It, Gt))
    assert not plan.take_oob_flag()
    worst = [0.0, 0.0, 0.0]
    for w in range(64):
        c_ref, gI_ref, gg_ref = _plain_f64(sets, bw, w)
        e = [float(np.max(np.abs(per[w] - c_ref) / np.abs(c_ref))),
             np.linalg.norm(gI[w] - gI_ref) / np.linalg.norm(gI_ref),
             float(np.max(np.abs(gg[w] - gg_ref) / np.abs(gg_ref)))]
        worst = [max(a, b) for a, b in zip(worst, e)]
        assert e[0] <= 1e-12 and e[1] <= 1e-9 and e[2] <= 1e-9, (w, e)
    print(f"\n64 sets x 3 datasets (nxy=4096, 5e6 points), double vs single calls: chi2 {worst[0]:.1e}, "
          f"dI {worst[1]:.1e}, geometry {worst[2]:.1e}")
    np.testing.assert_allclose(tot, per.sum(axis=1), rtol=1e-15)
    g = bw.geom[5]
    for s, d in enumerate(bw.sets):
        c_or = O.chi2_profile(bw.intensity[5], bw.Rmin, bw.dR, d['dxy'], d['nxy'], g[0], g[2], g[3],
                              1 / (d['nxy'] * d['dxy']), g[1], d['u'], d['v'], d['re'], d['im'], d['w'],
                              check_central=False)
        assert math.isclose(per[5, s], c_or, rel_tol=1e-12), (s, per[5, s], c_or)
    plan.close()


@pytest.mark.parametrize('form', ['f32_phase_f64', 'f32'])
def test_batched_64_sets_gradient_float_forms(grad_batch_workload, form):
    from galario_amd import torch_api as ta
    bw = grad_batch_workload
    f32 = torch.float32
    sets64 = [(ta.Dataset(d['u'], d['v'], d['re'], d['im'], d['w'], device=DEV), d['nxy'], d['dxy'])
              for d in bw.sets]
    sets = [(ta.Dataset(d['u'], d['v'], d['re'], d['im'], d['w'], device=DEV, dtype=f32), d['nxy'], d['dxy'])
            for d in bw.sets]
    mixed = form == 'f32_phase_f64'
    plan = ta.BatchedChi2Profile(sets, bw.nr, bw.Rmin, bw.dR, max_walkers=64, check_central_pixel=False,
                                 phase_f64=mixed)
    It = torch.tensor(bw.intensity, device=DEV, dtype=f32)
    Gt = torch.tensor(bw.geom, device=DEV, dtype=torch.float64 if mixed else f32)
    tot, per, gI, gg = (t.cpu().numpy() for t in plan.grad(It, Gt))
    c_fwd = plan(It, Gt).cpu().numpy()
    assert np.array_equal(c_fwd, tot)
    worst = [0.0, 0.0, 0.0]
    for w in range(0, 64, 4):
        c_ref, gI_ref, gg_ref = _plain_f64(sets64, bw, w)
        e = [abs(tot[w] - c_ref.sum()) / abs(c_ref.sum()), np.linalg.norm(gI[w] - gI_ref) / np.linalg.norm(gI_ref),
             float(np.max(np.abs(gg[w] - gg_ref) / np.abs(gg_ref)))]
        worst = [max(a, b) for a, b in zip(worst, e)]
    print(f"\n{form} vs double single calls (16 of 64 sets): chi2 {worst[0]:.1e}, dI {worst[1]:.1e}, "
          f"geometry {worst[2]:.1e}")
    bound = (1e-6, 1e-4, 1e-3) if mixed else (1e-5, 1e-3, 1e-2)
    assert all(e <= b for e, b in zip(worst, bound)), (worst, bound)
    plan.close()
