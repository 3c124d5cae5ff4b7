"""GPU: the gradient-sampler caller (galario's quickstart model, docs/quickstart.rst:84-174, on the batched
chi2 + adjoint with a vectorised HMC).

* the posterior's autograd gradient equals central finite differences of the same posterior (relative 1e-5
  on a smooth direction) -- the chain rule through the profile parameterisation and the batched adjoint;
* leapfrog with a tiny step conserves the Hamiltonian (acceptance ~ 1), which only holds with a correct
  gradient;
* chains started away from the truth climb in log p and keep a healthy acceptance rate;
* the posterior and its gradient equal the same posterior composed on tests/torch_reference.py (torch.fft and
  autograd on the CPU, no galario_amd code), relative 1e-7."""
import math

import numpy as np
import pytest

from oracle import galario_oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')
DEV = torch.device('cuda:0')


def quickstart_problem(seed=1, nvis=4000, nxy=256):
    from galario_amd import sampler as S
    from galario_amd import torch_api as ta
    rng = np.random.default_rng(seed)
    Rmin, dR, nR = 1e-4 * S.ARCSEC, 0.01 * S.ARCSEC, 300
    dxy = 0.05 * S.ARCSEC
    duv = 1 / (nxy * dxy)
    b = np.exp(rng.uniform(np.log(2 * duv), np.log(0.4 * nxy * duv), nvis))
    th = rng.uniform(0, 2 * np.pi, nvis)
    u, v = b * np.cos(th), b * np.sin(th)
    truth = np.array([10.0, 0.4, 50.0, 30.0, 0.05, -0.05])
    R = Rmin + dR * np.arange(nR)
    I = 10 ** truth[0] * np.exp(-0.5 * (R / (truth[1] * S.ARCSEC)) ** 2)
    vis = O.sample_profile(I, Rmin, dR, dxy, nxy, truth[2] * S.DEG, truth[4] * S.ARCSEC, truth[5] * S.ARCSEC, duv,
                           truth[3] * S.DEG, u, v, check_central=False)
    sig = 0.05 * np.abs(vis).max()
    re, im = vis.real + sig * rng.standard_normal(nvis), vis.imag + sig * rng.standard_normal(nvis)
    w = np.full(nvis, 1 / sig ** 2)
    ds = ta.Dataset(u, v, re, im, w, device=DEV)
    return ds, truth, (Rmin, dR, nR, nxy, dxy)


def test_posterior_gradient_vs_finite_differences():
    from galario_amd import sampler as S
    from galario_amd import torch_api as ta
    ds, truth, (Rmin, dR, nR, nxy, dxy) = quickstart_problem()
    W = 4
    plan = ta.BatchedChi2Profile([(ds, nxy, dxy)], nR, Rmin, dR, max_walkers=W, check_central_pixel=False)
    post = S.QuickstartPosterior(plan, [[1, 20], [0, 8], [0, 90], [0, 180], [-2, 2], [-2, 2]], Rmin, dR, nR)
    rng = np.random.default_rng(5)
    p = torch.tensor(truth[None, :] + 0.02 * rng.standard_normal((W, 6)) * np.array([1, 1, 10, 10, 1, 1]),
                     device=DEV).requires_grad_(True)
    lp = post(p)
    g, = torch.autograd.grad(lp.sum(), p)
    d = torch.tensor(rng.standard_normal((W, 6)), device=DEV) * torch.tensor([1e-3, 1e-3, 1e-1, 1e-1, 1e-3, 1e-3],
                                                                              device=DEV)
    h = 1e-4
    with torch.no_grad():
        fd = (post(p + h * d) - post(p - h * d)) / (2 * h)
    an = (g * d).sum(dim=1)
    rel = ((fd - an).abs() / an.abs()).max().item()
    print(f"\nposterior directional derivatives: max rel diff {rel:.2e}")
    assert rel < 1e-5
    plan.close()


def test_hmc_tiny_step_conserves_energy_and_chains_climb():
    from galario_amd import sampler as S
    from galario_amd import torch_api as ta
    ds, truth, (Rmin, dR, nR, nxy, dxy) = quickstart_problem()
    W = 16
    plan = ta.BatchedChi2Profile([(ds, nxy, dxy)], nR, Rmin, dR, max_walkers=W, check_central_pixel=False)
    post = S.QuickstartPosterior(plan, [[1, 20], [0, 8], [0, 90], [0, 180], [-2, 2], [-2, 2]], Rmin, dR, nR)
    gen = torch.Generator(device=DEV)
    gen.manual_seed(3)
    rng = np.random.default_rng(7)
    x0 = torch.tensor(truth[None, :] + rng.standard_normal((W, 6)) * np.array([0.01, 0.01, 1, 1, 0.005, 0.005]),
                      device=DEV)
    scale = torch.tensor([1e-3, 1e-3, 0.05, 0.05, 1e-3, 1e-3], device=DEV, dtype=torch.float64)
    _, acc_tiny, _ = S.hmc(post, x0, n_steps=3, step_size=1e-3 * scale, n_leapfrog=5, generator=gen)
    assert acc_tiny.min().item() > 0.99
    lp0 = post(x0).detach()
    samples, acc, lp = S.hmc(post, x0, n_steps=20, step_size=0.3 * scale, n_leapfrog=8, generator=gen)
    print(f"\nacceptance {acc.mean().item():.2f}; mean log p {lp0.mean().item():.1f} -> {lp.mean().item():.1f}")
    assert torch.isfinite(lp).all()
    assert acc.mean().item() > 0.3
    assert lp.mean().item() >= lp0.mean().item() - 5.0
    plan.close()


def test_posterior_gradient_vs_torch_reference():
    """QuickstartPosterior's autograd gradient against the same posterior built on tests/torch_reference.py
    (CPU float64 autograd through torch.fft, no galario_amd code): galario's GaussianProfile
    (docs/quickstart.rst:87-93, radial grid by np.linspace(..., endpoint=False)), the unit conversions of
    lnpostfn (:156-166) and the prior's Jacobian -p[0] (:188).  Log-posterior and its gradient with
    respect to all six parameters of every chain, 1e-7 relative."""
    from galario_amd import sampler as S
    from galario_amd import torch_api as ta
    import torch_reference as TR
    ds, truth, (Rmin, dR, nR, nxy, dxy) = quickstart_problem()
    W = 4
    plan = ta.BatchedChi2Profile([(ds, nxy, dxy)], nR, Rmin, dR, max_walkers=W, check_central_pixel=False)
    post = S.QuickstartPosterior(plan, [[1, 20], [0, 8], [0, 90], [0, 180], [-2, 2], [-2, 2]], Rmin, dR, nR)
    rng = np.random.default_rng(8)
    p0 = truth[None, :] + 0.02 * rng.standard_normal((W, 6)) * np.array([1, 1, 10, 10, 1, 1])
    p = torch.tensor(p0, device=DEV).requires_grad_(True)
    lp = post(p)
    g, = torch.autograd.grad(lp.sum(), p)
    lp, g = lp.detach().cpu().numpy(), g.cpu().numpy()
    # the observation in the caller's order (the Dataset keeps it sorted on the device)
    order = ds.order.long().cpu().numpy()
    obs = {}
    for k in ('u', 'v', 're', 'im', 'w'):
        a = np.empty(ds.n)
        a[order] = getattr(ds, k).cpu().numpy()
        obs[k] = torch.tensor(a, dtype=torch.float64)
    R = torch.tensor(np.linspace(Rmin, Rmin + dR * nR, nR, endpoint=False), dtype=torch.float64)
    for c in range(W):
        q = torch.tensor(p0[c], dtype=torch.float64).requires_grad_(True)
        f = 10.0 ** q[0] * torch.exp(-0.5 * (R / (q[1] * S.ARCSEC)) ** 2)
        geom = torch.stack([q[2] * S.DEG, q[3] * S.DEG, q[4] * S.ARCSEC, q[5] * S.ARCSEC])
        chi2 = TR.chi2_profile(f, geom, Rmin, dR, nxy, dxy, obs['u'], obs['v'], obs['re'], obs['im'], obs['w'])
        ref = -0.5 * chi2 - q[0]
        ref.backward()
        print(f"\nchain {c}: log p {lp[c]:.10e} ref {float(ref):.10e}; max rel grad diff "
              f"{np.max(np.abs(g[c] - q.grad.numpy()) / np.abs(q.grad.numpy())):.1e}")
        assert math.isclose(lp[c], float(ref), rel_tol=1e-7)
        np.testing.assert_allclose(g[c], q.grad.numpy(), rtol=1e-7)
    plan.close()
