"""Gradient-based sampling on the batched chi2 and its adjoint.

galario's quickstart fits a Gaussian ring with an emcee ensemble whose
log-posterior calls chi2Profile once per walker and step
(docs/quickstart.rst:84-174).  Here the same model runs through
``torch_api.BatchedChi2Profile.autograd``: one batched launch evaluates
every chain's chi2 and its gradient, torch's autograd carries the gradient
through the profile parameterisation, and a vectorised Hamiltonian Monte
Carlo moves all chains at once.

* ``gaussian_profiles``: galario's ``GaussianProfile`` for W parameter sets;
* ``QuickstartPosterior``: the quickstart's 6-parameter log-posterior
  (log10 f0, sigma ["], inc [deg], PA [deg], dRA ["], dDec ["]) with its
  uniform box prior and log-space Jacobian, for W chains at once;
* ``hmc``: leapfrog HMC for W independent chains of any differentiable
  log-density.
"""
from __future__ import annotations

import math

import torch

ARCSEC = 4.84813681109536e-06
DEG = 0.017453292519943295


def gaussian_profiles(f0, sigma, Rmin: float, dR: float, nR: int):
    """f0 [W] (Jy/sr), sigma [W] (rad) -> [W, nR] profiles f0 exp(-(R/sigma)^2 / 2), R = Rmin + k dR."""
    R = Rmin + dR * torch.arange(nR, dtype=f0.dtype, device=f0.device)
    return f0[:, None] * torch.exp(-0.5 * (R[None, :] / sigma[:, None]) ** 2)


class QuickstartPosterior:
    """log p(params) = -chi2 / 2 + log prior for W chains (params: [W, 6] tensor on the plan's device).

    ``plan`` is a BatchedChi2Profile over the observation(s), built with the
    same (nR, Rmin, dR) in radians.  The prior is uniform on ``p_ranges``
    (6 pairs, in the parameters' units) with the Jacobian -log10 f0 of the
    quickstart's log-space normalisation; outside it log p = -inf."""

    def __init__(self, plan, p_ranges, Rmin: float, dR: float, nR: int):
        self.plan, self.Rmin, self.dR, self.nR = plan, Rmin, dR, nR
        self.lo = torch.tensor([r[0] for r in p_ranges], dtype=torch.float64, device=plan.device)
        self.hi = torch.tensor([r[1] for r in p_ranges], dtype=torch.float64, device=plan.device)

    def __call__(self, p):
        inside = ((p >= self.lo) & (p <= self.hi)).all(dim=1)
        f0 = 10.0 ** p[:, 0]
        I = gaussian_profiles(f0, p[:, 1] * ARCSEC, self.Rmin, self.dR, self.nR)
        geom = torch.stack([p[:, 2] * DEG, p[:, 3] * DEG, p[:, 4] * ARCSEC, p[:, 5] * ARCSEC], dim=1)
        chi2 = self.plan.autograd(I.to(self.plan.dtype), geom.to(self.plan.dtype))
        logp = -0.5 * chi2 - p[:, 0]
        return torch.where(inside, logp, torch.full_like(logp, -math.inf))


def hmc(logp_fn, x0, n_steps: int, step_size, n_leapfrog: int = 10, generator=None):
    """Hamiltonian Monte Carlo for W independent chains.

    x0: [W, D] start; step_size: scalar or [D] (a diagonal scale); unit
    masses.  Returns (samples [n_steps, W, D], acceptance rate per chain
    [W], log p of the last states [W]).  Proposals whose log p is not finite
    are rejected."""
    x = x0.detach().clone()
    eps = torch.as_tensor(step_size, dtype=x.dtype, device=x.device)

    def logp_grad(z):
        z = z.detach().requires_grad_(True)
        lp = logp_fn(z)
        ok = torch.isfinite(lp)
        g, = torch.autograd.grad(torch.where(ok, lp, torch.zeros_like(lp)).sum(), z)
        return lp.detach(), g.detach()

    lp, g = logp_grad(x)
    out = torch.empty((n_steps,) + tuple(x.shape), dtype=x.dtype, device=x.device)
    acc = torch.zeros(x.shape[0], dtype=x.dtype, device=x.device)
    for s in range(n_steps):
        p0 = torch.randn(x.shape, dtype=x.dtype, device=x.device, generator=generator)
        xn, pn, gn = x, p0 + 0.5 * eps * g, g
        for k in range(n_leapfrog):
            xn = xn + eps * pn
            lpn, gn = logp_grad(xn)
            if k < n_leapfrog - 1:
                pn = pn + eps * gn
        pn = pn + 0.5 * eps * gn
        h0 = -lp + 0.5 * (p0 ** 2).sum(dim=1)
        h1 = -lpn + 0.5 * (pn ** 2).sum(dim=1)
        u = torch.rand(x.shape[0], dtype=x.dtype, device=x.device, generator=generator)
        take = torch.isfinite(h1) & (torch.log(u) < (h0 - h1))
        x = torch.where(take[:, None], xn, x)
        lp = torch.where(take, lpn, lp)
        g = torch.where(take[:, None], gn, g)
        acc += take.to(x.dtype)
        out[s] = x
    return out, acc / n_steps, lp
