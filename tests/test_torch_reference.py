"""CPU: the torch composition used for reference gradients is pinned to the
oracle (forward value) and to central finite differences of the oracle
(gradient), on a small problem."""
import math

import numpy as np
import pytest

from oracle import galario_oracle as O

torch = pytest.importorskip('torch')
import torch_reference as TR  # noqa: E402

arcsec = O.ARCSEC


def small_problem(seed=3, n=1500, N=128):
    rng = np.random.default_rng(seed)
    Rmin, dR, nr, dxy = 1e-4 * arcsec, 0.004 * arcsec, 200, 0.02 * arcsec
    R = Rmin + dR * np.arange(nr)
    I = 1e10 * (np.exp(-((R - 0.3 * arcsec) / (0.12 * arcsec)) ** 2) + 0.3 * np.exp(-R / (0.3 * arcsec)))
    duv = 1 / (N * dxy)
    b = rng.uniform(1.5, 0.4 * N, n) * duv
    th = rng.uniform(0, 2 * np.pi, n)
    u, v = b * np.cos(th), b * np.sin(th)
    geom = np.array([0.6, 1.1, 0.03 * arcsec, -0.02 * arcsec])
    vis = O.sample_profile(I, Rmin, dR, dxy, N, *geom[[0]], geom[2], geom[3], duv, geom[1], u, v,
                           check_central=False)
    sig = np.median(np.abs(vis))
    re = vis.real + sig * rng.standard_normal(n)
    im = vis.imag + sig * rng.standard_normal(n)
    w = np.full(n, 1 / sig ** 2)
    return dict(I=I, Rmin=Rmin, dR=dR, N=N, dxy=dxy, u=u, v=v, re=re, im=im, w=w, geom=geom)


def oracle_chi2(d, I=None, geom=None):
    I = d['I'] if I is None else I
    g = d['geom'] if geom is None else geom
    duv = 1 / (d['N'] * d['dxy'])
    return O.chi2_profile(I, d['Rmin'], d['dR'], d['dxy'], d['N'], g[0], g[2], g[3], duv, g[1], d['u'], d['v'],
                          d['re'], d['im'], d['w'], check_central=False)


def torch_chi2(d, I, geom):
    T = lambda a: torch.tensor(a, dtype=torch.float64)  # noqa: E731
    return TR.chi2_profile(I, geom, d['Rmin'], d['dR'], d['N'], d['dxy'], T(d['u']), T(d['v']), T(d['re']),
                           T(d['im']), T(d['w']))


def test_torch_reference_forward_matches_oracle():
    d = small_problem()
    I = torch.tensor(d['I'], dtype=torch.float64)
    g = torch.tensor(d['geom'], dtype=torch.float64)
    assert math.isclose(float(torch_chi2(d, I, g)), oracle_chi2(d), rel_tol=1e-11)


def test_torch_reference_gradient_vs_oracle_finite_differences():
    d = small_problem()
    I = torch.tensor(d['I'], dtype=torch.float64, requires_grad=True)
    g = torch.tensor(d['geom'], dtype=torch.float64, requires_grad=True)
    torch_chi2(d, I, g).backward()
    gI, gg = I.grad.numpy(), g.grad.numpy()
    rng = np.random.default_rng(0)
    # directional derivative along a random profile direction
    dirI = rng.standard_normal(len(d['I'])) * d['I'].max()
    h = 1e-6
    fd = (oracle_chi2(d, I=d['I'] + h * dirI) - oracle_chi2(d, I=d['I'] - h * dirI)) / (2 * h)
    assert math.isclose(fd, float(gI @ dirI), rel_tol=1e-6)
    for k in range(4):
        hk = 1e-6 * max(abs(d['geom'][k]), 1e-7)
        gp, gm = d['geom'].copy(), d['geom'].copy()
        gp[k] += hk
        gm[k] -= hk
        fdk = (oracle_chi2(d, geom=gp) - oracle_chi2(d, geom=gm)) / (2 * hk)
        assert math.isclose(fdk, gg[k], rel_tol=1e-5), (k, fdk, gg[k])
