"""GPU tests of the uv extent.

A spectrum built only within a dataset's largest baseline (the Dataset's own
calls and batched plans) gives the same numbers as the full spectrum (uvmax
= 0), at extents from a small part of the grid to nearly all of it.  A point
beyond a declared extent is reported like an out-of-grid point: NaN and the
flag, never a value read from cells that were not stored."""
import math

import numpy as np
import pytest

from oracle import galario_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')
arcsec = O.ARCSEC
DEV = 'cuda:0'


def workload(frac, nvis=40_000, nxy=1024, seed=3):
    """Baselines log-uniform up to `frac` of the half grid, a ring profile, noisy data."""
    rng = np.random.default_rng(seed)
    Rmin, dR, nr, dxy = 1e-4 * arcsec, 0.002 * arcsec, 600, 0.01 * arcsec
    R = Rmin + dR * np.arange(nr)
    I = 1e10 * (np.exp(-((R - 0.3 * arcsec) / (0.1 * arcsec)) ** 2) + 0.2 * np.exp(-R / (0.5 * arcsec)))
    duv = 1 / (nxy * dxy)
    b = np.exp(rng.uniform(np.log(duv), np.log(frac * (nxy // 2) * duv), nvis))
    th = rng.uniform(0, 2 * np.pi, nvis)
    u, v = b * np.cos(th), b * np.sin(th)
    re, im = rng.standard_normal(nvis), rng.standard_normal(nvis)
    w = rng.uniform(0.5, 2.0, nvis)
    geom = dict(inc=0.7, PA=2.1, dRA=0.02 * arcsec, dDec=-0.01 * arcsec)
    return dict(I=I, Rmin=Rmin, dR=dR, nr=nr, nxy=nxy, dxy=dxy, u=u, v=v, re=re, im=im, w=w, geom=geom)


def _dataset(d, dtype):
    from galario_amd import torch_api as ta
    return ta.Dataset(d['u'], d['v'], d['re'], d['im'], d['w'], device=DEV, dtype=dtype)


@pytest.mark.parametrize("frac", [0.15, 0.6, 0.97])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=['DP', 'SP'])
def test_extent_matches_full_spectrum(frac, dtype):
    from galario_amd import torch_api as ta
    d = workload(frac)
    ds = _dataset(d, dtype)
    assert 0 < ds.uvmax <= frac * (d['nxy'] // 2) / (d['nxy'] * d['dxy']) * (1 + 1e-6)
    I = torch.tensor(d['I'], dtype=dtype, device=DEV)
    grid = (I, d['Rmin'], d['dR'], d['nxy'], d['dxy'])
    kw = dict(check_central_pixel=False, **d['geom'])
    ta.take_oob_flag()
    # chi2 and visibilities: the stored cells are the same numbers, so the results are bit-identical
    c_cut = ds.chi2_profile(*grid, **kw).item()
    c_full = ta.chi2_profile(*grid, ds.u, ds.v, ds.re, ds.im, ds.w, uvmax=0., **kw).item()
    assert c_cut == c_full
    v_cut = ds.sample_profile(*grid, **kw)
    v_full = ta.sample_profile(*grid, ds.u, ds.v, perm=ds.order, uvmax=0., **kw)
    assert torch.equal(v_cut, v_full)
    src = [[1.0, 0.0, 0.0], [0.3, 0.05 * arcsec, -0.02 * arcsec]]
    kws = dict(check_central_pixel=False, inc=d['geom']['inc'], PA=d['geom']['PA'])
    m_cut = ds.chi2_profile_multi(*grid, src, **kws).item()
    m_full = ta.chi2_profile_multi(*grid, ds.u, ds.v, ds.re, ds.im, ds.w, src, uvmax=0., **kws).item()
    assert m_cut == m_full
    # gradient: chi2 in the fixed order; dI and the geometry by atomics, so to the last bits
    g_cut = ds.chi2_profile_grad(*grid, **kw)
    g_full = ta.chi2_profile_grad(*grid, ds.u, ds.v, ds.re, ds.im, ds.w, uvmax=0., **kw)
    assert g_cut[0].item() == g_full[0].item() == c_cut
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    gI_cut, gI_full = g_cut[1].cpu().numpy(), g_full[1].cpu().numpy()
    assert np.linalg.norm(gI_cut - gI_full) <= tol * np.linalg.norm(gI_full)
    np.testing.assert_allclose(g_cut[2].cpu().numpy(), g_full[2].cpu().numpy(), rtol=tol, atol=0)
    assert not ta.take_oob_flag()


def test_point_beyond_declared_extent_is_reported():
    from galario_amd import torch_api as ta
    d = workload(0.5, nvis=5_000)
    ds = _dataset(d, torch.float64)
    I = torch.tensor(d['I'], dtype=torch.float64, device=DEV)
    grid = (I, d['Rmin'], d['dR'], d['nxy'], d['dxy'])
    kw = dict(check_central_pixel=False, **d['geom'])
    ta.take_oob_flag()
    c = ta.chi2_profile(*grid, ds.u, ds.v, ds.re, ds.im, ds.w, uvmax=0.5 * ds.uvmax, **kw).item()
    assert math.isnan(c)
    assert ta.take_oob_flag()
    vis = ta.sample_profile(*grid, ds.u, ds.v, uvmax=0.5 * ds.uvmax, **kw).cpu().numpy()
    r = np.hypot(ds.u.cpu().numpy(), ds.v.cpu().numpy())
    beyond = r > 0.5 * ds.uvmax * (1 + 1e-6)       # (the bound carries a relative slack of 1e-12)
    assert beyond.sum() > 100
    assert np.isnan(vis[beyond]).all() and np.isfinite(vis[r < 0.5 * ds.uvmax * (1 - 1e-6)]).all()
    assert ta.take_oob_flag()
    # the exact extent itself is inside
    c = ta.chi2_profile(*grid, ds.u, ds.v, ds.re, ds.im, ds.w, uvmax=ds.uvmax, **kw).item()
    assert math.isfinite(c) and not ta.take_oob_flag()


def test_batched_plan_with_extent_matches_full_spectrum():
    from galario_amd import torch_api as ta
    d = workload(0.3, nvis=30_000)
    ds = _dataset(d, torch.float64)
    W = 3
    rng = np.random.default_rng(5)
    Ib = np.repeat(d['I'][None, :], W, 0) * np.exp(0.1 * rng.standard_normal((W, 1)))
    g0 = np.array([d['geom']['inc'], d['geom']['PA'], d['geom']['dRA'], d['geom']['dDec']])
    gb = g0[None, :] + np.array([0.05, 0.2, 0.01 * arcsec, 0.01 * arcsec]) * rng.standard_normal((W, 4))
    It = torch.tensor(Ib, device=DEV)
    Gt = torch.tensor(gb, device=DEV)
    plan = ta.BatchedChi2Profile([(ds, d['nxy'], d['dxy'])], d['nr'], d['Rmin'], d['dR'], max_walkers=W,
                                 check_central_pixel=False)
    chi2 = plan(It, Gt).cpu().numpy()
    c_g, _, gI, gg = plan.grad(It, Gt)
    for k in range(W):
        kw = dict(inc=gb[k, 0], PA=gb[k, 1], dRA=gb[k, 2], dDec=gb[k, 3], check_central_pixel=False)
        grid = (It[k].contiguous(), d['Rmin'], d['dR'], d['nxy'], d['dxy'])
        full = ta.chi2_profile_grad(*grid, ds.u, ds.v, ds.re, ds.im, ds.w, uvmax=0., **kw)
        assert chi2[k] == full[0].item() == c_g[k].item()
        gI_full = full[1].cpu().numpy()
        assert np.linalg.norm(gI[k].cpu().numpy() - gI_full) <= 1e-12 * np.linalg.norm(gI_full)
        np.testing.assert_allclose(gg[k].cpu().numpy(), full[2].cpu().numpy(), rtol=1e-12, atol=0)
    assert not plan.take_oob_flag()
