"""GPU: several sources sharing a profile, inc and PA at their own offsets (galario's multi-source recipe,
docs/tech-specs.rst:45-57).  Checked against the oracle's sum of per-source sampleProfile calls (linearity),
relative 1e-12 in double and 1e-5 in float; one source of flux 1 equals the single-source call bit for
bit; chi2 equals reduce_chi2 of the summed visibilities."""
import math

import numpy as np
import pytest

from oracle import galario_oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')
arcsec = O.ARCSEC
DEV = torch.device('cuda:0')


def problem(nvis=20000, nxy=512, seed=4):
    rng = np.random.default_rng(seed)
    Rmin, dR, nr, dxy = 1e-4 * arcsec, 0.002 * arcsec, 300, 0.01 * arcsec
    R = Rmin + dR * np.arange(nr)
    I = 1e10 * np.exp(-((R - 0.2 * arcsec) / (0.05 * arcsec)) ** 2)
    duv = 1 / (nxy * dxy)
    b = np.exp(rng.uniform(np.log(duv), np.log(0.45 * nxy * duv), nvis))
    th = rng.uniform(0, 2 * np.pi, nvis)
    u, v = b * np.cos(th), b * np.sin(th)
    src = np.array([[1.0, 0.3 * arcsec, -0.2 * arcsec], [0.5, -0.4 * arcsec, 0.1 * arcsec],
                    [0.25, 0.0, 0.0], [2.0, 0.05 * arcsec, 0.45 * arcsec]])
    re, im = rng.standard_normal(nvis), rng.standard_normal(nvis)
    w = rng.uniform(0.5, 2.0, nvis)
    return dict(I=I, Rmin=Rmin, dR=dR, nxy=nxy, dxy=dxy, duv=duv, u=u, v=v, re=re, im=im, w=w, src=src,
                inc=0.6, PA=1.2)


@pytest.mark.parametrize("dt, tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)], ids=['DP', 'SP'])
def test_multisource_equals_sum_of_single_sources(dt, tol):
    from galario_amd import torch_api as ta
    d = problem()
    ds = ta.Dataset(d['u'], d['v'], d['re'], d['im'], d['w'], device=DEV, dtype=dt)
    I = torch.tensor(d['I'], dtype=dt, device=DEV)
    vis = ds.sample_profile_multi(I, d['Rmin'], d['dR'], d['nxy'], d['dxy'], d['src'], PA=d['PA'], inc=d['inc'],
                                  check_central_pixel=False).cpu().numpy()
    ref = sum(f * O.sample_profile(d['I'], d['Rmin'], d['dR'], d['dxy'], d['nxy'], d['inc'], dra, ddec, d['duv'],
                                   d['PA'], d['u'], d['v'], check_central=False) for f, dra, ddec in d['src'])
    err = np.abs(vis - ref).max() / np.abs(ref).max()
    print(f"\nmulti-source vs sum of single sources ({dt}): {err:.2e}")
    assert err <= tol
    chi2 = ds.chi2_profile_multi(I, d['Rmin'], d['dR'], d['nxy'], d['dxy'], d['src'], PA=d['PA'], inc=d['inc'],
                                 check_central_pixel=False).item()
    chi2_ref = O.reduce_chi2(d['re'], d['im'], ref, d['w'])
    assert math.isclose(chi2, chi2_ref, rel_tol=max(tol, 1e-11))


def test_single_source_flux_one_is_the_plain_call():
    from galario_amd import torch_api as ta
    d = problem(nvis=5000)
    ds = ta.Dataset(d['u'], d['v'], d['re'], d['im'], d['w'], device=DEV)
    I = torch.tensor(d['I'], device=DEV)
    kw = dict(PA=d['PA'], inc=d['inc'], check_central_pixel=False)
    f, dra, ddec = 1.0, d['src'][1, 1], d['src'][1, 2]
    v1 = ds.sample_profile_multi(I, d['Rmin'], d['dR'], d['nxy'], d['dxy'], [[f, dra, ddec]], **kw)
    v0 = ds.sample_profile(I, d['Rmin'], d['dR'], d['nxy'], d['dxy'], dRA=dra, dDec=ddec, **kw)
    assert torch.equal(v1, v0)
    c1 = ds.chi2_profile_multi(I, d['Rmin'], d['dR'], d['nxy'], d['dxy'], [[f, dra, ddec]], **kw)
    c0 = ds.chi2_profile(I, d['Rmin'], d['dR'], d['nxy'], d['dxy'], dRA=dra, dDec=ddec, **kw)
    assert c1.item() == c0.item()
    # no sources: zero model
    z = ds.sample_profile_multi(I, d['Rmin'], d['dR'], d['nxy'], d['dxy'], np.zeros((0, 3)), **kw)
    assert torch.count_nonzero(z).item() == 0
