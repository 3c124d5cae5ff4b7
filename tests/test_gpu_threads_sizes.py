"""GPU: host-thread re-entrancy of the drop-in API (ctypes releases the GIL; each host thread has its own
stream and workspace) and the largest workgroup-FFT sizes (nxy = 8192 in double, 16384 in float), where
the row and adjoint-row passes use the most LDS."""
import math
import threading

import numpy as np
import pytest

import galario_amd
from oracle import galario_oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')
arcsec = O.ARCSEC


def case(seed, nxy=1024, nvis=20000):
    rng = np.random.default_rng(seed)
    Rmin, dR, nr, dxy = 1e-4 * arcsec, 0.002 * arcsec, 400, 0.01 * arcsec
    R = Rmin + dR * np.arange(nr)
    I = 1e10 * np.exp(-((R - (0.2 + 0.05 * seed) * arcsec) / (0.08 * arcsec)) ** 2)
    duv = 1 / (nxy * dxy)
    b = np.exp(rng.uniform(np.log(duv), np.log(0.45 * nxy * duv), nvis))
    th = rng.uniform(0, 2 * np.pi, nvis)
    u, v = b * np.cos(th), b * np.sin(th)
    re, im, w = rng.standard_normal(nvis), rng.standard_normal(nvis), rng.uniform(0.5, 2, nvis)
    kw = dict(inc=0.3 + 0.1 * seed, PA=0.5 * seed, dRA=0.01 * arcsec, dDec=-0.02 * arcsec, check_central_pixel=False)
    return (I, Rmin, dR, nxy, dxy, u, v, re, im, w), kw


def test_concurrent_host_threads_match_sequential():
    g = galario_amd.double
    cases = [case(s) for s in range(6)]
    seq = [g.chi2Profile(*a, **kw) for a, kw in cases]
    seqv = [g.sampleProfile(*a[:7], **kw) for a, kw in cases]
    out, outv, errs = [None] * 6, [None] * 6, []

    def work(k):
        try:
            for _ in range(3):
                a, kw = cases[k]
                out[k] = g.chi2Profile(*a, **kw)
                outv[k] = g.sampleProfile(*a[:7], **kw)
        except Exception as e:  # noqa: BLE001
            errs.append(e)

    ts = [threading.Thread(target=work, args=(k,)) for k in range(6)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs, errs
    assert out == seq
    for a, b in zip(outv, seqv):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("mod, nxy, tol", [(galario_amd.double, 8192, 1e-12), (galario_amd.single, 16384, 1e-5)],
                         ids=['DP8192', 'SP16384'])
def test_largest_sizes_against_oracle(mod, nxy, tol):
    a, kw = case(3, nxy=nxy, nvis=20000)
    dt = mod.real_dtype
    I, Rmin, dR, _, dxy, u, v, re, im, w = a
    cast = lambda x: np.ascontiguousarray(x, dtype=dt)  # noqa: E731
    vis = mod.sampleProfile(cast(I), Rmin, dR, nxy, dxy, cast(u), cast(v), **kw)
    ref = O.sample_profile(cast(I), Rmin, dR, dxy, nxy, kw['inc'], kw['dRA'], kw['dDec'], 1 / (nxy * dxy), kw['PA'],
                           cast(u), cast(v), check_central=False)
    err = np.abs(vis - ref).max() / np.abs(ref).max()
    print(f"\nnxy={nxy}: max|dV|/max|V| = {err:.2e}")
    assert err <= tol


def test_gradient_at_nxy_8192():
    """The adjoint passes at the largest double size (most LDS per workgroup): the gradient call's chi2
    equals the forward chi2, gradients repeat to 1e-12, and the profile gradient matches a central finite
    difference along a smooth direction (relative 1e-5; chi2 is quadratic in the profile)."""
    from galario_amd import torch_api as ta
    a, kw = case(5, nxy=8192, nvis=20000)
    I, Rmin, dR, nxy, dxy, u, v, re, im, w = a
    ds = ta.Dataset(u, v, re, im, w, device='cuda:0')
    It = torch.tensor(I, device='cuda:0')
    kw2 = {k: kw[k] for k in ('inc', 'PA', 'dRA', 'dDec', 'check_central_pixel')}
    c1, gI1, gg1 = ds.chi2_profile_grad(It, Rmin, dR, nxy, dxy, **kw2)
    c2, gI2, gg2 = ds.chi2_profile_grad(It, Rmin, dR, nxy, dxy, **kw2)
    cf = ds.chi2_profile(It, Rmin, dR, nxy, dxy, **kw2)
    assert c1.item() == c2.item() == cf.item()
    assert torch.isfinite(gI1).all() and torch.isfinite(gg1).all()
    assert torch.linalg.norm(gI1 - gI2) <= 1e-12 * torch.linalg.norm(gI1)
    # directional finite difference of the profile gradient (chi2 is linear-quadratic in I)
    d = torch.sin(torch.arange(len(I), device='cuda:0', dtype=torch.float64) / 17.0) * It
    h = 1e-6
    fd = (ds.chi2_profile(It + h * d, Rmin, dR, nxy, dxy, **kw2).item()
          - ds.chi2_profile(It - h * d, Rmin, dR, nxy, dxy, **kw2).item()) / (2 * h)
    an = float(gI1 @ d)
    print(f"\nnxy=8192 profile direction: fd {fd:.10e} analytic {an:.10e}")
    assert math.isclose(fd, an, rel_tol=1e-5)
