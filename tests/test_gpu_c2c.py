"""GPU: sampleImage against the complex-to-complex formulation of python/test_galario.py:114-172
(test_R2C_vs_C2C): full numpy fft2 of the shifted image, bilinear interpolation of Re, Im and |.| at
the rotated (u, v) in C2C pixel coordinates, phase from Re/Im, phase shift by the rotated offsets.
The reference's own tolerance: relative 1e-6 (real part), imaginary part with atol = |mean(Re)| rtol.
The R2C path (half spectrum, mirror for u < 0) and this C2C path must agree to that level; both are
independent of the oracle used elsewhere."""
import numpy as np
import pytest

import galario_amd
from oracle import galario_oracle as O

pytestmark = pytest.mark.gpu

PARS = [dict(dRA=0., dDec=0.4, PA=2., nxy=1024), dict(dRA=-3.5, dDec=7.2, PA=-23., nxy=2048),
        dict(dRA=2.3, dDec=3.2, PA=88., nxy=4096), dict(dRA=0., dDec=0., PA=145., nxy=1024)]


def sampling_points(n, maxuv):
    """create_sampling_points (python/utils.py:351-364): seed 42, uvdist in [maxuv/100, maxuv]."""
    rng = np.random.RandomState(42)
    uvdist = rng.uniform(low=maxuv / 100., high=maxuv, size=n)
    phi = rng.uniform(low=0., high=2. * np.pi, size=n)
    return uvdist * np.cos(phi), uvdist * np.sin(phi)


def reference_image(size, x0, y0, sigma_x=50., sigma_y=30.):
    """create_reference_image (python/utils.py:324-348) with correct_axes."""
    x = (np.linspace(0., size - 1, size) - size / 2.)
    xx, yy = np.meshgrid(-x, -x)
    return np.exp(-(xx - x0) ** 2. / sigma_x - (yy - y0) ** 2. / sigma_y)


def bilinear(f, x, y):
    """int_bilin_MT (python/utils.py:390-407): x column, y row, in pixels."""
    fx, fy = np.floor(x).astype(np.int64), np.floor(y).astype(np.int64)
    u, t = x - fx, y - fy
    y0, y1, y2, y3 = f[fy, fx], f[fy + 1, fx], f[fy + 1, fx + 1], f[fy, fx + 1]
    return t * u * (y0 - y1 + y2 - y3) + t * (y1 - y0) + u * (y3 - y0) + y0


@pytest.mark.parametrize("pars", PARS, ids=[str(i) for i in range(4)])
def test_r2c_vs_c2c(pars):
    u, v = sampling_points(1000, 3.e3)
    maxuv = np.max(np.hypot(u, v)) * 4.8          # matrix_size (python/utils.py:410-424)
    nxy = pars['nxy']
    du = maxuv / nxy
    img = reference_image(nxy, -5., 2.)
    PA, dRA, dDec = pars['PA'] * O.DEG, pars['dRA'] * O.ARCSEC, pars['dDec'] * O.ARCSEC
    c, s = np.cos(PA), np.sin(PA)
    urot, vrot = u * c - v * s, u * s + v * c
    dRArot, dDecrot = dRA * c - dDec * s, dRA * s + dDec * c
    ft = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(img)))
    xi, yi = nxy / 2. + urot / du, nxy / 2. + vrot / du
    re, im, amp = bilinear(ft.real, xi, yi), bilinear(ft.imag, xi, yi), bilinear(np.abs(ft), xi, yi)
    ph = np.angle(re + 1j * im)
    vis = amp * (np.cos(ph) + 1j * np.sin(ph))
    theta = urot * 2. * np.pi * dRArot + vrot * 2. * np.pi * dDecrot
    vis = vis * (np.cos(theta) + 1j * np.sin(theta))
    got = galario_amd.double.sampleImage(img, 1. / nxy / du, u, v, dRA=dRA, dDec=dDec, PA=PA)
    rtol = 1e-6
    bad_re = ~np.isclose(got.real, vis.real, rtol, 0)
    bad_im = ~np.isclose(got.imag, vis.imag, rtol, np.abs(np.mean(got.real)) * rtol)
    print(f"\nnxy={nxy}: max rel |dV|/max|V| {np.abs(got - vis).max() / np.abs(vis).max():.2e}, "
          f"{int(bad_re.sum())} / {int(bad_im.sum())} points outside rtol (re / im)")
    assert not bad_re.any() and not bad_im.any()
