"""Pins the CPU oracle (oracle/galario_oracle.py) with checks independent of it.

Each stage is compared with a formulation that does not share its code:
np.interp on a meshgrid (the reference test's sweep_ref, python/utils.py:245-321),
closed forms, long-double brute-force DFTs, a direct Fourier sum, and the C
restatement.  Parameter sets and tolerances follow python/test_galario.py.
A few tests plant a mistake in a copy of an oracle function and check that
the pinning comparison catches it.
"""
import math

import numpy as np
import pytest

from oracle import galario_oracle as O
from oracle import port

arcsec, deg = O.ARCSEC, O.DEG


# ---------------------------------------------------------------- helpers (independent formulations)
def radial_profile(Rmin, dR, nrad, mode, gauss_width):
    """python/utils.py:161-172."""
    R = np.linspace(Rmin, Rmin + dR * (nrad - 1), nrad)
    if mode == 'Gauss':
        return np.exp(-(R / gauss_width) ** 2)
    return np.cos(2. * np.pi * R / gauss_width) ** 2. * np.exp(-(R / gauss_width) ** 2)


def sweep_interp(I, Rmin, dR, nxy, dxy, inc):
    """Image by np.interp on a linspace meshgrid (the formulation of python/utils.py:245-321),
    zero beyond the last radial node, central pixel by the trapezoid closed form below."""
    nrad = len(I)
    grid = np.linspace(Rmin, Rmin + dR * (nrad - 1), nrad)
    x = np.linspace(0.5, -0.5 + 1. / nxy, nxy) * dxy * nxy
    y = x.copy()
    xx, yy = np.meshgrid(x / np.cos(inc), y)
    r = np.sqrt(xx ** 2 + yy ** 2)
    img = np.interp(r, grid, I, left=np.nan, right=0.0)
    below = np.isnan(img)          # r < Rmin: the reference extrapolates the first cell
    img[below] = I[0] + (r[below] - Rmin) * (I[1] - I[0]) / dR
    img[nxy // 2, nxy // 2] = central_trapezoid(I, Rmin, dR, dxy)
    return img * dxy ** 2


def central_trapezoid(I, Rmin, dR, dxy):
    """Mean of I over the central pixel: trapezoid of 2 I(R) R over the nodes
    Rmin, Rmin+dR, ..., Rmin+iIN dR, dxy/2 (I linear in the last cell), / ((dxy/2)^2 - Rmin^2)."""
    iIN = int(math.floor((dxy / 2 - Rmin) / dR))
    nodes = list(Rmin + dR * np.arange(iIN + 1)) + [dxy / 2]
    vals = list(I[:iIN + 1]) + [I[iIN] + (I[iIN + 1] - I[iIN]) / dR * (dxy / 2 - nodes[iIN])]
    f = [2 * a * b for a, b in zip(vals, nodes)]
    T = sum(0.5 * (f[k] + f[k + 1]) * (nodes[k + 1] - nodes[k]) for k in range(len(f) - 1))
    return T / ((dxy / 2) ** 2 - Rmin ** 2)


# ---------------------------------------------------------------- sweep
@pytest.mark.parametrize("Rmin, dR, nrad, nxy, dxy, inc, mode",
                         [(1e-6, 0.001, 2000, 1024, 0.2, 20., 'Gauss'),
                          (1e-6, 0.001, 2000, 2048, 0.2, 44.23, 'Cos-Gauss'),
                          (1e-6, 0.001, 2000, 2048, 0.5, 20., 'Gauss'),
                          (1e-6, 0.001, 2000, 1024, 0.3, 20., 'Gauss')])
def test_sweep_vs_interp_meshgrid(Rmin, dR, nrad, nxy, dxy, inc, mode):
    """python/test_galario.py:63-111 parameter sets, rtol 1e-12."""
    Rmin, dR, dxy, inc = Rmin * arcsec, dR * arcsec, dxy * arcsec, np.radians(inc)
    I = radial_profile(Rmin, dR, nrad, mode, gauss_width=dxy * 6)
    img = O.sweep(I, Rmin, dR, nxy, dxy, inc)
    ref = sweep_interp(I, Rmin, dR, nxy, dxy, inc)
    np.testing.assert_allclose(img, ref, rtol=1e-12, atol=1e-300)


def test_central_pixel_linear_profile_closed_form():
    """For I(R) = a + bR the trapezoid sum has the closed form
    exact integral + sum over cells of h^3 f''/12, f = 2 I R (f'' = 4b)."""
    a, b = 2.5, -0.7 / arcsec
    Rmin, dR, dxy = 1e-4 * arcsec, 0.003 * arcsec, 0.05 * arcsec
    nr = 100
    I = a + b * (Rmin + dR * np.arange(nr))
    iIN = int(math.floor((dxy / 2 - Rmin) / dR))
    R1, R2 = Rmin, dxy / 2
    exact = 2 * (a * (R2 ** 2 - R1 ** 2) / 2 + b * (R2 ** 3 - R1 ** 3) / 3)
    hlast = R2 - (Rmin + iIN * dR)
    trap = exact + (4 * b) / 12 * (iIN * dR ** 3 + hlast ** 3)
    expected = dxy ** 2 * trap / (R2 ** 2 - R1 ** 2)
    got = O.central_pixel_value(I, Rmin, dR, dxy)
    assert abs(got - expected) <= 1e-12 * abs(expected)


def test_sweep_constant_profile_known_answer():
    """A constant profile gives dxy^2 * a on every swept pixel and on the central pixel."""
    a = 3.0
    Rmin, dR, dxy, nxy, nr = 1e-4 * arcsec, 0.001 * arcsec, 0.02 * arcsec, 64, 300
    img = O.sweep(np.full(nr, a), Rmin, dR, nxy, dxy, 0.3)
    rmax = O.sweep_rmax(nr, Rmin, dR, nxy, dxy)
    assert math.isclose(img[nxy // 2, nxy // 2], a * dxy ** 2, rel_tol=1e-13)
    o = nxy // 2 - rmax
    box = img[o:o + 2 * rmax, o:o + 2 * rmax]
    nz = box[box != 0]
    np.testing.assert_allclose(nz, a * dxy ** 2, rtol=1e-14)
    assert np.all(img[:o] == 0) and np.all(img[:, :o] == 0)


def test_central_pixel_guard_and_range():
    """CHECK_CENTRAL_PIXEL (src/galario.cpp:113-121) and the profile-range check."""
    with pytest.raises(ValueError, match=r"Expect \(dxy/2-Rmin\)/dR > 5"):
        O.check_central_pixel(0.01 * arcsec, 1e-4 * arcsec, 0.002 * arcsec)
    O.check_central_pixel(0.2 * arcsec, 1e-6 * arcsec, 0.001 * arcsec)
    with pytest.raises(ValueError):
        O.central_pixel_value(np.ones(3), 1e-6 * arcsec, 0.001 * arcsec, 0.2 * arcsec)


def test_planted_sweep_mistake_is_caught(monkeypatch):
    """A dropped 1/cos(inc) in a copy of the sweep must fail the meshgrid comparison."""
    Rmin, dR, dxy, nxy, inc = 1e-6 * arcsec, 0.001 * arcsec, 0.2 * arcsec, 256, np.radians(30.)
    I = radial_profile(Rmin, dR, 2000, 'Gauss', gauss_width=dxy * 6)
    monkeypatch.setattr(O.math, 'cos', lambda x: 1.0)
    bad = O.sweep(I, Rmin, dR, nxy, dxy, inc)
    monkeypatch.undo()
    ref = sweep_interp(I, Rmin, dR, nxy, dxy, inc)
    assert not np.allclose(bad, ref, rtol=1e-12, atol=0)


# ---------------------------------------------------------------- shifts and FFT
def test_fftshifts_bitwise():
    rng = np.random.default_rng(3)
    a = rng.random((64, 64))
    assert np.array_equal(O.fftshift2(a), np.fft.fftshift(a))
    c = rng.random((64, 33)) + 1j * rng.random((64, 33))
    assert np.array_equal(O.fftshift_axis0(c), np.fft.fftshift(c, axes=0))


@pytest.mark.parametrize("n", [2, 4, 8, 16])
def test_rfft2_vs_long_double_brute_force(n):
    rng = np.random.default_rng(n)
    img = rng.standard_normal((n, n))
    ref = O.brute_rfft2(img)
    got = O.rfft2(img)
    err = np.abs(got - ref.astype(np.complex128)).max() / np.abs(ref).max()
    assert err < 5e-15          # O(eps log N)


def test_rfft2_known_answers():
    N = 32
    d = np.zeros((N, N)); d[0, 0] = 1.0
    assert np.allclose(O.rfft2(d), 1.0, rtol=0, atol=1e-15)
    c = np.full((N, N), 2.0)
    F = O.rfft2(c)
    assert abs(F[0, 0] - 2.0 * N * N) < 1e-12 and np.abs(F).sum() - abs(F[0, 0]) < 1e-10
    k0, k1 = 3, 5
    n0, n1 = np.meshgrid(np.arange(N), np.arange(N), indexing='ij')
    cosimg = np.cos(2 * np.pi * (k0 * n0 + k1 * n1) / N)
    F = O.rfft2(cosimg)
    assert abs(F[k0, k1] - N * N / 2) < 1e-10 and abs(F[N - k0, 0]) < 1e-10
    big = np.abs(F) > 1e-8
    assert big.sum() == 1


def test_rfft2_parseval():
    rng = np.random.default_rng(7)
    N = 128
    x = rng.standard_normal((N, N))
    F = O.rfft2(x)
    C = N // 2 + 1
    w = np.full(C, 2.0); w[0] = 1.0; w[-1] = 1.0
    energy = (np.abs(F) ** 2 * w[None, :]).sum() / (N * N)
    assert math.isclose(energy, (x ** 2).sum(), rel_tol=1e-12)


# ---------------------------------------------------------------- rotation, interpolation, phase, chi2
def test_uv_rotate_closed_form():
    rng = np.random.default_rng(1)
    u, v = rng.standard_normal(100), rng.standard_normal(100)
    dRA, dDec, ur, vr = O.uv_rotate(0.0, 0.3, -0.2, u, v)
    assert (dRA, dDec) == (0.3, -0.2) and np.array_equal(ur, u) and np.array_equal(vr, v)
    PA = 0.7
    dRA, dDec, ur, vr = O.uv_rotate(PA, 0.3, -0.2, u, v)
    z = (u + 1j * v) * np.exp(1j * PA)
    np.testing.assert_allclose(ur, z.real, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(vr, z.imag, rtol=1e-13, atol=1e-15)
    zo = (0.3 - 0.2j) * np.exp(1j * PA)
    assert math.isclose(dRA, zo.real, rel_tol=1e-14) and math.isclose(dDec, zo.imag, rel_tol=1e-14)


def int_bilin(f, x, y):
    """Bilinear interpolation loop of python/utils.py:390-407 (rows y, columns x)."""
    out = np.zeros(len(x))
    for i in range(len(x)):
        t = y[i] - np.floor(y[i])
        s = x[i] - np.floor(x[i])
        r, c = int(np.floor(y[i])), int(np.floor(x[i]))
        y0, y1, y2, y3 = f[r, c], f[r + 1, c], f[r + 1, c + 1], f[r, c + 1]
        out[i] = t * s * (y0 - y1 + y2 - y3) + t * (y1 - y0) + s * (y3 - y0) + y0
    return out


def test_interpolate_vs_separate_bilinears():
    """python/test_galario.py:175-219: Re, Im, |.| interpolated separately; Im flipped for u<0."""
    rng = np.random.default_rng(5)
    N, du = 256, 1000. / 256
    img = rng.standard_normal((N, N))
    S = O.spectrum_of_image(img)
    n = 2000
    b = rng.uniform(10, 400, n); th = rng.uniform(0, 2 * np.pi, n)
    u, v = b * np.cos(th), b * np.sin(th)
    indu = np.abs(u) / du
    indv = N / 2 + np.sign(u) * v / du
    Re = int_bilin(S.real, indu, indv)
    Im = int_bilin(S.imag, indu, indv)
    Amp = int_bilin(np.abs(S), indu, indv)
    Im[u < 0] *= -1
    ph = np.angle(Re + 1j * Im)
    ref = Amp * np.exp(1j * ph)
    got = O.interpolate(S, 1.0, u, v, du)
    np.testing.assert_allclose(got.real, ref.real, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    np.testing.assert_allclose(got.imag, ref.imag, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_interpolate_at_nodes_is_the_grid_value():
    rng = np.random.default_rng(6)
    N, du = 128, 10.0
    S = rng.standard_normal((N, N // 2 + 1)) + 1j * rng.standard_normal((N, N // 2 + 1))
    cols = rng.integers(0, N // 2, 300)
    rows = rng.integers(1, N - 1, 300)
    u = cols * du
    v = (rows - N / 2) * du
    got = O.interpolate(S, 1.0, u, v, du)
    np.testing.assert_allclose(got, S[rows, cols], rtol=1e-14)
    neg = cols > 0
    got_n = O.interpolate(S, 1.0, -u[neg], -v[neg], du)     # u<0: Hermitian fold, conjugate phase
    np.testing.assert_allclose(got_n, np.conj(S[rows[neg], cols[neg]]), rtol=1e-14)


def test_apply_phase_closed_form():
    rng = np.random.default_rng(8)
    n = 1000
    u, v = rng.uniform(-3e3, 3e3, n), rng.uniform(-3e3, 3e3, n)
    vis = rng.random(n) * 10 + 1j * rng.random(n) * 30
    dRA, dDec = 0.4 * arcsec, 10. * arcsec
    got = O.apply_phase(dRA, dDec, u, v, vis)
    ref = vis * np.exp(2j * np.pi * (u * dRA + v * dDec))
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-13)
    assert np.array_equal(O.apply_phase(0.0, 0.0, u, v, vis), vis)


def test_reduce_chi2_vs_sum():
    rng = np.random.default_rng(9)
    n = 1000
    x = rng.random(n) + 1j * rng.random(n)
    y = rng.random(n) + 1j * rng.random(n)
    w = rng.random(n)
    ref = np.sum(((x.real - y.real) ** 2 + (x.imag - y.imag) ** 2) * w)
    assert math.isclose(O.reduce_chi2(x.real, x.imag, y, w), ref, rel_tol=1e-14)


# ---------------------------------------------------------------- whole pipeline
def test_pipeline_vs_direct_dft_at_grid_nodes():
    """At (u, v) on grid nodes the FFT + interpolation path equals the direct Fourier sum."""
    rng = np.random.default_rng(11)
    N, dxy = 256, 0.02 * arcsec
    Rmin, dR, nr = 1e-4 * arcsec, 0.001 * arcsec, 800
    I = radial_profile(Rmin, dR, nr, 'Cos-Gauss', gauss_width=0.3 * arcsec) * 1e10
    img = O.sweep(I, Rmin, dR, N, dxy, 0.5)
    duv = 1 / (N * dxy)
    cols = rng.integers(-N // 2 + 2, N // 2 - 1, 500)
    rows = rng.integers(-N // 2 + 2, N // 2 - 1, 500)
    u, v = cols * duv, rows * duv
    dRA, dDec = 0.013 * arcsec, -0.021 * arcsec
    got = O.sample_image(img, 1.0, dRA, dDec, duv, 0.0, u, v)
    ref = O.direct_dft_vis(img, dxy, u, v, dRA, dDec, 0.0)
    err = np.abs(got - ref) / np.abs(ref).max()
    assert err.max() < 1e-12


def test_profile_equals_image_of_sweep_and_origin_lower():
    rng = np.random.default_rng(12)
    N, dxy = 256, 0.02 * arcsec
    Rmin, dR, nr = 1e-4 * arcsec, 0.001 * arcsec, 800
    I = radial_profile(Rmin, dR, nr, 'Gauss', gauss_width=0.3 * arcsec)
    duv = 1 / (N * dxy)
    b = rng.uniform(2, 100, 400) * duv; th = rng.uniform(0, 2 * np.pi, 400)
    u, v = b * np.cos(th), b * np.sin(th)
    vp = O.sample_profile(I, Rmin, dR, dxy, N, 0.4, 0.1 * arcsec, 0.2 * arcsec, duv, 0.3, u, v)
    img = O.sweep(I, Rmin, dR, N, dxy, 0.4)
    vi = O.sample_image(img, 1.0, 0.1 * arcsec, 0.2 * arcsec, duv, 0.3, u, v)
    assert np.array_equal(vp, vi)
    # origin='lower' of the vertically flipped (and rolled) image gives the same visibilities
    img_low = np.roll(np.flipud(img), 1, 0)
    vl = O.sample_image(img_low, -1.0, 0.1 * arcsec, 0.2 * arcsec, duv, 0.3, u, v)
    np.testing.assert_allclose(vl, vi, rtol=1e-10, atol=1e-12 * np.abs(vi).max())


def test_c_port_matches_numpy_oracle():
    rng = np.random.default_rng(13)
    N, dxy = 512, 0.01 * arcsec
    Rmin, dR, nr = 1e-4 * arcsec, 0.002 * arcsec, 1000
    I = radial_profile(Rmin, dR, nr, 'Gauss', gauss_width=0.5 * arcsec)
    duv = 1 / (N * dxy)
    b = np.exp(rng.uniform(np.log(duv), np.log(200 * duv), 5000)); th = rng.uniform(0, 2 * np.pi, 5000)
    u, v = b * np.cos(th), b * np.sin(th)
    re, im, w = rng.standard_normal(5000), rng.standard_normal(5000), rng.random(5000)
    args = (I, Rmin, dR, dxy, N, 0.7, 0.02 * arcsec, 0.02 * arcsec, duv, 2.1, u, v)
    vo = O.sample_profile(*args, check_central=False)
    vp = port.sample_profile(*args)
    assert np.abs(vo - vp).max() <= 1e-13 * np.abs(vo).max()
    co = O.reduce_chi2(re, im, vo, w)
    cp = port.chi2_profile(*args, re, im, w)
    assert math.isclose(co, cp, rel_tol=1e-12)


# ---------------------------------------------------------------- golden fixtures
def test_oracle_vs_golden_direct_dft():
    """tests/golden/dft_nodes.npz: long-double direct Fourier sum at grid nodes."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'dft_nodes.npz'))
    img = O.sweep(g['intensity'], float(g['Rmin']), float(g['dR']), int(g['nxy']), float(g['dxy']), float(g['inc']))
    assert np.array_equal(img, g['image'])
    duv = 1 / (int(g['nxy']) * float(g['dxy']))
    vis = O.sample_image(img, 1.0, float(g['dRA']), float(g['dDec']), duv, 0.0, g['u'], g['v'])
    ref = g['vis_re'] + 1j * g['vis_im']
    assert np.abs(vis - ref).max() <= 1e-13 * np.abs(ref).max()


# ---------------------------------------------------------------- the float32 branches
# The oracle's single-precision paths (galario's float build: dreal = float, with the C++'s promotions to
# double) are pinned against the independent float64 formulations above, at the float bar 1e-5.
@pytest.mark.parametrize("Rmin, dR, nrad, nxy, dxy, inc, mode",
                         [(1e-6, 0.001, 2000, 1024, 0.2, 20., 'Gauss'),
                          (1e-6, 0.001, 2000, 2048, 0.2, 44.23, 'Cos-Gauss'),
                          (1e-6, 0.001, 2000, 1024, 0.3, 20., 'Gauss')])
def test_float32_sweep_vs_interp_meshgrid(Rmin, dR, nrad, nxy, dxy, inc, mode):
    """Float32 sweep (box and central pixel) against np.interp on the float64 meshgrid: 1e-5 of the
    largest pixel; a dropped central pixel exceeds that bound."""
    Rmin, dR, dxy, inc = Rmin * arcsec, dR * arcsec, dxy * arcsec, np.radians(inc)
    I = radial_profile(Rmin, dR, nrad, mode, gauss_width=dxy * 6)
    ref = sweep_interp(I, Rmin, dR, nxy, dxy, inc)
    img = O.sweep(I.astype(np.float32), Rmin, dR, nxy, dxy, inc, dtype=np.float32)
    assert img.dtype == np.float32
    scale = np.abs(ref).max()
    assert np.abs(img - ref).max() <= 1e-5 * scale
    c = nxy // 2
    assert abs(img[c, c] - ref[c, c]) <= 1e-5 * abs(ref[c, c])
    bad = img.copy()
    bad[c, c] = 0
    assert np.abs(bad - ref).max() > 1e-5 * scale


@pytest.mark.parametrize("Rmin, dR, dxy", [(1e-4, 0.003, 0.05), (1e-6, 0.001, 0.2), (1e-4, 2e-5, 0.01)])
def test_float32_central_pixel_vs_trapezoid(Rmin, dR, dxy):
    """The float32 central-pixel branch (its own promotions to double) against the float64 trapezoid
    closed form of a smooth profile: 1e-5 relative."""
    Rmin, dR, dxy = Rmin * arcsec, dR * arcsec, dxy * arcsec
    nr = int((dxy / 2 - Rmin) / dR) + 10
    R = Rmin + dR * np.arange(nr)
    I = np.exp(-(R / (dxy * 0.7)) ** 2) + 0.2
    ref = dxy ** 2 * central_trapezoid(I, Rmin, dR, dxy)
    got = O.central_pixel_value(I.astype(np.float32), Rmin, dR, dxy, dtype=np.float32)
    assert isinstance(got, np.float32)
    assert abs(float(got) - ref) <= 1e-5 * abs(ref), (float(got), ref)


def test_float32_pipeline_vs_direct_dft_at_grid_nodes():
    """The float32 pipeline (sweep, shift, FFT, interpolation, phase, chi2) at grid-node uv points against
    the float64 direct Fourier sum of the float64 image and math.fsum: visibilities 1e-5 of max|V|,
    chi2 1e-5 relative."""
    rng = np.random.default_rng(21)
    N, dxy = 128, 0.03 * arcsec
    Rmin, dR, nr = 1e-4 * arcsec, 0.001 * arcsec, 2500
    R = Rmin + dR * np.arange(nr)
    I = 1e10 * (np.exp(-((R - 0.6 * arcsec) / (0.2 * arcsec)) ** 2) + 0.3 * np.exp(-R / (0.5 * arcsec)))
    inc, dRA, dDec = 0.5, 0.02 * arcsec, -0.01 * arcsec
    duv = 1 / (N * dxy)
    u = rng.integers(-N // 2 + 2, N // 2 - 1, 300) * duv
    v = rng.integers(-N // 2 + 2, N // 2 - 1, 300) * duv
    img64 = O.sweep(I, Rmin, dR, N, dxy, inc, check_central=False)
    ref = O.direct_dft_vis(img64, dxy, u, v, dRA, dDec, 0.0)
    f = np.float32
    vis = O.sample_profile(I.astype(f), f(Rmin), f(dR), f(dxy), N, f(inc), f(dRA), f(dDec), f(duv), f(0.0),
                           u.astype(f), v.astype(f), dtype=f, check_central=False)
    assert vis.dtype == np.complex64
    assert np.abs(vis - ref).max() <= 1e-5 * np.abs(ref).max()
    re, im = ref.real + rng.standard_normal(300) * 1e6, ref.imag + rng.standard_normal(300) * 1e6
    w = rng.uniform(0.5, 2.0, 300) * 1e-12
    c_ref = math.fsum(w * ((ref.real - re) ** 2 + (ref.imag - im) ** 2))
    c32 = O.reduce_chi2(re.astype(f), im.astype(f), vis, w.astype(f), dtype=f)
    assert abs(float(c32) - c_ref) <= 1e-5 * c_ref
