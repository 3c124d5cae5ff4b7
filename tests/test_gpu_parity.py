"""GPU parity: the HIP library (through galario_amd's C ABI) against the CPU oracle.

Tolerances (stated per test):
  * bit-exact where galario's arithmetic is restated without an FFT in between:
    sweep image, fftshift, fftshift_axis0, uv_rotate;
  * double: visibilities max|dV| <= 1e-12 max|V| and ||dV|| <= 1e-12 ||V||,
    chi2 relative 1e-12 (the FFT engines differ: pocketfft vs the gfx950
    Stockham FFT, both O(eps log N));
  * single: 1e-5 relative to max|V| and for chi2, against the oracle run in
    float32 (the reference's single build).
Parameter sets follow python/test_galario.py.
"""
import math
import os

import numpy as np
import pytest

import galario_amd
from galario_amd import synthetic as syn
from oracle import galario_oracle as O

pytestmark = pytest.mark.gpu

gD, gS = galario_amd.double, galario_amd.single
arcsec, deg = O.ARCSEC, O.DEG
HERE = os.path.dirname(os.path.abspath(__file__))


def assert_vis_close(got, ref, tol):
    scale = np.abs(ref).max()
    err = np.abs(got - ref)
    assert err.max() <= tol * scale, f"max err {err.max() / scale:.3e} > {tol:.1e} (relative to max|V|)"
    assert np.linalg.norm(got - ref) <= tol * np.linalg.norm(ref), "norm-wise error"


def radial_profile(Rmin, dR, nrad, mode, gauss_width, dtype=np.float64):
    R = np.linspace(Rmin, Rmin + dR * (nrad - 1), nrad)
    if mode == 'Gauss':
        I = np.exp(-(R / gauss_width) ** 2)
    else:
        I = np.cos(2. * np.pi * R / gauss_width) ** 2. * np.exp(-(R / gauss_width) ** 2)
    return I.astype(dtype)


def uv_points(n, maxuv, seed=42, dtype=np.float64):
    """create_sampling_points (python/utils.py:351-364), with a local generator."""
    rng = np.random.RandomState(seed)
    uvdist = rng.uniform(low=maxuv / 100., high=maxuv, size=n)
    phi = rng.uniform(low=0., high=2. * np.pi, size=n)
    return (uvdist * np.cos(phi)).astype(dtype), (uvdist * np.sin(phi)).astype(dtype)


# ---------------------------------------------------------------- sweep: bit-exact
@pytest.mark.parametrize("Rmin, dR, nrad, nxy, dxy, inc, mode",
                         [(1e-6, 0.001, 2000, 1024, 0.2, 20., 'Gauss'),
                          (1e-6, 0.001, 2000, 2048, 0.2, 44.23, 'Cos-Gauss'),
                          (1e-6, 0.001, 2000, 2048, 0.5, 20., 'Gauss'),
                          (1e-6, 0.001, 2000, 1024, 0.3, 20., 'Gauss')])
@pytest.mark.parametrize("mod", [gD, gS], ids=['DP', 'SP'])
def test_sweep_bit_exact(mod, Rmin, dR, nrad, nxy, dxy, inc, mode):
    """python/test_galario.py:63-111 sets; the GPU image equals the oracle's bit for bit."""
    dt = mod.real_dtype
    Rmin, dR, dxy, inc = Rmin * arcsec, dR * arcsec, dxy * arcsec, np.radians(inc)
    I = radial_profile(Rmin, dR, nrad, mode, dxy * 6, dt)
    got = mod.sweep(I, dt(Rmin), dt(dR), nxy, dt(dxy), dt(inc))
    ref = O.sweep(I, Rmin, dR, nxy, dxy, inc, dtype=dt)
    assert got.dtype == dt and got.shape == (nxy, nxy)
    assert np.array_equal(got, ref)


# ---------------------------------------------------------------- FFT and shifts
@pytest.mark.parametrize("size", [1024, 1000, 64, 2, 4, 8192])
@pytest.mark.parametrize("mod, tol", [(gD, 1e-13), (gS, 1e-5)], ids=['DP', 'SP'])
def test_fft2d(mod, tol, size):
    """python/test_galario.py:222-244, plus the generic-size path (1000, 64) and N=8192."""
    if size == 8192 and mod is gS:
        pytest.skip("covered in double")
    rng = np.random.default_rng(size)
    img = rng.standard_normal((size, size)).astype(mod.real_dtype)
    got = mod._fft2d(img)
    ref = np.fft.rfft2(img.astype(np.float64))
    assert got.shape == (size, size // 2 + 1)
    assert_vis_close(got, ref, tol)


@pytest.mark.parametrize("mod", [gD, gS], ids=['DP', 'SP'])
def test_fftshifts_bitwise(mod):
    """python/test_galario.py:247-289."""
    rng = np.random.default_rng(0)
    img = rng.random((1024, 1024)).astype(mod.real_dtype)
    got = mod._fftshift(img)
    assert np.array_equal(got.view(mod.real_dtype)[:, :-2], np.fft.fftshift(img))
    m = (rng.random((1024, 513)) + 1j * rng.random((1024, 513))).astype(mod.complex_dtype)
    ref = np.fft.fftshift(m, axes=0)
    mod._fftshift_axis0(m)
    assert np.array_equal(m, ref)


# ---------------------------------------------------------------- per-stage expert functions
@pytest.mark.parametrize("mod", [gD, gS], ids=['DP', 'SP'])
def test_uv_rotate_bit_exact(mod):
    dt = mod.real_dtype
    u, v = uv_points(1000, 3e3, dtype=dt)
    for PA in (0.0, 2. * deg, -23. * deg, 88. * deg):
        dRA, dDec = dt(2.3 * arcsec), dt(3.2 * arcsec)
        got = mod.uv_rotate(dt(PA), dRA, dDec, u, v)
        ref = O.uv_rotate(PA, dRA, dDec, u, v, dtype=dt)
        assert got[0] == ref[0] and got[1] == ref[1]
        assert np.array_equal(got[2], ref[2]) and np.array_equal(got[3], ref[3])


@pytest.mark.parametrize("mod, rtol, atol", [(gS, 2e-4, 1e-5), (gD, 1e-13, 1e-13)], ids=['SP', 'DP'])
def test_interpolate(mod, rtol, atol):
    """python/test_galario.py:175-219 (tolerances relative to max|V| for atol)."""
    dt, ct = mod.real_dtype, mod.complex_dtype
    size, maxuv = 1024, 1000.
    rng = np.random.default_rng(1)
    img = rng.standard_normal((size, size))
    ft = np.fft.fftshift(np.fft.rfft2(np.fft.fftshift(img)), axes=0).astype(ct)
    u, v = uv_points(10000, maxuv / 2.2, dtype=dt)
    du = dt(maxuv / size)
    got = mod.interpolate(ft, du, u, v)
    ref = O.interpolate(ft, 1.0, u, v, du, dtype=dt)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(got.real, ref.real, rtol=rtol, atol=atol * scale)
    np.testing.assert_allclose(got.imag, ref.imag, rtol=rtol, atol=atol * scale)
    got_l = mod.interpolate(ft, du, u, v, origin='lower')
    ref_l = O.interpolate(ft, -1.0, u, v, du, dtype=dt)
    np.testing.assert_allclose(got_l.real, ref_l.real, rtol=rtol, atol=atol * scale)


@pytest.mark.parametrize("mod, tol", [(gS, 1e-5), (gD, 1e-13)], ids=['SP', 'DP'])
@pytest.mark.parametrize("pars", [(0., 0.4), (-3.5, 7.2), (2.3, 3.2), (0., 0.)])
def test_apply_phase_vis(mod, tol, pars):
    """python/test_galario.py:292-328."""
    dt, ct = mod.real_dtype, mod.complex_dtype
    dRA, dDec = dt(pars[0] * arcsec), dt(pars[1] * arcsec)
    u, v = uv_points(10000, 3e3, dtype=dt)
    rng = np.random.default_rng(2)
    vis = (rng.random(10000) * 10 + 1j * rng.random(10000) * 30).astype(ct)
    got = mod.apply_phase_vis(dRA, dDec, u, v, vis)
    ref = O.apply_phase(dRA, dDec, u, v, vis, dtype=dt)
    assert_vis_close(got, ref, tol)


@pytest.mark.parametrize("mod, tol", [(gS, 1e-6), (gD, 1e-15)], ids=['SP', 'DP'])
def test_reduce_chi2(mod, tol):
    """python/test_galario.py:331-345."""
    dt = mod.real_dtype
    rng = np.random.default_rng(3)
    n = 1000
    x = (3. * rng.random(n) + 2.8 + 1j * (rng.random(n) + 8.2)).astype(mod.complex_dtype)
    y = (8. * rng.uniform(0.5, 3., n) + 5.7 + 1j * (rng.uniform(0., 6., n) + 21.2)).astype(mod.complex_dtype)
    w = rng.uniform(0., 1e4, n).astype(dt)
    w /= w.sum()
    ref = math.fsum((((x.real.astype(float) - y.real) ** 2 + (x.imag.astype(float) - y.imag) ** 2) * w).tolist())
    got = mod.reduce_chi2(x.real.copy(), x.imag.copy(), w, y.copy())
    assert math.isclose(got, ref, rel_tol=tol)


# ---------------------------------------------------------------- whole path, reference parameter sets
PARS = [{'dRA': 0., 'dDec': 0.4, 'PA': 2., 'nxy': 1024},
        {'dRA': -3.5, 'dDec': 7.2, 'PA': -23., 'nxy': 2048},
        {'dRA': 2.3, 'dDec': 3.2, 'PA': 88., 'nxy': 4096},
        {'dRA': 0., 'dDec': 0., 'PA': 145., 'nxy': 1024}]


@pytest.mark.parametrize("pars", PARS, ids=[str(i) for i in range(4)])
def test_all_double(pars):
    """python/test_galario.py:431-504: sampleImage/sampleProfile/chi2Image/chi2Profile
    against the oracle at 1e-12 (visibilities relative to max|V|, chi2 relative)."""
    nsamples = 1000
    u, v = uv_points(nsamples, 3e3)
    uvdist = np.hypot(u, v)
    maxuv = uvdist.max() * 4.8
    dxy = 1. / maxuv
    nxy = pars['nxy']
    Rmin, dR, nrad, inc = dxy / 100., dxy / 10.5, 10000, 20. * deg
    dRA, dDec, PA = pars['dRA'] * arcsec, pars['dDec'] * arcsec, pars['PA'] * deg
    I = radial_profile(Rmin, dR, nrad, 'Gauss', dxy * 10)
    duv = 1 / (dxy * nxy)
    img = O.sweep(I, Rmin, dR, nxy, dxy, inc)
    vis_i = gD.sampleImage(img, dxy, u, v, dRA=dRA, dDec=dDec, PA=PA)
    ref_i = O.sample_image(img, 1.0, dRA, dDec, duv, PA, u, v)
    assert_vis_close(vis_i, ref_i, 1e-12)
    vis_p = gD.sampleProfile(I, Rmin, dR, nxy, dxy, u, v, dRA=dRA, dDec=dDec, PA=PA, inc=inc)
    assert_vis_close(vis_p, ref_i, 1e-12)
    assert np.array_equal(vis_p, vis_i) or np.abs(vis_p - vis_i).max() <= 1e-13 * np.abs(vis_i).max()
    rng = np.random.default_rng(4)
    re, im = rng.random(nsamples) * 3 + 2.8, rng.random(nsamples) + 8.2
    w = rng.uniform(0, 1e4, nsamples); w /= w.sum()
    c_ref = O.reduce_chi2(re, im, ref_i, w)
    c_i = gD.chi2Image(img, dxy, u, v, re, im, w, dRA=dRA, dDec=dDec, PA=PA)
    c_p = gD.chi2Profile(I, Rmin, dR, nxy, dxy, u, v, re, im, w, dRA=dRA, dDec=dDec, PA=PA, inc=inc)
    assert math.isclose(c_i, c_ref, rel_tol=1e-12)
    assert math.isclose(c_p, c_ref, rel_tol=1e-12)


@pytest.mark.parametrize("nxy", [1024, 2048, 4096, 8192])
@pytest.mark.parametrize("rmax", [37, 255, 256])
def test_profile_column_window(nxy, rmax):
    """The column transform has two forms: with the profile's support inside a 512-row window
    (rmax <= 255 pixels, N >= 1024) one 512-point transform per residue class of output rows, else
    the full N-point transform.  Both, at the window's edge (rmax 255 and 256), against the oracle at
    1e-12 (visibilities relative to max|V|, chi2 relative); float at 1e-5 for rmax 255."""
    nvis = 2000
    dxy = 0.01 * arcsec
    duv = 1 / (nxy * dxy)
    rng = np.random.default_rng(rmax + nxy)
    b = rng.uniform(0.5 * duv, 0.45 * nxy * duv, nvis)
    th = rng.uniform(0, 2 * np.pi, nvis)
    u, v = b * np.cos(th), b * np.sin(th)
    nr = 400
    Rmin = 1e-4 * arcsec
    dR = (rmax * dxy - Rmin) / nr * (1 - 1e-9)       # ceil((Rmin + nr dR) / dxy) = rmax
    assert math.ceil((Rmin + nr * dR) / dxy) == rmax
    R = Rmin + dR * np.arange(nr)
    I = 1e10 * (np.exp(-((R - 0.3 * R[-1]) / (0.2 * R[-1])) ** 2) + 0.2 * np.exp(-R / R[-1]))
    inc, PA, dRA, dDec = 0.7, 0.4, 0.02 * arcsec, -0.03 * arcsec
    ref = O.sample_profile(I, Rmin, dR, dxy, nxy, inc, dRA, dDec, duv, PA, u, v, check_central=False)
    got = gD.sampleProfile(I, Rmin, dR, nxy, dxy, u, v, inc=inc, PA=PA, dRA=dRA, dDec=dDec,
                           check_central_pixel=False)
    assert_vis_close(got, ref, 1e-12)
    re, im = rng.standard_normal(nvis), rng.standard_normal(nvis)
    w = rng.uniform(0.5, 2.0, nvis)
    c_ref = O.reduce_chi2(re, im, ref, w)
    c = gD.chi2Profile(I, Rmin, dR, nxy, dxy, u, v, re, im, w, inc=inc, PA=PA, dRA=dRA, dDec=dDec,
                       check_central_pixel=False)
    assert math.isclose(c, c_ref, rel_tol=1e-12)
    if rmax == 255 and nxy <= 4096:
        f = np.float32
        ref32 = O.sample_profile(I.astype(f), f(Rmin), f(dR), f(dxy), nxy, f(inc), f(dRA), f(dDec), f(duv),
                                 f(PA), u.astype(f), v.astype(f), check_central=False, dtype=f)
        got32 = gS.sampleProfile(I.astype(f), f(Rmin), f(dR), nxy, f(dxy), u.astype(f), v.astype(f), inc=f(inc),
                                 PA=f(PA), dRA=f(dRA), dDec=f(dDec), check_central_pixel=False)
        assert_vis_close(got32, ref32, 1e-5)


def test_all_single():
    """Single precision at 1e-5 against the float32 oracle (par1 of python/test_galario.py)."""
    dt = np.float32
    u, v = uv_points(1000, 3e3, dtype=dt)
    maxuv = float(np.hypot(u, v).max()) * 4.8
    dxy = dt(1. / maxuv)
    nxy = 1024
    Rmin, dR, inc = dt(dxy / 100.), dt(dxy / 10.5), dt(20. * deg)
    I = radial_profile(Rmin, dR, 10000, 'Gauss', dxy * 10, dt)
    dRA, dDec, PA = dt(0.), dt(0.4 * arcsec), dt(2. * deg)
    duv = dt(1 / (dxy * nxy))
    ref = O.sample_profile(I, Rmin, dR, dxy, nxy, inc, dRA, dDec, duv, PA, u, v, dtype=dt)
    got = gS.sampleProfile(I, Rmin, dR, nxy, dxy, u, v, dRA=dRA, dDec=dDec, PA=PA, inc=inc)
    assert got.dtype == np.complex64
    assert_vis_close(got, ref, 1e-5)
    img = O.sweep(I, Rmin, dR, nxy, dxy, inc, dtype=dt)
    got_i = gS.sampleImage(img, dxy, u, v, dRA=dRA, dDec=dDec, PA=PA)
    assert_vis_close(got_i, ref, 1e-5)


def test_image_origin():
    """python/test_galario.py:348-427: origin='lower' of the flipped image == origin='upper'."""
    rng = np.random.default_rng(5)
    nxy, dxy = 1024, 6.42956326721e-08
    u, v = uv_points(1000, 3e3)
    img = rng.random((nxy, nxy)) * (np.hypot(*np.meshgrid(np.arange(nxy) - 512, np.arange(nxy) - 512)) < 100)
    img_low = np.roll(np.flipud(img), 1, 0).copy()
    up = gD.sampleImage(img, dxy, u, v, dRA=0.5 * arcsec, dDec=-3. * arcsec, PA=10. * deg)
    low = gD.sampleImage(img_low, dxy, u, v, dRA=0.5 * arcsec, dDec=-3. * arcsec, PA=10. * deg, origin='lower')
    assert_vis_close(low, up, 1e-10)
    ref = O.sample_image(img_low, -1.0, 0.5 * arcsec, -3. * arcsec, 1 / (nxy * dxy), 10. * deg, u, v)
    assert_vis_close(low, ref, 1e-12)


# ---------------------------------------------------------------- workloads at full size
def test_profile_workload_double_4096():
    """Profile fit at nxy=4096, Nvis=1e6 (guard off), double: 1e-12 against the oracle."""
    def model(wl):
        return O.sample_profile(wl.intensity, wl.Rmin, wl.dR, wl.dxy, wl.nxy, wl.inc, wl.dRA, wl.dDec, wl.duv,
                                wl.PA, wl.u, wl.v, check_central=False)
    wl = syn.ring_profile_workload(model, nvis=1_000_000, nxy=4096, seed=1)
    ref_vis = model(wl)
    got_vis = gD.sampleProfile(wl.intensity, wl.Rmin, wl.dR, wl.nxy, wl.dxy, wl.u, wl.v, check_central_pixel=False,
                               **wl.kwargs())
    assert_vis_close(got_vis, ref_vis, 1e-12)
    ref = O.reduce_chi2(wl.re, wl.im, ref_vis, wl.w)
    got = gD.chi2Profile(*wl.args(), check_central_pixel=False, **wl.kwargs())
    assert math.isclose(got, ref, rel_tol=1e-12)
    # the error of the whole path against a direct Fourier sum, at 2000 points (reported by -s)
    img = O.sweep(wl.intensity, wl.Rmin, wl.dR, wl.nxy, wl.dxy, wl.inc, check_central=False)
    sel = slice(0, 2000)
    dft = O.direct_dft_vis(img, wl.dxy, wl.u[sel], wl.v[sel], wl.dRA, wl.dDec, wl.PA)
    print(f"\n|V_gpu - V_dft|max / max|V| = {np.abs(got_vis[sel] - dft).max() / np.abs(dft).max():.3e}")


def test_profile_workload_single_4096():
    dt = np.float32

    def model(wl):
        return O.sample_profile(wl.intensity, wl.Rmin, wl.dR, wl.dxy, wl.nxy, wl.inc, wl.dRA, wl.dDec, wl.duv,
                                wl.PA, wl.u, wl.v, check_central=False)
    wl = syn.ring_profile_workload(model, nvis=1_000_000, nxy=4096, seed=1)
    f = lambda a: np.ascontiguousarray(a, dtype=dt)  # noqa: E731
    I, u, v, re, im, w = map(f, (wl.intensity, wl.u, wl.v, wl.re, wl.im, wl.w))
    sc = {k: dt(x) for k, x in wl.kwargs().items()}
    Rmin, dR, dxy = dt(wl.Rmin), dt(wl.dR), dt(wl.dxy)
    duv = dt(1 / (dxy * wl.nxy))
    ref_vis = O.sample_profile(I, Rmin, dR, dxy, wl.nxy, sc['inc'], sc['dRA'], sc['dDec'], duv, sc['PA'], u, v,
                               dtype=dt, check_central=False)
    got_vis = gS.sampleProfile(I, Rmin, dR, wl.nxy, dxy, u, v, check_central_pixel=False, **sc)
    assert_vis_close(got_vis, ref_vis, 1e-5)
    ref = O.reduce_chi2(re, im, ref_vis, w, dtype=dt)
    got = gS.chi2Profile(I, Rmin, dR, wl.nxy, dxy, u, v, re, im, w, check_central_pixel=False, **sc)
    assert math.isclose(got, ref, rel_tol=1e-5)


def test_image_workload_double_1024():
    """Image fit at nxy=1024, Nvis=1e5, double."""
    def model(wl):
        return O.sample_image(wl.image, 1.0, wl.dRA, wl.dDec, 1 / (wl.dxy * wl.image.shape[0]), wl.PA, wl.u, wl.v)
    wl = syn.ring_image_workload(model, nvis=100_000, nxy=1024, seed=0)
    ref_vis = model(wl)
    got_vis = gD.sampleImage(wl.image, wl.dxy, wl.u, wl.v, dRA=wl.dRA, dDec=wl.dDec, PA=wl.PA)
    assert_vis_close(got_vis, ref_vis, 1e-12)
    ref = O.reduce_chi2(wl.re, wl.im, ref_vis, wl.w)
    got = gD.chi2Image(wl.image, wl.dxy, wl.u, wl.v, wl.re, wl.im, wl.w, dRA=wl.dRA, dDec=wl.dDec, PA=wl.PA)
    assert math.isclose(got, ref, rel_tol=1e-12)


def test_golden_direct_dft_nodes():
    """tests/golden/dft_nodes.npz (long-double direct Fourier sum at grid nodes)."""
    g = np.load(os.path.join(HERE, 'golden', 'dft_nodes.npz'))
    N, dxy = int(g['nxy']), float(g['dxy'])
    ref = g['vis_re'] + 1j * g['vis_im']
    got = gD.sampleProfile(g['intensity'], float(g['Rmin']), float(g['dR']), N, dxy, g['u'], g['v'],
                           dRA=float(g['dRA']), dDec=float(g['dDec']), inc=float(g['inc']))
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    assert np.array_equal(gD.sweep(g['intensity'], float(g['Rmin']), float(g['dR']), N, dxy, float(g['inc'])),
                          g['image'])


# ---------------------------------------------------------------- edge cases
def test_edges_and_errors():
    rng = np.random.default_rng(6)
    N, dxy = 256, 0.02 * arcsec
    duv = 1 / (N * dxy)
    img = rng.random((N, N))
    S = O.spectrum_of_image(img)
    # exactly on the last column (|u|/duv = nxy/2, weight q = 0 beyond it): finite and equal to the oracle
    u = np.array([N / 2 * duv, -N / 2 * duv, 3 * duv])
    v = np.array([0.0, 0.0, (N / 2 - 1.5) * duv])
    got = gD.sampleImage(img, dxy, u, v)
    ref = O.sample_spectrum(S, 1.0, 0., 0., duv, 0., u, v)
    assert np.all(np.isfinite(got))
    assert_vis_close(got, ref, 1e-12)
    # exactly on the last row (nxy/2 + v/duv = nxy-1, weight t = 0 beyond it) -- the oracle's flat
    # indexing (like galario's) would read past the buffer there; check against the node value instead
    last = gD.sampleImage(img, dxy, np.array([3 * duv]), np.array([(N / 2 - 1) * duv]))
    assert np.isfinite(last).all()
    assert abs(last[0] - S[N - 1, 3]) <= 1e-12 * np.abs(S).max()
    # one point outside the grid -> ValueError
    with pytest.raises(ValueError, match="outside the Fourier grid"):
        gD.sampleImage(img, dxy, np.array([0.3 * duv, (N / 2 + 0.5) * duv]), np.array([0., 0.]))
    with pytest.raises(ValueError, match="outside the Fourier grid"):
        gD.chi2Image(img, dxy, np.array([0.5 * duv]), np.array([N * duv]), np.ones(1), np.ones(1), np.ones(1))
    # a single point, PA = 0 and no offset (both skip paths)
    one = gD.sampleImage(img, dxy, np.array([5.3 * duv]), np.array([-7.1 * duv]))
    assert_vis_close(one, O.sample_spectrum(S, 1.0, 0., 0., duv, 0., [5.3 * duv], [-7.1 * duv]), 1e-12)


@pytest.mark.parametrize("N", [4, 16, 100, 1000])
def test_generic_sizes_sample_image(N):
    """Even sizes outside the workgroup FFT (not a power of two, or < 128): direct-DFT path.
    Points lie within radius nxy/2 - 1.05 cells, so they stay on the grid for any PA."""
    rng = np.random.default_rng(N)
    dxy = 0.05 * arcsec
    duv = 1 / (N * dxy)
    img = rng.random((N, N))
    n = 500
    b = rng.uniform(0, N / 2 - 1.05, n) * duv
    th = rng.uniform(0, 2 * np.pi, n)
    u, v = b * np.cos(th), b * np.sin(th)
    got = gD.sampleImage(img, dxy, u, v, dRA=0.01 * arcsec, PA=0.3)
    ref = O.sample_image(img, 1.0, 0.01 * arcsec, 0., duv, 0.3, u, v)
    assert_vis_close(got, ref, 1e-12)


@pytest.mark.parametrize("nxy", [512, 4096])
def test_central_pixel_with_many_inner_bins(nxy):
    """A fine profile (dR = dxy/700) puts ~350 profile bins inside the central
    pixel, more than the row pass has threads (64 at nxy=512, 256 at 4096):
    the central value then reads some bins from LDS and the rest from the
    profile in memory.  sampleProfile and chi2Profile against the oracle at
    1e-12, as in test_all_double."""
    nsamples = 2000
    u, v = uv_points(nsamples, 3e3)
    dxy = 1. / (np.hypot(u, v).max() * 4.8)
    Rmin, dR, nrad, inc = dxy / 1000., dxy / 700., 3000, 30. * deg
    assert math.floor((dxy / 2 - Rmin) / dR) > 256
    dRA, dDec, PA = 1.1 * arcsec, -0.7 * arcsec, 35. * deg
    I = radial_profile(Rmin, dR, nrad, 'Gauss', dxy * 2)
    duv = 1 / (dxy * nxy)
    img = O.sweep(I, Rmin, dR, nxy, dxy, inc)
    ref = O.sample_image(img, 1.0, dRA, dDec, duv, PA, u, v)
    vis = gD.sampleProfile(I, Rmin, dR, nxy, dxy, u, v, dRA=dRA, dDec=dDec, PA=PA, inc=inc)
    assert_vis_close(vis, ref, 1e-12)
    rng = np.random.default_rng(6)
    re, im = ref.real + rng.standard_normal(nsamples), ref.imag + rng.standard_normal(nsamples)
    w = rng.uniform(0.5, 2.0, nsamples)
    c = gD.chi2Profile(I, Rmin, dR, nxy, dxy, u, v, re, im, w, dRA=dRA, dDec=dDec, PA=PA, inc=inc)
    assert math.isclose(c, O.reduce_chi2(re, im, ref, w), rel_tol=1e-12)


def test_profile_larger_than_image():
    """rmax capped at nxy/2: the swept box is the whole image."""
    Rmin, dR, nr, N, dxy = 1e-4 * arcsec, 0.002 * arcsec, 3000, 256, 0.01 * arcsec
    I = radial_profile(Rmin, dR, nr, 'Gauss', 2.0 * arcsec)
    assert O.sweep_rmax(nr, Rmin, dR, N, dxy) == N // 2
    img = gD.sweep(I, Rmin, dR, N, dxy, 0.2, check_central_pixel=False)
    assert np.array_equal(img, O.sweep(I, Rmin, dR, N, dxy, 0.2, check_central=False))
    u, v = uv_points(2000, 0.45 / dxy)
    got = gD.sampleProfile(I, Rmin, dR, N, dxy, u, v, inc=0.2, PA=0.5, check_central_pixel=False)
    ref = O.sample_profile(I, Rmin, dR, dxy, N, 0.2, 0., 0., 1 / (N * dxy), 0.5, u, v, check_central=False)
    assert_vis_close(got, ref, 1e-12)


def test_uvtable_head_realistic_data():
    """First 2000 rows of galario's docs/uvtable.txt (tests/golden/uvtable_head.npz)."""
    g = np.load(os.path.join(HERE, 'golden', 'uvtable_head.npz'))
    u, v, re, im, w = g['u'], g['v'], g['re'], g['im'], g['w']
    nxy, dxy = gD.get_image_size(u, v)
    Rmin, dR, nr = 1e-4 * arcsec, 0.0005 * arcsec, 4000
    I = 1e10 * np.exp(-((Rmin + dR * np.arange(nr)) / (0.5 * arcsec)) ** 2)
    duv = 1 / (nxy * dxy)
    got = gD.chi2Profile(I, Rmin, dR, nxy, dxy, u, v, re, im, w, inc=0.5, PA=1.0, dRA=0.1 * arcsec)
    vis = O.sample_profile(I, Rmin, dR, dxy, nxy, 0.5, 0.1 * arcsec, 0., duv, 1.0, u, v)
    assert math.isclose(got, O.reduce_chi2(re, im, vis, w), rel_tol=1e-12)
