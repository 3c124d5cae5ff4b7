"""CPU tests of the boundary: the C ABI exports, and the host-side logic of the
Python API (argument checks, error types and messages) -- nothing here
launches GPU work; every case fails or returns before the library touches a
device."""
import ctypes
import os
import re

import numpy as np
import pytest

import galario_amd
from galario_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'galario_amd.h')


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(galario_[a-z0-9_]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert len(names) >= 38
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_package_layout_mirrors_galario():
    """python/__init__.py.in:20-31 and python/libcommon.pyx:32-39."""
    assert galario_amd.HAVE_CUDA is False and galario_amd.HAVE_HIP is True
    for mod in (galario_amd.double, galario_amd.single):
        for name in mod.__all__:
            assert hasattr(mod, name), name
    assert galario_amd.arcsec == 4.84813681109536e-06 and galario_amd.deg == 0.017453292519943295
    assert galario_amd.double.real_dtype is np.float64 and galario_amd.single.real_dtype is np.float32


@pytest.mark.parametrize("mod", [galario_amd.double, galario_amd.single], ids=['DP', 'SP'])
def test_exception(mod):
    """python/test_galario.py:617-628."""
    dt = mod.real_dtype
    with pytest.raises(ValueError, match="dimension.*is less than 2"):
        mod._fft2d(np.ones((1, 1), dtype=dt))
    with pytest.raises(ValueError, match="Expect a square image"):
        mod._fft2d(np.ones((10, 12), dtype=dt))
    with pytest.raises(ValueError, match="dimension.*is odd"):
        mod._fft2d(np.ones((9, 9), dtype=dt))


def test_input_checks_before_any_gpu_work():
    g = galario_amd.double
    a = g.arcsec
    u = np.linspace(1e3, 1e4, 10)
    I = np.ones(100)
    with pytest.raises(ValueError, match="is odd"):
        g.sampleProfile(I, 1e-6 * a, 0.001 * a, 1023, 0.2 * a, u, u)
    with pytest.raises(ValueError, match=r"Expect \(dxy/2-Rmin\)/dR > 5"):
        g.sampleProfile(I, 1e-4 * a, 0.002 * a, 4096, 0.01 * a, u, u)
    with pytest.raises(ValueError, match=r"Expect \(dxy/2-Rmin\)/dR > 5"):
        g.chi2Profile(I, 1e-4 * a, 0.002 * a, 4096, 0.01 * a, u, u, u, u, u)
    with pytest.raises(ValueError, match="central pixel needs"):
        g.sweep(np.ones(3), 1e-6 * a, 0.001 * a, 64, 0.2 * a)
    with pytest.raises(AssertionError, match="expect Rmin < dxy"):
        g.sampleProfile(I, 0.3 * a, 0.001 * a, 64, 0.2 * a, u, u)
    with pytest.raises(AssertionError, match="Expect origin='upper' or 'lower'"):
        g.sampleImage(np.ones((8, 8)), 0.1 * a, u, u, origin='middle')
    with pytest.raises(AssertionError, match="Wrong array length: vis_obs_im"):
        g.chi2Profile(I, 1e-6 * a, 0.001 * a, 64, 0.2 * a, u, u, u, u[:5], u)
    with pytest.raises(ValueError, match="Expect a square image"):
        g.sampleImage(np.ones((8, 10)), 0.1 * a, u, u)
    with pytest.raises(ValueError, match="Expect a square image"):
        g.chi2Image(np.ones((8, 10)), 0.1 * a, u, u, u, u, u)


def test_buffer_contract_like_cython_memoryviews():
    g = galario_amd.double
    u = np.linspace(1e3, 1e4, 10)
    with pytest.raises(ValueError, match="Buffer dtype mismatch"):
        g.uv_rotate(0.1, 0., 0., u.astype(np.float32), u)
    with pytest.raises(ValueError, match="C-contiguous"):
        g.uv_rotate(0.1, 0., 0., np.linspace(0, 1, 20)[::2], u)
    with pytest.raises(ValueError, match="wrong number of dimensions"):
        g.uv_rotate(0.1, 0., 0., u.reshape(2, 5), u)
    with pytest.raises(IndexError):
        g.reduce_chi2(np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0, complex))
    with pytest.raises(AssertionError):
        galario_amd.single.uv_rotate(0.1, 0., 0., u[:5].astype(np.float32), u.astype(np.float32))


def test_host_helpers():
    """get_image_size, check_image_size, get_coords_meshgrid, set_v_origin (pure numpy in galario too)."""
    g = galario_amd.double
    rng = np.random.default_rng(0)
    b = rng.uniform(1e4, 1e6, 1000); th = rng.uniform(0, 2 * np.pi, 1000)
    u, v = b * np.cos(th), b * np.sin(th)
    nxy, dxy = g.get_image_size(u, v)
    uvd = np.hypot(u, v)
    duv = 1 / (0.6 / uvd.min()) / 5.
    assert nxy == int(2 ** np.ceil(np.log2(uvd.max() * 2 * 2.5 / duv)))
    assert np.isclose(dxy, 1 / (nxy * duv), rtol=1e-15)
    assert g.check_image_size(u, v, nxy, dxy, 1 / (nxy * dxy))
    assert g.set_v_origin('upper') == 1. and g.set_v_origin('lower') == -1.


@pytest.mark.parametrize("nxy, inc, dxy, Dx, Dy, mod, tol",
                         [(1000, 20., 2e-3, -2., 0., galario_amd.single, 1e-6),
                          (1000, 33.4, 1e-8, 0.23, -1.23, galario_amd.double, 1e-15)], ids=['SP', 'DP'])
def test_get_coords_meshgrid(nxy, inc, dxy, Dx, Dy, mod, tol):
    """python/test_galario.py:631-656."""
    real = mod.real_dtype
    x = (np.linspace(0.5, -0.5 + 1. / float(nxy), nxy, dtype=real)) * dxy * nxy
    y = x.copy()
    x_m, y_m = np.meshgrid((x - Dx) / np.cos(inc), y - Dy)
    R_m = np.sqrt(x_m ** 2. + y_m ** 2.)
    got = mod.get_coords_meshgrid(nxy, nxy, dxy, inc, Dx=Dx, Dy=Dy, origin='upper')
    for a, b in zip((x, y, x_m, y_m, R_m), got):
        np.testing.assert_allclose(a, b, atol=0, rtol=tol)


def test_threads_setting():
    g = galario_amd.double
    t0 = g.threads()
    assert g.threads(8) == 8 and g.threads() == 8
    g.threads(t0)


def test_batch_plan_argument_checks_before_any_gpu_work():
    """The batched form validates its datasets on the host (no GPU needed)."""
    from galario_amd import _lib
    a = galario_amd.arcsec
    create = _lib.lib.galario_batch_create_f64
    plan = ctypes.c_void_p()

    # No device pointer is ever passed: the datasets have no points (null arrays are valid then), except the
    # one whose null arrays are the error under test.  A check that regressed would therefore fail on a later
    # host check or plan no work, never hand a kernel a bogus address.
    def ds(nxy, dxy, nd=0, ptr=None):
        return _lib.DatasetF64(nxy, dxy, 1.0 / (nxy * dxy), nd, ptr, ptr, ptr, ptr, ptr)

    def status(sets, nr=100, Rmin=1e-4 * a, dR=0.0005 * a, max_w=4, cc=0):
        arr = (_lib.DatasetF64 * max(len(sets), 1))(*sets)
        return create(len(sets), arr, nr, Rmin, dR, max_w, cc, ctypes.byref(plan))

    with pytest.raises(ValueError, match="number of datasets"):
        _lib.check(status([]))
    with pytest.raises(ValueError, match="not supported by the batched form"):
        _lib.check(status([ds(1000, 0.01 * a)]))
    with pytest.raises(ValueError, match="is odd"):
        _lib.check(status([ds(1023, 0.01 * a)]))
    with pytest.raises(ValueError, match="null array"):
        _lib.check(status([ds(512, 0.01 * a, nd=10)]))
    with pytest.raises(ValueError, match=r"Expect \(dxy/2-Rmin\)/dR > 5"):
        _lib.check(status([ds(512, 0.01 * a)], dR=0.002 * a, cc=1))
    with pytest.raises(ValueError, match="max_walkers"):
        _lib.check(status([ds(512, 0.01 * a)], max_w=0))
    with pytest.raises(ValueError, match="not a batch plan"):
        _lib.check(_lib.lib.galario_batch_info(ctypes.c_void_p(0), 1, 0, None, None, None))
    with pytest.raises(ValueError, match="not a batch plan"):
        _lib.check(_lib.lib.galario_batch_tile(ctypes.c_void_p(0), 1, 0, None, None))
    with pytest.raises(ValueError, match="not a batch plan"):
        _lib.check(_lib.lib.galario_batch_set_tile(ctypes.c_void_p(0), 1, 0, 1, 3072))


def test_gradient_argument_checks_before_any_gpu_work():
    """The gradient entry point rejects unsupported sizes on the host.  Every array is null: the size checks
    come first, and null arrays are themselves rejected before any GPU work."""
    a = galario_amd.arcsec
    f = _lib.lib.galario_chi2_profile_grad_dev_f64
    p = None
    common = (0.01 * a, 1024, 0.5, 0.0, 0.0, 1 / (1024 * 0.01 * a), 0.3, 10, p, p, p, p, p, 0, p, p, p, 0.0)
    with pytest.raises(ValueError, match="up to 2048"):
        _lib.check(f(None, 3000, p, 1e-4 * a, 0.0005 * a, *common))
    with pytest.raises(ValueError, match="not supported by the gradient"):
        _lib.check(f(None, 1000, p, 1e-4 * a, 0.0005 * a, 0.01 * a, 1000, *common[2:]))
    with pytest.raises(ValueError, match="null output pointer"):
        _lib.check(f(None, 1000, p, 1e-4 * a, 0.0005 * a, *common))


def test_device_entry_points_reject_null_arrays():
    """The _dev forms reject null device arrays on the host (after their size checks)."""
    a = galario_amd.arcsec
    f = _lib.lib.galario_chi2_profile_dev_f64
    args = (None, 1000, None, 1e-4 * a, 0.0005 * a, 0.01 * a, 1024, 0.5, 0.0, 0.0, 1 / (1024 * 0.01 * a), 0.3, 10,
            None, None, None, None, None, 0, None, 0.0)
    with pytest.raises(ValueError, match="null intensity pointer"):
        _lib.check(f(*args))
    with pytest.raises(ValueError, match="is odd"):
        _lib.check(f(*args[:6], 1023, *args[7:]))
    g = _lib.lib.galario_sample_profile_dev_f64
    with pytest.raises(ValueError, match="null intensity pointer"):
        _lib.check(g(None, 1000, None, 1e-4 * a, 0.0005 * a, 0.01 * a, 1024, 0.5, 0.0, 0.0, 1 / (1024 * 0.01 * a),
                     0.3, 10, None, None, 0, None, None, 0.0))


def test_read_probe_argument_checks_before_any_gpu_work():
    """The read-rate reference rejects bad sizes and null buffers on the host."""
    f = _lib.lib.galario_read_probe
    with pytest.raises(ValueError, match="null buffer"):
        _lib.check(f(None, None, 16, 1, None))
    with pytest.raises(ValueError, match="multiple of 16"):
        _lib.check(f(None, None, 24, 1, None))
    with pytest.raises(ValueError, match="grid must be"):
        _lib.check(f(None, None, 32, 0, None))
    with pytest.raises(ValueError, match="grid must be"):
        _lib.check(f(None, None, 32, 70000, None))


def test_galario_import_path():
    """The package name and layout of galario (python/__init__.py.in:20-31, docs/quickstart.rst:72,85,144-145):
    ``galario.double`` / ``galario.single`` are the HIP-backed modules themselves."""
    import galario
    import galario.double
    import galario.single
    from galario import arcsec, au, cgs_to_Jy, deg, pc
    from galario.double import chi2Profile, get_image_size, sampleProfile  # noqa: F401
    assert galario.double is galario_amd.double and galario.single is galario_amd.single
    assert galario.__version__ == '1.2.2' and galario.HAVE_CUDA is False
    assert (arcsec, deg) == (4.84813681109536e-06, 0.017453292519943295)
    assert (au, pc, cgs_to_Jy) == (galario_amd.double.au, galario_amd.double.pc, 1e23)
    assert chi2Profile is galario_amd.double.chi2Profile
    for mod in (galario.double, galario.single):
        assert set(mod.__all__) >= {'sampleImage', 'chi2Image', 'sampleProfile', 'chi2Profile', 'get_image_size'}
