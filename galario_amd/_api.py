"""The galario Python API (python/libcommon.pyx), bound to the HIP library.

``populate(namespace, double)`` defines, in the namespace of
``galario_amd.double`` (double=True) or ``galario_amd.single``, the names of
galario's ``__all__`` (python/libcommon.pyx:32-39) with the same signatures,
defaults, argument checks, messages and exception types.  Array arguments
follow the Cython ``dreal[::1]`` / ``dreal[:, ::1]`` contract: exactly the
module's dtype, C-contiguous, else ValueError.

Extensions (keyword-only, default keeps galario's behaviour):
``check_central_pixel=True`` on sweep / sampleProfile / chi2Profile.  With
False the central-pixel guard (src/galario.cpp:113-121) is skipped; the
central pixel's arithmetic itself is still checked to stay inside the profile.

Points that fall outside the Fourier grid raise ValueError (galario reads
outside its buffer there, src/galario.cpp:749-815 has no bounds check).
"""
from __future__ import annotations

import ctypes
import weakref

import numpy as np

from ._lib import check as _check, lib

# CONSTANTS, python/libcommon.pyx:42-47
arcsec = 4.84813681109536e-06       # radians
deg = 0.017453292519943295          # radians
cgs_to_Jy = 1e23                    # 1 Jy = 1.0e-23 erg/(s cm^2 Hz)
pc = 3.0856775815e18                # cm (IAU 2015 Resolution B2)
au = 1.49597870700e13               # cm (IAU 2012 Resolution B1)

ALL = ['arcsec', 'deg', 'cgs_to_Jy', 'pc', 'au',
       '_init', '_cleanup', 'set_v_origin',
       'ngpus', 'use_gpu', 'threads',
       'check_obs', 'check_image_size', 'get_image_size',
       'sampleImage', 'sampleProfile', 'chi2Image', 'chi2Profile',
       'get_coords_meshgrid',
       'sweep', 'uv_rotate', 'interpolate', 'apply_phase_vis', 'reduce_chi2',
       '_fft2d', '_fftshift', '_fftshift_axis0']


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def _typed(x, dtype, ndim: int, name: str) -> np.ndarray:
    """Accept what a Cython typed memoryview `dtype[::1]` (ndim 1) or
    `dtype[:, ::1]` (ndim 2) accepts; raise ValueError like Cython otherwise."""
    a = x if isinstance(x, np.ndarray) else np.asarray(memoryview(x))
    if a.ndim != ndim:
        raise ValueError(f"Buffer has wrong number of dimensions (expected {ndim}, got {a.ndim})")
    if a.dtype != dtype:
        exp = {np.dtype(np.float64): 'double', np.dtype(np.float32): 'float',
               np.dtype(np.complex128): 'double complex', np.dtype(np.complex64): 'float complex'}[np.dtype(dtype)]
        raise ValueError(f"Buffer dtype mismatch, expected '{exp}' but got '{a.dtype}' for argument '{name}'")
    if not a.flags.c_contiguous:
        raise ValueError("ndarray is not C-contiguous")
    return a


def _first(a: np.ndarray) -> None:
    """`&a[0]` of a Cython memoryview raises IndexError on an empty buffer."""
    if a.shape[0] == 0:
        raise IndexError("Out of bounds on buffer access (axis 0)")


def populate(ns: dict, double: bool) -> None:
    real = np.float64 if double else np.float32
    cplx = np.complex128 if double else np.complex64
    suf = 'f64' if double else 'f32'
    cR = ctypes.c_double if double else ctypes.c_float
    L = {name: getattr(lib, f'galario_{name}_{suf}') for name in (
        'sample_profile', 'sample_image', 'chi2_profile', 'chi2_image', 'sweep', 'uv_rotate',
        'copy_input', 'fft2d', 'fftshift', 'fftshift_axis0', 'interpolate',
        'apply_phase_sampled', 'reduce_chi2')}

    def _init():
        """Initialises the library (python/libcommon.pyx:97-99)."""
        _check(lib.galario_init())

    def _cleanup():
        """Releases this thread's GPU workspaces (python/libcommon.pyx:102-104)."""
        _check(lib.galario_cleanup())

    def set_v_origin(origin):
        """Sets the Dec (v) axis orientation given the matrix origin (python/libcommon.pyx:107-116)."""
        if origin == 'upper':
            return 1.
        elif origin == 'lower':
            return -1.
        else:
            raise AssertionError("Expect origin='upper' or 'lower', got {}".format(origin))

    def ngpus():
        """Number of GPUs visible to HIP (python/libcommon.pyx:125-135)."""
        n = ctypes.c_int(0)
        _check(lib.galario_ngpus(ctypes.byref(n)))
        return n.value

    def use_gpu(device_id):
        """Select the GPU used by the calling thread (python/libcommon.pyx:138-160)."""
        _check(lib.galario_use_gpu(int(device_id)))

    def threads(num=0):
        """Set (num > 0) and get the thread setting (python/libcommon.pyx:162-199).

        Kept for API compatibility: the HIP kernels use launch shapes fixed per
        kernel, so the value does not change any computation."""
        return lib.galario_threads(int(num))

    def check_obs(vis_obs_re, vis_obs_im, vis_obs_w, vis=None, u=None, v=None):
        """Checks whether the observed visibilities are consistent (python/libcommon.pyx:208-220)."""
        nd = len(vis_obs_re)
        assert len(vis_obs_im) == nd, "Wrong array length: vis_obs_im."
        assert len(vis_obs_w) == nd, "Wrong array length: vis_obs_w."
        if vis is not None and len(vis):
            assert len(vis) == nd, "Wrong array length: vis."
        if u is not None and len(u):
            assert len(u) == nd, "Wrong array length: u"
        if v is not None and len(v):
            assert len(v) == nd, "Wrong array length: v"
        return True

    def check_image_size(u, v, nxy, dxy, duv, PB=0, verbose=False):
        """Nyquist checks of the (u, v) sampling (python/libcommon.pyx:223-288).

        The last two asserts keep galario's semantics: they pass if any point
        is within range; points outside the grid are caught by the GPU kernel."""
        assert len(u) == len(v), "Wrong array length: u, v must have same length."
        uvdist = np.hypot(u, v)
        MRS = 0.6 / np.min(uvdist)
        max_uv = np.max(uvdist) * 2.
        FOV = nxy * dxy
        FOV_to_MRS = FOV / MRS
        UVFOV_to_MAXUV = 1 / (max_uv * dxy)
        FOV_to_MRS_str = "Nxy * dxy / MRS = {} must be > 1 at the very least".format(FOV_to_MRS)
        UVFOV_to_MAXUV_str = "Nxy * duv / (2*max(u,v)) = {} must be > 2 for Nyquist sampling".format(UVFOV_to_MAXUV)
        if PB != 0:
            FOV_to_PB = FOV / PB
            FOV_to_PB_str = "Nxy * dxy / PB = {} must be > 1".format(FOV_to_PB)
        if verbose:
            print(FOV_to_MRS_str)
            print(UVFOV_to_MAXUV_str)
            if PB != 0:
                print(FOV_to_PB_str)
        assert FOV_to_MRS > 1, FOV_to_MRS_str
        assert UVFOV_to_MAXUV > 2, UVFOV_to_MAXUV_str
        if PB != 0:
            assert FOV_to_PB > 1, FOV_to_PB_str
        assert np.max(np.abs(u) / duv <= nxy // 2 + 1)
        assert np.max(np.abs(v) / duv <= nxy // 2)
        return True

    def get_image_size(u, v, PB=0, f_min=5., f_max=2.5, verbose=False):
        """Recommended (nxy, dxy) for the (u, v) points (python/libcommon.pyx:291-361)."""
        uvdist = np.hypot(u, v)
        MRS = 0.6 / np.min(uvdist)
        duv = 1 / MRS / f_min
        max_uv = np.max(uvdist) * 2. * f_max
        nxy = int(2 ** np.ceil(np.log2(max_uv / duv)))
        nxy_MRS = nxy
        dxy = 1 / (nxy * duv)
        if PB != 0:
            while dxy * nxy / PB < 1.:
                nxy *= 2
        if verbose:
            print("dxy:{:e}arcsec\tnxy_MRS:{}".format(dxy / arcsec, nxy_MRS))
            print("nxy_MRS: matrix size to have FOV > f_min * MRS, where f_min:{} and MRS:{:e}arcsec".format(
                f_min, MRS / arcsec))
            if PB != 0:
                print("nxy_FOV:{}".format(nxy))
                print("nxy_FOV: matrix size to have FOV > PB")
        return nxy, dxy

    def sampleImage(image, dxy, u, v, dRA=0., dDec=0., PA=0., check=False, origin='upper'):
        """Synthetic visibilities of a model image at (u, v) (python/libcommon.pyx:371-443)."""
        image = _typed(image, real, 2, 'image')
        u = _typed(u, real, 1, 'u')
        v = _typed(v, real, 1, 'v')
        nxy = image.shape[0]
        duv = 1 / (dxy * nxy)
        if check:
            check_image_size(u, v, nxy, dxy, duv)
        vis = np.zeros(len(u), dtype=cplx)
        v_origin = set_v_origin(origin)
        _first(image); _first(u); _first(v)
        assert len(v) >= len(u), "Wrong array length: v"
        _check(L['sample_image'](image.shape[0], image.shape[1], _ptr(image), v_origin, dRA, dDec, duv, PA, len(u),
                                _ptr(u), _ptr(v), _ptr(vis)))
        return vis

    def sampleProfile(intensity, Rmin, dR, nxy, dxy, u, v, dRA=0., dDec=0., PA=0., inc=0., check=False, *,
                      check_central_pixel=True):
        """Synthetic visibilities of an axisymmetric profile (python/libcommon.pyx:446-530)."""
        intensity = _typed(intensity, real, 1, 'intensity')
        u = _typed(u, real, 1, 'u')
        v = _typed(v, real, 1, 'v')
        assert Rmin < dxy, ("For the interpolation of the image center, expect Rmin < dxy, but got Rmin={}, "
                            "dxy={}".format(Rmin, dxy))
        duv = 1 / (dxy * nxy)
        if check:
            check_image_size(u, v, nxy, dxy, duv)
        vis = np.zeros(len(u), dtype=cplx)
        _first(intensity); _first(u); _first(v)
        assert len(v) >= len(u), "Wrong array length: v"
        _check(L['sample_profile'](len(intensity), _ptr(intensity), Rmin, dR, dxy, int(nxy), inc, dRA, dDec, duv, PA,
                                  len(u), _ptr(u), _ptr(v), _ptr(vis), int(bool(check_central_pixel))))
        return vis

    def chi2Image(image, dxy, u, v, vis_obs_re, vis_obs_im, vis_obs_w, dRA=0., dDec=0., PA=0., check=False,
                  origin='upper'):
        """chi2 of a model image against observed visibilities (python/libcommon.pyx:533-626)."""
        image = _typed(image, real, 2, 'image')
        u = _typed(u, real, 1, 'u')
        v = _typed(v, real, 1, 'v')
        vis_obs_re = _typed(vis_obs_re, real, 1, 'vis_obs_re')
        vis_obs_im = _typed(vis_obs_im, real, 1, 'vis_obs_im')
        vis_obs_w = _typed(vis_obs_w, real, 1, 'vis_obs_w')
        check_obs(vis_obs_re, vis_obs_im, vis_obs_w, u=u, v=v)
        nxy = image.shape[0]
        duv = 1 / (dxy * nxy)
        if check:
            check_image_size(u, v, nxy, dxy, duv)
        v_origin = set_v_origin(origin)
        for a in (image, u, v, vis_obs_re, vis_obs_im, vis_obs_w):
            _first(a)
        out = cR()
        _check(L['chi2_image'](image.shape[0], image.shape[1], _ptr(image), v_origin, dRA, dDec, duv, PA, len(u),
                              _ptr(u), _ptr(v), _ptr(vis_obs_re), _ptr(vis_obs_im), _ptr(vis_obs_w),
                              ctypes.byref(out)))
        return out.value

    def chi2Profile(intensity, Rmin, dR, nxy, dxy, u, v, vis_obs_re, vis_obs_im, vis_obs_w, dRA=0., dDec=0., PA=0.,
                    inc=0., check=False, *, check_central_pixel=True):
        """chi2 of an axisymmetric profile against observed visibilities (python/libcommon.pyx:629-728)."""
        intensity = _typed(intensity, real, 1, 'intensity')
        u = _typed(u, real, 1, 'u')
        v = _typed(v, real, 1, 'v')
        vis_obs_re = _typed(vis_obs_re, real, 1, 'vis_obs_re')
        vis_obs_im = _typed(vis_obs_im, real, 1, 'vis_obs_im')
        vis_obs_w = _typed(vis_obs_w, real, 1, 'vis_obs_w')
        check_obs(vis_obs_re, vis_obs_im, vis_obs_w, u=u, v=v)
        assert Rmin < dxy, ("For the interpolation of the image center, expect Rmin < dxy, but got Rmin={}, "
                            "dxy={}".format(Rmin, dxy))
        duv = 1 / (dxy * nxy)
        if check:
            check_image_size(u, v, nxy, dxy, duv)
        for a in (intensity, u, v, vis_obs_re, vis_obs_im, vis_obs_w):
            _first(a)
        out = cR()
        _check(L['chi2_profile'](len(intensity), _ptr(intensity), Rmin, dR, dxy, int(nxy), inc, dRA, dDec, duv, PA,
                                len(u), _ptr(u), _ptr(v), _ptr(vis_obs_re), _ptr(vis_obs_im), _ptr(vis_obs_w),
                                int(bool(check_central_pixel)), ctypes.byref(out)))
        return out.value

    def get_coords_meshgrid(nrow, ncol, dxy=1., inc=0., Dx=0., Dy=0., origin='upper'):
        """(R.A., Dec.) coordinate mesh grid (python/libcommon.pyx:731-797)."""
        v_origin = set_v_origin(origin)
        x = (np.linspace(0.5, -0.5 + 1. / float(ncol), ncol, dtype=real)) * ncol * dxy
        y = (np.linspace(0.5, -0.5 + 1. / float(nrow), nrow, dtype=real)) * nrow * dxy * v_origin
        x_m, y_m = np.meshgrid((x - Dx) / np.cos(inc), (y - Dy))
        R_m = np.hypot(x_m, y_m)
        return x, y, x_m, y_m, R_m

    def sweep(intensity, Rmin, dR, nxy, dxy, inc=0., *, check_central_pixel=True):
        """2D image of an axisymmetric profile (python/libcommon.pyx:800-851)."""
        intensity = _typed(intensity, real, 1, 'intensity')
        assert Rmin < dxy, "For the interpolation algorithm, Rmin must be smaller than dxy. " \
                           "Currently Rmin={}\t dxy={}".format(Rmin, dxy)
        image = np.empty((nxy, nxy // 2 + 1), dtype=cplx, order='C')
        _first(intensity)
        _check(L['sweep'](len(intensity), _ptr(intensity), Rmin, dR, int(nxy), dxy, inc, _ptr(image),
                         int(bool(check_central_pixel))))
        return np.ascontiguousarray(image.view(dtype=real)[:, :-2])

    def uv_rotate(PA, dRA, dDec, u, v):
        """(u, v) and (dRA, dDec) rotated by the position angle (python/libcommon.pyx:854-906)."""
        u = _typed(u, real, 1, 'u')
        v = _typed(v, real, 1, 'v')
        nd = len(u)
        assert nd == len(v)
        dRArot = cR()
        dDecrot = cR()
        urot = np.copy(u, order='C')
        vrot = np.copy(v, order='C')
        _first(u); _first(v)
        _check(L['uv_rotate'](PA, dRA, dDec, ctypes.byref(dRArot), ctypes.byref(dDecrot), nd, _ptr(u), _ptr(v),
                             _ptr(urot), _ptr(vrot)))
        return dRArot.value, dDecrot.value, urot, vrot

    def interpolate(r2cFT, duv, u, v, origin='upper'):
        """Amplitude/phase interpolation of an R2C spectrum at (u, v) (python/libcommon.pyx:909-950)."""
        r2cFT = _typed(r2cFT, cplx, 2, 'r2cFT')
        u = _typed(u, real, 1, 'u')
        v = _typed(v, real, 1, 'v')
        vis = np.empty(len(u), dtype=cplx, order='C')
        v_origin = set_v_origin(origin)
        _first(r2cFT); _first(u); _first(v)
        assert len(v) >= len(u), "Wrong array length: v"
        _check(L['interpolate'](r2cFT.shape[0], r2cFT.shape[1], _ptr(r2cFT), v_origin, len(u), _ptr(u), _ptr(v), duv,
                               _ptr(vis)))
        return vis

    def apply_phase_vis(dRA, dDec, u, v, vis):
        """Phase shift of sampled visibilities by (dRA, dDec) (python/libcommon.pyx:953-991)."""
        u = _typed(u, real, 1, 'u')
        v = _typed(v, real, 1, 'v')
        vis = _typed(vis, cplx, 1, 'vis')
        vis_out = np.copy(vis, order='C')
        _first(u); _first(v)
        assert len(u) >= len(vis) and len(v) >= len(vis), "Wrong array length: u, v"
        _check(L['apply_phase_sampled'](dRA, dDec, len(vis), _ptr(u), _ptr(v), _ptr(vis_out)))
        return vis_out

    def reduce_chi2(vis_obs_re, vis_obs_im, vis_obs_w, vis):
        """chi2 of observed and model visibilities (python/libcommon.pyx:994-1027)."""
        vis_obs_re = _typed(vis_obs_re, real, 1, 'vis_obs_re')
        vis_obs_im = _typed(vis_obs_im, real, 1, 'vis_obs_im')
        vis_obs_w = _typed(vis_obs_w, real, 1, 'vis_obs_w')
        vis = _typed(vis, cplx, 1, 'vis')
        check_obs(vis_obs_re, vis_obs_im, vis_obs_w, vis)
        for a in (vis_obs_re, vis_obs_im, vis_obs_w, vis):
            _first(a)
        out = cR()
        _check(L['reduce_chi2'](len(vis), _ptr(vis_obs_re), _ptr(vis_obs_im), _ptr(vis), _ptr(vis_obs_w),
                               ctypes.byref(out)))
        return out.value

    def _owned_padded(nx, ny, ptr):
        """ndarray view of a library buffer (nx x (ny/2+1) complex), freed with galario_free."""
        C = ny // 2 + 1
        nbytes = nx * C * np.dtype(cplx).itemsize
        buf = (ctypes.c_char * nbytes).from_address(ptr)
        arr = np.frombuffer(buf, dtype=cplx).reshape(nx, C)
        weakref.finalize(arr, lib.galario_free, ctypes.c_void_p(ptr))
        return arr

    def _copy_input(image):
        nx, ny = image.shape[0], image.shape[1]
        res = ctypes.c_void_p()
        _check(L['copy_input'](nx, ny, _ptr(image), ctypes.byref(res)))
        return _owned_padded(nx, ny, res.value)

    def _fft2d(image):
        """Wrapper for the 2D Real to Complex FFT (python/libcommon.pyx:1030-1039)."""
        image = _typed(image, real, 2, 'image')
        nx, ny = image.shape[0], image.shape[1]
        _first(image)
        res = _copy_input(image)
        _check(L['fft2d'](nx, ny, _ptr(res)))
        res.flags.writeable = True
        return res

    def _fftshift(matrix):
        """Swap of the four quarters of a real matrix, in the R2C padded layout
        (python/libcommon.pyx:1042-1066); equivalent to numpy.fft.fftshift(matrix)."""
        matrix = _typed(matrix, real, 2, 'matrix')
        nx, ny = matrix.shape[0], matrix.shape[1]
        assert nx % 2 == 0 and ny % 2 == 0, "Expect even matrix size but got {}".format(matrix.shape)
        _first(matrix)
        res = _copy_input(matrix)
        _check(L['fftshift'](nx, ny, _ptr(res)))
        return res

    def _fftshift_axis0(matrix):
        """Swap the upper and lower halves of a complex matrix, in place
        (python/libcommon.pyx:1069-1082); equivalent to numpy.fft.fftshift(matrix, axes=0)."""
        matrix = _typed(matrix, cplx, 2, 'matrix')
        assert matrix.shape[0] % 2 == 0, "Axis 0 of `matrix` has to be even "
        _first(matrix)
        _check(L['fftshift_axis0'](matrix.shape[0], matrix.shape[1], _ptr(matrix)))

    local = dict(locals())
    consts = dict(arcsec=arcsec, deg=deg, cgs_to_Jy=cgs_to_Jy, pc=pc, au=au)
    for name in ALL:
        ns[name] = consts[name] if name in consts else local[name]
    ns['real_dtype'] = real
    ns['complex_dtype'] = cplx
    ns['__all__'] = list(ALL)
