"""CPU baseline: galario's CPU chi2Profile path, C/OpenMP loops + scipy.fft.

TEST INFRASTRUCTURE ONLY (the checker / the timed CPU comparison).

Stages and data layout follow ``sample_profile`` / ``sample_h`` /
``chi2_profile`` of ``src/galario.cpp`` (1404-1436, 1324-1350, 1615-1656):
a fresh padded nxy x 2(nxy/2+1) buffer per call, memset + sweep + central
pixel, quadrant shift, in-place R2C FFT, row shift, rotation, interpolation,
phase, chi2.  FFTW is absent from this image, so the R2C FFT is
``scipy.fft.rfft2(workers=threads)`` -- the engine the CPU baseline names.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np
import scipy.fft

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, '_build', 'libgalario_port.so')
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def build() -> str:
    """Compile the C restatement (make -C oracle); returns the library path."""
    subprocess.run(['make', '-s', '-C', _HERE], check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        L.port_threads.argtypes = [ctypes.c_int]
        L.port_threads.restype = ctypes.c_int
        L.port_alloc_image.argtypes = [ctypes.c_int]
        L.port_alloc_image.restype = ctypes.c_void_p
        L.port_free.argtypes = [ctypes.c_void_p]
        L.port_create_image.argtypes = [ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
        L.port_shift.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.port_shift_axis0.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.port_uv_rotate.argtypes = [ctypes.c_double] * 3 + [_dp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.port_interpolate.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_double,
                                       ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_void_p]
        L.port_apply_phase.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp,
                                       ctypes.c_void_p]
        L.port_reduce_chi2.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_void_p, _dp]
        L.port_reduce_chi2.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def threads(n: int = 0) -> int:
    return lib().port_threads(int(n))


def sample_profile(intensity, Rmin, dR, dxy, nxy, inc, dRA, dDec, duv, PA, u, v, workers=None):
    """Visibilities of a radial profile (double), reference stage order."""
    L = lib()
    workers = workers or threads()
    intensity = np.ascontiguousarray(intensity, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    nd = len(u)
    C = nxy // 2 + 1
    buf = L.port_alloc_image(nxy)
    try:
        L.port_create_image(len(intensity), _p(intensity), Rmin, dR, nxy, dxy, inc, buf)
        L.port_shift(nxy, nxy, buf)
        padded = np.ctypeslib.as_array((ctypes.c_double * (nxy * 2 * C)).from_address(buf))
        padded = padded.reshape(nxy, 2 * C)
        spec = scipy.fft.rfft2(padded[:, :nxy], workers=workers)
        padded.view(np.complex128)[:, :] = spec          # in-place result, FFTW layout
        L.port_shift_axis0(nxy, C, buf)
        urot = np.empty(nd)
        vrot = np.empty(nd)
        dRAr = ctypes.c_double()
        dDecr = ctypes.c_double()
        L.port_uv_rotate(PA, dRA, dDec, ctypes.byref(dRAr), ctypes.byref(dDecr), nd,
                         _p(u), _p(v), _p(urot), _p(vrot))
        vis = np.empty(nd, dtype=np.complex128)
        L.port_interpolate(nxy, C, buf, 1.0, nd, _p(urot), _p(vrot), duv,
                           vis.ctypes.data_as(ctypes.c_void_p))
        L.port_apply_phase(dRAr.value, dDecr.value, nd, _p(urot), _p(vrot),
                           vis.ctypes.data_as(ctypes.c_void_p))
        return vis
    finally:
        L.port_free(buf)


def chi2_profile(intensity, Rmin, dR, dxy, nxy, inc, dRA, dDec, duv, PA, u, v, re, im, w,
                 workers=None):
    """chi2 of a radial profile (double), reference stage order."""
    vis = sample_profile(intensity, Rmin, dR, dxy, nxy, inc, dRA, dDec, duv, PA, u, v, workers)
    re = np.ascontiguousarray(re, dtype=np.float64)
    im = np.ascontiguousarray(im, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    return lib().port_reduce_chi2(len(vis), _p(re), _p(im), vis.ctypes.data_as(ctypes.c_void_p), _p(w))
