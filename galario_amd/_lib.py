"""ctypes binding of the C ABI in include/galario_amd.h (libgalario_amd.so).

The library is loaded from this package directory; there is no fallback: if
it is missing, importing galario_amd fails with an ImportError naming the
build command.
"""
from __future__ import annotations

import ctypes
import importlib.util
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libgalario_amd.so')

OK, ERR_INVALID, ERR_ALLOC, ERR_RUNTIME = 0, 1, 2, 3

_c_int = ctypes.c_int
_c_double = ctypes.c_double
_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)


def _preload_torch_hip_runtime() -> None:
    """Use the HIP runtime that ships with PyTorch when PyTorch is installed.

    PyTorch-ROCm wheels carry their own libamdhip64.so (soname
    libamdhip64.so.7, requested by torch as libamdhip64.so).  If this library
    were loaded first it would bind /opt/rocm's copy, torch would later load
    its own, and the process would hold two HIP runtimes (torch then reports
    no GPU).  Loading torch's copy first, globally, makes both share it."""
    spec = importlib.util.find_spec('torch')
    if spec is None or not spec.submodule_search_locations:
        return
    for base in spec.submodule_search_locations:
        path = os.path.join(base, 'lib', 'libamdhip64.so')
        if os.path.exists(path):
            ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
            return


class DatasetF64(ctypes.Structure):
    """galario_dataset_f64 (include/galario_amd.h)."""
    _fields_ = [('nxy', ctypes.c_int), ('dxy', ctypes.c_double), ('duv', ctypes.c_double), ('nd', ctypes.c_int),
                ('u', _vp), ('v', _vp), ('vis_obs_re', _vp), ('vis_obs_im', _vp), ('weights', _vp),
                ('uvmax', ctypes.c_double)]


class DatasetF32(ctypes.Structure):
    """galario_dataset_f32 (include/galario_amd.h)."""
    _fields_ = [('nxy', ctypes.c_int), ('dxy', ctypes.c_float), ('duv', ctypes.c_double), ('nd', ctypes.c_int),
                ('u', _vp), ('v', _vp), ('vis_obs_re', _vp), ('vis_obs_im', _vp), ('weights', _vp),
                ('uvmax', ctypes.c_double)]


BATCH_EAGER = 1
BATCH_MODE_GRAD, BATCH_MODE_F64_PHASE = 1, 2


def _load():
    _preload_torch_hip_runtime()
    if not os.path.exists(LIB_PATH):
        raise ImportError(f'galario_amd: {LIB_PATH} is missing; build it with '
                          f'`python galario_amd/build.py` (hipcc, gfx950)')
    lib = ctypes.CDLL(LIB_PATH)
    lib.galario_last_error.restype = ctypes.c_char_p
    lib.galario_last_error.argtypes = []
    lib.galario_init.argtypes = []
    lib.galario_cleanup.argtypes = []
    lib.galario_threads.argtypes = [_c_int]
    lib.galario_threads.restype = _c_int
    lib.galario_ngpus.argtypes = [_ip]
    lib.galario_use_gpu.argtypes = [_c_int]
    lib.galario_free.argtypes = [_vp]
    lib.galario_free.restype = None
    lib.galario_take_oob_flag.argtypes = [_ip]
    lib.galario_read_probe.argtypes = [_vp, _vp, ctypes.c_longlong, _c_int, _vp]
    lib.galario_stage_timer.argtypes = [_c_int]
    lib.galario_stage_repeat.argtypes = [_c_int, _c_int]
    lib.galario_stage_times.argtypes = [ctypes.POINTER(ctypes.c_double), _c_int, _ip]
    lib.galario_batch_destroy.argtypes = [_vp]
    lib.galario_batch_info.argtypes = [_vp, _c_int, _c_int, _ip, _ip, _ip]
    lib.galario_batch_set_chunk.argtypes = [_vp, _c_int, _c_int, _c_int, _c_int]
    lib.galario_batch_tile.argtypes = [_vp, _c_int, _c_int, _ip, _ip]
    lib.galario_batch_set_tile.argtypes = [_vp, _c_int, _c_int, _c_int, _c_int]
    lib.galario_batch_take_oob_flag.argtypes = [_vp, _ip]
    lib.galario_batch_chi2_profile_f32x.argtypes = [_vp, _vp, _c_int, _vp, _vp, _vp, _vp, _c_int]
    lib.galario_batch_chi2_profile_grad_f32x.argtypes = [_vp, _vp, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _c_int]
    for suf, real, ds in (('f64', ctypes.c_double, DatasetF64), ('f32', ctypes.c_float, DatasetF32)):
        R = real
        f = getattr(lib, f'galario_batch_create_{suf}')
        f.argtypes = [_c_int, ctypes.POINTER(ds), _c_int, R, R, _c_int, _c_int, ctypes.POINTER(_vp)]
        f.restype = _c_int
        f = getattr(lib, f'galario_batch_chi2_profile_{suf}')
        f.argtypes = [_vp, _vp, _c_int, _vp, _vp, _vp, _vp, _c_int]
        f.restype = _c_int
        f = getattr(lib, f'galario_batch_chi2_profile_grad_{suf}')
        f.argtypes = [_vp, _vp, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _c_int]
        f.restype = _c_int
        sig = {
            'sample_profile': [_c_int, _vp, R, R, R, _c_int, R, R, R, R, R, _c_int, _vp, _vp, _vp, _c_int],
            'sample_image': [_c_int, _c_int, _vp, R, R, R, R, R, _c_int, _vp, _vp, _vp],
            'chi2_profile': [_c_int, _vp, R, R, R, _c_int, R, R, R, R, R, _c_int, _vp, _vp, _vp, _vp, _vp,
                             _c_int, _vp],
            'chi2_image': [_c_int, _c_int, _vp, R, R, R, R, R, _c_int, _vp, _vp, _vp, _vp, _vp, _vp],
            'sweep': [_c_int, _vp, R, R, _c_int, R, R, _vp, _c_int],
            'uv_rotate': [R, R, R, _vp, _vp, _c_int, _vp, _vp, _vp, _vp],
            'copy_input': [_c_int, _c_int, _vp, ctypes.POINTER(_vp)],
            'fft2d': [_c_int, _c_int, _vp],
            'fftshift': [_c_int, _c_int, _vp],
            'fftshift_axis0': [_c_int, _c_int, _vp],
            'interpolate': [_c_int, _c_int, _vp, R, _c_int, _vp, _vp, R, _vp],
            'apply_phase_sampled': [R, R, _c_int, _vp, _vp, _vp],
            'reduce_chi2': [_c_int, _vp, _vp, _vp, _vp, _vp],
            # the device-resident forms end with uvmax (double; <= 0: no uv extent)
            'chi2_profile_dev': [_vp, _c_int, _vp, R, R, R, _c_int, R, R, R, R, R, _c_int, _vp, _vp, _vp, _vp,
                                 _vp, _c_int, _vp, _c_double],
            'chi2_profile_noalloc_dev': [_vp, _c_int, _vp, R, R, R, _c_int, R, R, R, R, R, _c_int, _vp, _vp, _vp,
                                         _vp, _vp, _c_int, _vp, _c_double],
            'chi2_profile_grad_dev': [_vp, _c_int, _vp, R, R, R, _c_int, R, R, R, R, R, _c_int, _vp, _vp, _vp,
                                      _vp, _vp, _c_int, _vp, _vp, _vp, _c_double],
            'sample_profile_dev': [_vp, _c_int, _vp, R, R, R, _c_int, R, R, R, R, R, _c_int, _vp, _vp, _c_int,
                                   _vp, _vp, _c_double],
            'sample_profile_multi_dev': [_vp, _c_int, _vp, R, R, R, _c_int, R, R, R, _c_int, _vp, _c_int, _vp, _vp,
                                         _c_int, _vp, _vp, _c_double],
            'chi2_profile_multi_dev': [_vp, _c_int, _vp, R, R, R, _c_int, R, R, R, _c_int, _vp, _c_int, _vp, _vp,
                                       _vp, _vp, _vp, _c_int, _vp, _c_double],
        }
        for name, args in sig.items():
            f = getattr(lib, f'galario_{name}_{suf}')
            f.argtypes = args
            f.restype = _c_int
    return lib


lib = _load()


def check(status: int) -> None:
    """Raise the Python exception galario's `except +` mapping gives
    (python/galario_defs.pxd:22-46, docs/py-api.rst:72-82)."""
    if status == OK:
        return
    msg = lib.galario_last_error().decode(errors='replace')
    if status == ERR_INVALID:
        raise ValueError(msg)
    if status == ERR_ALLOC:
        raise MemoryError(msg)
    raise RuntimeError(msg)
