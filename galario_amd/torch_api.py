"""Device-resident entry points on torch tensors (PyTorch-ROCm as plumbing).

Arrays are torch tensors already on the GPU (float64 for the double path,
float32 for the single path); work is enqueued on torch's current stream
through the C ABI's ``_dev`` functions and nothing is copied to the host.
Out-of-grid (u, v) points make the results NaN; ``take_oob_flag()``
synchronises and reports whether that happened since the last call.
"""
from __future__ import annotations

import ctypes

import torch

from ._lib import check as _check, lib

_SUF = {torch.float64: 'f64', torch.float32: 'f32'}


def _stream_ptr(device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _p(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _check_tensors(dtype, *ts):
    dev = ts[0].device
    if dev.type != 'cuda':
        raise ValueError('galario_amd.torch_api expects tensors on a GPU device')
    for t in ts:
        if t.dtype != dtype:
            raise ValueError(f'expected {dtype} tensors, got {t.dtype}')
        if t.device != dev:
            raise ValueError('all tensors must be on the same device')
        if not t.is_contiguous():
            raise ValueError('tensors must be contiguous')
    return dev


def chi2_profile(intensity, Rmin, dR, nxy, dxy, u, v, vis_obs_re, vis_obs_im, vis_obs_w, dRA=0., dDec=0., PA=0.,
                 inc=0., check_central_pixel=True, out=None):
    """chi2Profile on device tensors; returns a 1-element float64 tensor (device)."""
    dt = intensity.dtype
    dev = _check_tensors(dt, intensity, u, v, vis_obs_re, vis_obs_im, vis_obs_w)
    nd = u.numel()
    if not (v.numel() == nd == vis_obs_re.numel() == vis_obs_im.numel() == vis_obs_w.numel()):
        raise ValueError('u, v, vis_obs_re, vis_obs_im, vis_obs_w must have the same length')
    if out is None:
        out = torch.empty(1, dtype=torch.float64, device=dev)
    f = getattr(lib, f'galario_chi2_profile_dev_{_SUF[dt]}')
    with torch.cuda.device(dev):
        _check(f(_stream_ptr(dev), intensity.numel(), _p(intensity), Rmin, dR, dxy, int(nxy), inc, dRA, dDec,
                 1.0 / (dxy * nxy), PA, nd, _p(u), _p(v), _p(vis_obs_re), _p(vis_obs_im), _p(vis_obs_w),
                 int(bool(check_central_pixel)), _p(out)))
    return out


def sample_profile(intensity, Rmin, dR, nxy, dxy, u, v, dRA=0., dDec=0., PA=0., inc=0., check_central_pixel=True,
                   out=None):
    """sampleProfile on device tensors; returns a complex tensor (device)."""
    dt = intensity.dtype
    dev = _check_tensors(dt, intensity, u, v)
    nd = u.numel()
    if v.numel() != nd:
        raise ValueError('u and v must have the same length')
    if out is None:
        out = torch.empty(nd, dtype=torch.complex128 if dt == torch.float64 else torch.complex64, device=dev)
    f = getattr(lib, f'galario_sample_profile_dev_{_SUF[dt]}')
    with torch.cuda.device(dev):
        _check(f(_stream_ptr(dev), intensity.numel(), _p(intensity), Rmin, dR, dxy, int(nxy), inc, dRA, dDec,
                 1.0 / (dxy * nxy), PA, nd, _p(u), _p(v), int(bool(check_central_pixel)), _p(out)))
    return out


def take_oob_flag() -> bool:
    """Synchronise and return (and clear) whether a point fell outside the grid."""
    out = ctypes.c_int(0)
    _check(lib.galario_take_oob_flag(ctypes.byref(out)))
    return bool(out.value)


STAGES = ('row_pass', 'column_pass', 'sample', 'final_sum')


def stage_timer(enable: bool) -> None:
    """Start (True) / stop (False) per-stage HIP-event timing of this thread's device calls."""
    _check(lib.galario_stage_timer(int(bool(enable))))


def stage_times():
    """(summed milliseconds per stage as a dict, number of timed calls)."""
    ms = (ctypes.c_double * len(STAGES))()
    n = ctypes.c_int(0)
    _check(lib.galario_stage_times(ms, len(STAGES), ctypes.byref(n)))
    return dict(zip(STAGES, list(ms))), n.value
