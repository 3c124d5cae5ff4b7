"""Device-resident entry points on torch tensors (PyTorch-ROCm as plumbing).

Arrays are torch tensors already on the GPU (float64 for the double path,
float32 for the single path); work is enqueued on torch's current stream
through the C ABI's ``_dev`` functions and nothing is copied to the host.
Out-of-grid (u, v) points make the results NaN; ``take_oob_flag()``
synchronises and reports whether that happened since the last call.

``Dataset`` uploads an observation once, in a locality order: the points are
sorted by the Morton code of their Hermitian-folded position (|u|, sign(u) v),
so that the threads of a wavefront gather neighbouring spectrum cells (a
rotation by PA keeps neighbours together except across the new fold line).
chi2 does not depend on the order; visibilities are returned in the caller's
order (the kernel writes point i to perm[i]).  A Dataset also records its uv
extent (the largest sqrt(u^2 + v^2)); rotation keeps that radius, so its calls
build the spectrum only within it.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

import weakref

from ._lib import BATCH_EAGER, BATCH_MODE_F64_PHASE, BATCH_MODE_GRAD, DatasetF32, DatasetF64, check as _check, lib

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
                 inc=0., check_central_pixel=True, out=None, uvmax=0.):
    """chi2Profile on device tensors; returns a 1-element float64 tensor (device).

    uvmax (> 0): a bound on sqrt(u^2 + v^2); the spectrum is then built only
    within it (points beyond it count as outside the grid)."""
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
                 int(bool(check_central_pixel)), _p(out), float(uvmax)))
    return out


def sample_profile(intensity, Rmin, dR, nxy, dxy, u, v, dRA=0., dDec=0., PA=0., inc=0., check_central_pixel=True,
                   out=None, perm=None, uvmax=0.):
    """sampleProfile on device tensors; returns a complex tensor (device).  uvmax as for chi2_profile."""
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
                 1.0 / (dxy * nxy), PA, nd, _p(u), _p(v), int(bool(check_central_pixel)),
                 _p(perm) if perm is not None else None, _p(out), float(uvmax)))
    return out


def take_oob_flag() -> bool:
    """Synchronise and return (and clear) whether a point fell outside the grid."""
    out = ctypes.c_int(0)
    _check(lib.galario_take_oob_flag(ctypes.byref(out)))
    return bool(out.value)


STAGES = ('row_pass', 'column_pass', 'sample', 'adjoint_column_pass', 'adjoint_row_pass', 'reduce')


def stage_timer(enable: bool) -> None:
    """Start (True) / stop (False) per-stage HIP-event timing of this thread's device calls."""
    _check(lib.galario_stage_timer(int(bool(enable))))


def stage_repeat(stage, reps: int) -> None:
    """While timing, launch one forward kernel ('row_pass', 'column_pass' or 'sample'; None for none)
    `reps` times back to back, so its stage time / reps is its average launch duration."""
    idx = -1 if stage is None else STAGES.index(stage)
    _check(lib.galario_stage_repeat(idx, int(reps)))


def read_probe(buf: torch.Tensor, grid: int, out: torch.Tensor) -> None:
    """HBM read-rate reference: `grid` workgroups read all of the contiguous device
    tensor `buf` (a multiple of 16 bytes) and write one sum each to `out`
    (float64, at least `grid` elements), on the current stream."""
    if buf.device.type != 'cuda' or out.device != buf.device or out.dtype != torch.float64:
        raise ValueError('read_probe expects a device buffer and a float64 output on the same device')
    if not (buf.is_contiguous() and out.is_contiguous()) or out.numel() < grid:
        raise ValueError('read_probe expects contiguous tensors and at least `grid` outputs')
    nbytes = buf.numel() * buf.element_size()
    _check(lib.galario_read_probe(_stream_ptr(buf.device), _p(buf), nbytes, int(grid), _p(out)))


def stage_times():
    """(summed milliseconds per stage as a dict, number of timed calls)."""
    ms = (ctypes.c_double * len(STAGES))()
    n = ctypes.c_int(0)
    _check(lib.galario_stage_times(ms, len(STAGES), ctypes.byref(n)))
    return dict(zip(STAGES, list(ms))), n.value


def _sources(sources, dt, dev):
    src = torch.as_tensor(sources, dtype=dt, device=dev).reshape(-1, 3).contiguous()
    return src, src.shape[0]


def sample_profile_multi(intensity, Rmin, dR, nxy, dxy, u, v, sources, PA=0., inc=0., check_central_pixel=True,
                         out=None, perm=None, uvmax=0.):
    """Visibilities of several sources with this profile, inc and PA: ``sources`` [K, 3] = (flux, dRA, dDec);
    sum_k flux_k sampleProfile(..., dRA_k, dDec_k) from one spectrum (complex device tensor)."""
    dt = intensity.dtype
    dev = _check_tensors(dt, intensity, u, v)
    src, k = _sources(sources, dt, dev)
    nd = u.numel()
    if out is None:
        out = torch.empty(nd, dtype=torch.complex128 if dt == torch.float64 else torch.complex64, device=dev)
    f = getattr(lib, f'galario_sample_profile_multi_dev_{_SUF[dt]}')
    with torch.cuda.device(dev):
        _check(f(_stream_ptr(dev), intensity.numel(), _p(intensity), Rmin, dR, dxy, int(nxy), inc, 1.0 / (dxy * nxy),
                 PA, k, _p(src), nd, _p(u), _p(v), int(bool(check_central_pixel)),
                 _p(perm) if perm is not None else None, _p(out), float(uvmax)))
    return out


def chi2_profile_multi(intensity, Rmin, dR, nxy, dxy, u, v, vis_obs_re, vis_obs_im, vis_obs_w, sources, PA=0., inc=0.,
                       check_central_pixel=True, out=None, uvmax=0.):
    """chi2 of the multi-source model (see ``sample_profile_multi``); 1-element float64 device tensor."""
    dt = intensity.dtype
    dev = _check_tensors(dt, intensity, u, v, vis_obs_re, vis_obs_im, vis_obs_w)
    src, k = _sources(sources, dt, dev)
    if out is None:
        out = torch.empty(1, dtype=torch.float64, device=dev)
    f = getattr(lib, f'galario_chi2_profile_multi_dev_{_SUF[dt]}')
    with torch.cuda.device(dev):
        _check(f(_stream_ptr(dev), intensity.numel(), _p(intensity), Rmin, dR, dxy, int(nxy), inc, 1.0 / (dxy * nxy),
                 PA, k, _p(src), u.numel(), _p(u), _p(v), _p(vis_obs_re), _p(vis_obs_im), _p(vis_obs_w),
                 int(bool(check_central_pixel)), _p(out), float(uvmax)))
    return out


def _spread16(x: np.ndarray) -> np.ndarray:
    """The 16 low bits of x moved to the even bit positions of a uint32."""
    x = x.astype(np.uint32) & np.uint32(0xFFFF)
    x = (x | (x << np.uint32(8))) & np.uint32(0x00FF00FF)
    x = (x | (x << np.uint32(4))) & np.uint32(0x0F0F0F0F)
    x = (x | (x << np.uint32(2))) & np.uint32(0x33333333)
    x = (x | (x << np.uint32(1))) & np.uint32(0x55555555)
    return x


def locality_key(u, v) -> np.ndarray:
    """The 32-bit Morton code of the quantised folded position (|u|, sign(u) v)."""
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    s = np.copysign(1.0, u)
    fu, fv = np.abs(u), s * v
    r = max(float(np.max(np.abs(fu), initial=0.0)), float(np.max(np.abs(fv), initial=0.0)), 1e-300)
    qu = np.clip(fu / r * 65535.0, 0, 65535).astype(np.uint32)
    qv = np.clip((fv / r + 1.0) * 0.5 * 65535.0, 0, 65535).astype(np.uint32)
    return _spread16(qu) | (_spread16(qv) << np.uint32(1))


def locality_order(u, v) -> np.ndarray:
    """Permutation sorting (u, v) by locality_key, stably: two stable passes over the
    key's 16-bit halves, low then high (numpy sorts 16-bit integers by radix), which
    give the permutation of a stable sort on the whole key."""
    key = locality_key(u, v)
    lo = (key & np.uint32(0xFFFF)).astype(np.uint16)
    hi = (key >> np.uint32(16)).astype(np.uint16)
    o = np.argsort(lo, kind='stable')
    return o[np.argsort(hi[o], kind='stable')]


class Dataset:
    """An observation (u, v [, Re, Im, w]) resident on the GPU in locality order.

    Arrays may be numpy arrays or tensors; they are copied to ``device`` in
    ``dtype`` (torch.float64 for the double path, torch.float32 for single)."""

    def __init__(self, u, v, vis_obs_re=None, vis_obs_im=None, vis_obs_w=None, device='cuda',
                 dtype=torch.float64, sort=True):
        to_np = lambda a: a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)  # noqa: E731
        u, v = to_np(u).astype(np.float64), to_np(v).astype(np.float64)
        if u.shape != v.shape or u.ndim != 1:
            raise ValueError('u and v must be 1-D arrays of the same length')
        self.n = len(u)
        order = locality_order(u, v) if sort else np.arange(self.n)
        dev = torch.device(device)
        self.device, self.dtype = dev, dtype
        order_t = torch.from_numpy(order.astype(np.int64)).to(dev)
        # uploaded in the caller's order, gathered into the locality order on the device
        put = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)[order_t].to(dtype).contiguous()  # noqa: E731
        self.u, self.v = put(u), put(v)
        # bound on sqrt(u^2 + v^2) of the stored values: the spectrum is built only within it
        us, vs = (a.astype(np.float32 if dtype == torch.float32 else np.float64).astype(np.float64) for a in (u, v))
        self.uvmax = float(np.sqrt(np.max(us * us + vs * vs, initial=0.0)))
        self.has_obs = vis_obs_re is not None
        if self.has_obs:
            re, im, w = (to_np(a).astype(np.float64) for a in (vis_obs_re, vis_obs_im, vis_obs_w))
            if not (re.shape == im.shape == w.shape == u.shape):
                raise ValueError('vis_obs_re, vis_obs_im, vis_obs_w must have the length of u')
            self.re, self.im, self.w = put(re), put(im), put(w)
        self.order = order_t.to(torch.int32)   # sorted position i <- caller's point order[i]

    def chi2_profile(self, intensity, Rmin, dR, nxy, dxy, dRA=0., dDec=0., PA=0., inc=0.,
                     check_central_pixel=True, out=None):
        if not self.has_obs:
            raise ValueError('this Dataset has no observed visibilities')
        return chi2_profile(intensity, Rmin, dR, nxy, dxy, self.u, self.v, self.re, self.im, self.w, dRA=dRA,
                            dDec=dDec, PA=PA, inc=inc, check_central_pixel=check_central_pixel, out=out,
                            uvmax=self.uvmax)

    def sample_profile(self, intensity, Rmin, dR, nxy, dxy, dRA=0., dDec=0., PA=0., inc=0.,
                       check_central_pixel=True, out=None):
        """Visibilities in the caller's original point order."""
        return sample_profile(intensity, Rmin, dR, nxy, dxy, self.u, self.v, dRA=dRA, dDec=dDec, PA=PA, inc=inc,
                              check_central_pixel=check_central_pixel, out=out, perm=self.order, uvmax=self.uvmax)

    def sample_profile_multi(self, intensity, Rmin, dR, nxy, dxy, sources, PA=0., inc=0., check_central_pixel=True,
                             out=None):
        """Multi-source visibilities in the caller's original point order."""
        return sample_profile_multi(intensity, Rmin, dR, nxy, dxy, self.u, self.v, sources, PA=PA, inc=inc,
                                    check_central_pixel=check_central_pixel, out=out, perm=self.order,
                                    uvmax=self.uvmax)

    def chi2_profile_multi(self, intensity, Rmin, dR, nxy, dxy, sources, PA=0., inc=0., check_central_pixel=True,
                           out=None):
        if not self.has_obs:
            raise ValueError('this Dataset has no observed visibilities')
        return chi2_profile_multi(intensity, Rmin, dR, nxy, dxy, self.u, self.v, self.re, self.im, self.w, sources,
                                  PA=PA, inc=inc, check_central_pixel=check_central_pixel, out=out,
                                  uvmax=self.uvmax)


def chi2_profile_grad(intensity, Rmin, dR, nxy, dxy, u, v, vis_obs_re, vis_obs_im, vis_obs_w, dRA=0., dDec=0., PA=0.,
                      inc=0., check_central_pixel=True, uvmax=0.):
    """chi2 and its gradient on device tensors.

    Returns (chi2 [1], grad_intensity [nr], grad_geom [4] = d/d(inc, PA, dRA, dDec)), float64 device tensors."""
    dt = intensity.dtype
    dev = _check_tensors(dt, intensity, u, v, vis_obs_re, vis_obs_im, vis_obs_w)
    nd = u.numel()
    if not (v.numel() == nd == vis_obs_re.numel() == vis_obs_im.numel() == vis_obs_w.numel()):
        raise ValueError('u, v, vis_obs_re, vis_obs_im, vis_obs_w must have the same length')
    chi2 = torch.empty(1, dtype=torch.float64, device=dev)
    gI = torch.empty(intensity.numel(), dtype=torch.float64, device=dev)
    gg = torch.empty(4, dtype=torch.float64, device=dev)
    f = getattr(lib, f'galario_chi2_profile_grad_dev_{_SUF[dt]}')
    with torch.cuda.device(dev):
        _check(f(_stream_ptr(dev), intensity.numel(), _p(intensity), Rmin, dR, dxy, int(nxy), inc, dRA, dDec,
                 1.0 / (dxy * nxy), PA, nd, _p(u), _p(v), _p(vis_obs_re), _p(vis_obs_im), _p(vis_obs_w),
                 int(bool(check_central_pixel)), _p(chi2), _p(gI), _p(gg), float(uvmax)))
    return chi2, gI, gg


def _dataset_grad(self, intensity, Rmin, dR, nxy, dxy, dRA=0., dDec=0., PA=0., inc=0., check_central_pixel=True):
    if not self.has_obs:
        raise ValueError('this Dataset has no observed visibilities')
    return chi2_profile_grad(intensity, Rmin, dR, nxy, dxy, self.u, self.v, self.re, self.im, self.w, dRA=dRA,
                             dDec=dDec, PA=PA, inc=inc, check_central_pixel=check_central_pixel, uvmax=self.uvmax)


Dataset.chi2_profile_grad = _dataset_grad


class Chi2ProfileFunction(torch.autograd.Function):
    """chi2 of a radial profile against a Dataset, differentiable with respect to
    the profile and to geom = (inc, PA, dRA, dDec).

    The forward pass computes the gradient together with chi2 when any input
    needs it (one fused device call), and the backward pass scales it by the
    incoming gradient.  geom values are read on the host (a device tensor is
    synchronised)."""

    @staticmethod
    def forward(ctx, intensity, geom, dataset, Rmin, dR, nxy, dxy, check_central_pixel):
        inc, PA, dRA, dDec = (float(x) for x in geom.detach().cpu().tolist())
        I = intensity.detach().contiguous()
        kw = dict(dRA=dRA, dDec=dDec, PA=PA, inc=inc, check_central_pixel=check_central_pixel)
        if ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
            chi2, gI, gg = dataset.chi2_profile_grad(I, Rmin, dR, nxy, dxy, **kw)
            ctx.save_for_backward(gI, gg)
            ctx.dtypes = (intensity.dtype, geom.dtype, geom.device)
        else:
            chi2 = dataset.chi2_profile(I, Rmin, dR, nxy, dxy, **kw)
        return chi2.reshape(())

    @staticmethod
    def backward(ctx, gout):
        gI, gg = ctx.saved_tensors
        idt, gdt, gdev = ctx.dtypes
        dI = (gout * gI).to(idt) if ctx.needs_input_grad[0] else None
        dg = (gout * gg).to(device=gdev, dtype=gdt) if ctx.needs_input_grad[1] else None
        return dI, dg, None, None, None, None, None, None


def chi2_profile_autograd(dataset, intensity, geom, Rmin, dR, nxy, dxy, check_central_pixel=True):
    """Differentiable chi2Profile: intensity (nr) on the dataset's device, geom = tensor
    (inc, PA, dRA, dDec) in radians.  Returns a 0-dim float64 tensor."""
    return Chi2ProfileFunction.apply(intensity, geom, dataset, Rmin, dR, nxy, dxy, check_central_pixel)


def _destroy_plan(handle: int) -> None:
    lib.galario_batch_destroy(ctypes.c_void_p(handle))


class BatchedChi2Profile:
    """chi2Profile of many walkers' parameter sets over several datasets.

    ``sets``: sequence of ``(dataset, nxy, dxy)`` with ``Dataset`` objects
    holding observations (all in the same dtype, on the same device).
    Datasets on the same grid share one spectrum per walker.  A call takes
    ``intensity`` [W, nr] and ``geom`` [W, 4] = (inc, PA, dRA, dDec) device
    tensors in the datasets' dtype and returns the per-walker chi2 summed over
    the datasets (float64 [W]); ``return_sets=True`` also returns the
    per-dataset values [W, S].  The first call for a given W times chunk
    sizes (synchronising) and captures the launch sequence in a HIP graph;
    later calls replay it on torch's current stream.  Points outside a grid
    make the walker's chi2 NaN (``take_oob_flag``).

    ``phase_f64=True`` (float32 datasets only) is the FP32-bulk form with FP64
    geometry: profiles, spectra and dataset streams stay float32 while
    ``geom`` is float64 and the rotation, grid position, interpolation, phase
    argument and its sincos, residual and chi2 run in double on the float
    values (the step galario performs in dreal, src/galario.cpp:877-884)."""

    def __init__(self, sets, nr, Rmin, dR, max_walkers, check_central_pixel=True, phase_f64=False):
        sets = list(sets)
        if not sets:
            raise ValueError('at least one dataset is needed')
        ds0 = sets[0][0]
        self.dtype, self.device = ds0.dtype, ds0.device
        self.nr, self.nsets, self.max_walkers = int(nr), len(sets), int(max_walkers)
        suf = _SUF[self.dtype]
        self.phase_f64 = bool(phase_f64)
        if self.phase_f64 and suf != 'f32':
            raise ValueError('phase_f64 is the float32 datasets\' mixed form')
        self.geom_dtype = torch.float64 if self.phase_f64 else self.dtype
        St = DatasetF64 if suf == 'f64' else DatasetF32
        arr = (St * len(sets))()
        self._keep = []
        for k, (ds, nxy, dxy) in enumerate(sets):
            if not ds.has_obs:
                raise ValueError('every dataset needs observed visibilities')
            if ds.dtype != self.dtype or ds.device != self.device:
                raise ValueError('all datasets must share dtype and device')
            self._keep.append(ds)
            arr[k] = St(int(nxy), dxy, 1.0 / (dxy * nxy), ds.n, ds.u.data_ptr(), ds.v.data_ptr(),
                        ds.re.data_ptr(), ds.im.data_ptr(), ds.w.data_ptr(), ds.uvmax)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _check(getattr(lib, f'galario_batch_create_{suf}')(len(sets), arr, self.nr, Rmin, dR, self.max_walkers,
                                                               int(bool(check_central_pixel)), ctypes.byref(h)))
        self._h = h
        self._fin = weakref.finalize(self, _destroy_plan, h.value)
        x = 'x' if self.phase_f64 else ''
        self._run = getattr(lib, f'galario_batch_chi2_profile_{suf}{x}')
        self._run_grad = getattr(lib, f'galario_batch_chi2_profile_grad_{suf}{x}')

    def _mode(self, grad):
        return (BATCH_MODE_GRAD if grad else 0) | (BATCH_MODE_F64_PHASE if self.phase_f64 else 0)

    def _inputs(self, intensity, geom):
        _check_tensors(self.dtype, intensity)
        _check_tensors(self.geom_dtype, geom)
        if intensity.device != self.device or geom.device != self.device:
            raise ValueError('parameters must be on the datasets\' device')
        if intensity.dim() != 2 or intensity.shape[1] != self.nr:
            raise ValueError(f'intensity must be [W, {self.nr}]')
        W = intensity.shape[0]
        if tuple(geom.shape) != (W, 4):
            raise ValueError('geom must be [W, 4] = (inc, PA, dRA, dDec)')
        return W

    def __call__(self, intensity, geom, return_sets=False, eager=False):
        W = self._inputs(intensity, geom)
        chi2 = torch.empty(W, dtype=torch.float64, device=self.device)
        per = torch.empty(W, self.nsets, dtype=torch.float64, device=self.device) if return_sets else None
        with torch.cuda.device(self.device):
            _check(self._run(self._h, _stream_ptr(self.device), W, _p(intensity), _p(geom),
                             _p(per) if per is not None else None, _p(chi2), BATCH_EAGER if eager else 0))
        return (chi2, per) if return_sets else chi2

    def grad(self, intensity, geom, eager=False):
        """(chi2 [W], chi2 per dataset [W, S], dchi2/dintensity [W, nr], dchi2/d(inc, PA, dRA, dDec) [W, 4])."""
        W = self._inputs(intensity, geom)
        f64 = dict(dtype=torch.float64, device=self.device)
        chi2, per = torch.empty(W, **f64), torch.empty(W, self.nsets, **f64)
        gI, gg = torch.empty(W, self.nr, **f64), torch.empty(W, 4, **f64)
        with torch.cuda.device(self.device):
            _check(self._run_grad(self._h, _stream_ptr(self.device), W, _p(intensity), _p(geom), _p(per), _p(chi2),
                                  _p(gI), _p(gg), BATCH_EAGER if eager else 0))
        return chi2, per, gI, gg

    def autograd(self, intensity, geom):
        """Per-walker chi2 [W] differentiable with respect to intensity [W, nr] and geom [W, 4]."""
        return _BatchedChi2Function.apply(self, intensity, geom)

    def info(self, nwalkers, grad=False):
        """(chunk size, walkers per sample workgroup) picked for nwalkers, and whether its graph is captured."""
        c, k, g = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        _check(lib.galario_batch_info(self._h, int(nwalkers), self._mode(grad), ctypes.byref(c), ctypes.byref(k),
                                      ctypes.byref(g)))
        return c.value, k.value, bool(g.value)

    def set_chunk(self, nwalkers, chunk, walkers_per_sample=4, grad=False):
        """Override the timed launch shape for nwalkers."""
        _check(lib.galario_batch_set_chunk(self._h, int(nwalkers), self._mode(grad), int(chunk),
                                           int(walkers_per_sample)))

    def tile(self, nwalkers, grad=False):
        """(256-point blocks per forward sample workgroup, LDS window cells of the gradient sample kernel)
        picked for nwalkers; zeros before the first call."""
        pb, wn = ctypes.c_int(0), ctypes.c_int(0)
        _check(lib.galario_batch_tile(self._h, int(nwalkers), self._mode(grad), ctypes.byref(pb), ctypes.byref(wn)))
        return pb.value, wn.value

    def set_tile(self, nwalkers, point_blocks=1, grad_window=3072, grad=False):
        """Fix the sample kernel's tile for nwalkers (the first call's timing keeps it)."""
        _check(lib.galario_batch_set_tile(self._h, int(nwalkers), self._mode(grad), int(point_blocks),
                                          int(grad_window)))

    def take_oob_flag(self) -> bool:
        out = ctypes.c_int(0)
        with torch.cuda.device(self.device):
            _check(lib.galario_batch_take_oob_flag(self._h, ctypes.byref(out)))
        return bool(out.value)

    def close(self):
        self._fin()


class _BatchedChi2Function(torch.autograd.Function):
    @staticmethod
    def forward(ctx, plan, intensity, geom):
        I = intensity.detach().contiguous()
        g = geom.detach().to(I.device, plan.geom_dtype).contiguous()
        if ctx.needs_input_grad[1] or ctx.needs_input_grad[2]:
            chi2, _, gI, gg = plan.grad(I, g)
            ctx.save_for_backward(gI, gg)
        else:
            chi2 = plan(I, g)
        ctx.geom_device = geom.device
        ctx.dtypes = (intensity.dtype, geom.dtype)
        return chi2

    @staticmethod
    def backward(ctx, go):
        gI, gg = ctx.saved_tensors
        go = go.to(gI.dtype)[:, None]
        dI = (go * gI).to(ctx.dtypes[0]) if ctx.needs_input_grad[1] else None
        dg = (go * gg).to(ctx.geom_device, ctx.dtypes[1]) if ctx.needs_input_grad[2] else None
        return None, dI, dg
