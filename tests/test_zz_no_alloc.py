"""GPU: the allocation-free chi2 entry point (galario_chi2_profile_noalloc_dev_*) through the C ABI.

Checked against the existing device entry point galario_chi2_profile_dev_* on
the same device arrays, on two calls in a row, the second with a smaller
input (smaller nxy, fewer points, another geometry): the two entry points run
the same kernels, so chi2 must be bit-identical (stricter than the 1e-12
relative bar of the double path).  A float call is checked the same way.
Sizes beyond the library's scratch and unsupported nxy are rejected on the
host with ValueError, before any GPU work.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')

import galario_amd  # noqa: E402  (its import sets up the current device)
from galario_amd import _lib  # noqa: E402
from galario_amd import synthetic as syn  # noqa: E402

DEV = torch.device('cuda:0')
A = syn.ARCSEC


def problem(nxy, nvis, nr, dR, seed, dtype):
    rng = np.random.default_rng(seed)
    Rmin, dxy = 1e-4 * A, 0.01 * A
    I = syn.ring_envelope_profile(nr, Rmin, dR)
    duv = 1.0 / (nxy * dxy)
    u, v = syn.log_uniform_uv(rng, nvis, 2 * duv, 0.45 * nxy * duv)
    re, im = rng.standard_normal(nvis) * 1e-3, rng.standard_normal(nvis) * 1e-3
    w = rng.uniform(0.5, 2.0, nvis) * 1e6
    T = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)  # noqa: E731
    geom = (rng.uniform(0.2, 1.0), rng.uniform(-3, 3), rng.uniform(-0.05, 0.05) * A, rng.uniform(-0.05, 0.05) * A)
    return dict(nxy=nxy, dxy=dxy, Rmin=Rmin, dR=dR, I=T(I), u=T(u), v=T(v), re=T(re), im=T(im), w=T(w), geom=geom,
                uvmax=float(np.hypot(u, v).max()))


def call(name, suf, p, out):
    f = getattr(_lib.lib, f'galario_{name}_{suf}')
    inc, PA, dRA, dDec = p['geom']
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    stream = ctypes.c_void_p(torch.cuda.current_stream(DEV).cuda_stream)
    return f(stream, p['I'].numel(), ptr(p['I']), p['Rmin'], p['dR'], p['dxy'], p['nxy'], inc, dRA, dDec,
             1.0 / (p['nxy'] * p['dxy']), PA, p['u'].numel(), ptr(p['u']), ptr(p['v']), ptr(p['re']), ptr(p['im']),
             ptr(p['w']), 0, ptr(out), p['uvmax'])


@pytest.mark.parametrize('suf, dtype', [('f64', torch.float64), ('f32', torch.float32)])
def test_noalloc_matches_device_entry_point_on_two_calls(suf, dtype):
    big = problem(4096, 1_000_000, 1000, 0.002 * A, 1, dtype)
    small = problem(1024, 50_000, 400, 0.004 * A, 2, dtype)
    out = torch.full((2,), float('nan'), dtype=torch.float64, device=DEV)
    ref = torch.full((2,), float('nan'), dtype=torch.float64, device=DEV)
    for k, p in enumerate((big, small)):
        print(f"\n{suf}: nxy={p['nxy']} nvis={p['u'].numel()}", flush=True)
        _lib.check(call('chi2_profile_noalloc_dev', suf, p, out[k:]))
        _lib.check(call('chi2_profile_dev', suf, p, ref[k:]))
    torch.cuda.synchronize()
    got, exp = out.cpu().numpy(), ref.cpu().numpy()
    print(f"{suf}: allocation-free {got}, device entry point {exp}")
    assert np.all(np.isfinite(got))
    assert np.array_equal(got, exp)
    # the first input again after the smaller one: the array keeps nothing that matters between calls
    _lib.check(call('chi2_profile_noalloc_dev', suf, big, out[1:]))
    torch.cuda.synchronize()
    assert out[1].item() == exp[0]


def test_noalloc_rejects_what_does_not_fit():
    """nxy = 8192 in double needs more scratch than the library holds; nxy = 1000 is not a power of two."""
    p = problem(1024, 1000, 400, 0.004 * A, 3, torch.float64)
    out = torch.zeros(1, dtype=torch.float64, device=DEV)
    with pytest.raises(ValueError, match="bytes of scratch"):
        _lib.check(call('chi2_profile_noalloc_dev', 'f64', dict(p, nxy=8192), out))
    with pytest.raises(ValueError, match="power of two"):
        _lib.check(call('chi2_profile_noalloc_dev', 'f64', dict(p, nxy=1000), out))
    assert out.item() == 0.0
