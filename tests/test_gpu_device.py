"""GPU tests of the device-resident forms (C ABI _dev entry points on torch
tensors, the locality-ordered Dataset) against the host-pointer drop-in path
and the oracle."""
import math

import numpy as np
import pytest

import galario_amd
from oracle import galario_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')
arcsec = O.ARCSEC


def workload(nvis=50_000, nxy=1024, seed=7, dtype=np.float64):
    rng = np.random.default_rng(seed)
    Rmin, dR, nr, dxy = 1e-4 * arcsec, 0.002 * arcsec, 600, 0.01 * arcsec
    R = Rmin + dR * np.arange(nr)
    I = 1e10 * (np.exp(-((R - 0.3 * arcsec) / (0.1 * arcsec)) ** 2) + 0.2 * np.exp(-R / (0.5 * arcsec)))
    duv = 1 / (nxy * dxy)
    b = np.exp(rng.uniform(np.log(duv), np.log(0.45 * nxy * duv), nvis))
    th = rng.uniform(0, 2 * np.pi, nvis)
    u, v = b * np.cos(th), b * np.sin(th)
    re, im = rng.standard_normal(nvis), rng.standard_normal(nvis)
    w = rng.uniform(0.5, 2.0, nvis)
    kw = dict(inc=0.7, PA=2.1, dRA=0.02 * arcsec, dDec=-0.01 * arcsec)
    cast = lambda a: np.ascontiguousarray(a, dtype=dtype)  # noqa: E731
    return dict(I=cast(I), Rmin=Rmin, dR=dR, nxy=nxy, dxy=dxy, u=cast(u), v=cast(v), re=cast(re), im=cast(im),
                w=cast(w), kw=kw)


@pytest.mark.parametrize("dtype, tol", [(np.float64, 1e-12), (np.float32, 1e-5)], ids=['DP', 'SP'])
def test_device_forms_match_host_path(dtype, tol):
    from galario_amd import torch_api as ta
    g = galario_amd.double if dtype == np.float64 else galario_amd.single
    d = workload(dtype=dtype)
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    dev = torch.device('cuda:0')
    T = lambda a: torch.from_numpy(a).to(dev)  # noqa: E731
    args = (d['I'], d['Rmin'], d['dR'], d['nxy'], d['dxy'], d['u'], d['v'])
    vis_h = g.sampleProfile(*args, check_central_pixel=False, **d['kw'])
    chi_h = g.chi2Profile(*args, d['re'], d['im'], d['w'], check_central_pixel=False, **d['kw'])
    tI, tu, tv, tre, tim, tw = map(T, (d['I'], d['u'], d['v'], d['re'], d['im'], d['w']))
    vis_d = ta.sample_profile(tI, d['Rmin'], d['dR'], d['nxy'], d['dxy'], tu, tv, check_central_pixel=False,
                              **d['kw'])
    chi_d = ta.chi2_profile(tI, d['Rmin'], d['dR'], d['nxy'], d['dxy'], tu, tv, tre, tim, tw,
                            check_central_pixel=False, **d['kw'])
    assert np.array_equal(vis_d.cpu().numpy(), vis_h)        # same kernels, same order: bit-identical
    assert math.isclose(chi_d.item(), chi_h, rel_tol=tol)
    assert not ta.take_oob_flag()
    # locality-ordered dataset: visibilities come back in the caller's order
    ds = ta.Dataset(d['u'], d['v'], d['re'], d['im'], d['w'], device=dev, dtype=tdt)
    vis_s = ds.sample_profile(tI, d['Rmin'], d['dR'], d['nxy'], d['dxy'], check_central_pixel=False, **d['kw'])
    assert np.array_equal(vis_s.cpu().numpy(), vis_h)
    chi_s = ds.chi2_profile(tI, d['Rmin'], d['dR'], d['nxy'], d['dxy'], check_central_pixel=False, **d['kw'])
    assert math.isclose(chi_s.item(), chi_h, rel_tol=tol)
    # against the oracle (double only)
    if dtype == np.float64:
        ref = O.sample_profile(d['I'], d['Rmin'], d['dR'], d['dxy'], d['nxy'], d['kw']['inc'], d['kw']['dRA'],
                               d['kw']['dDec'], 1 / (d['nxy'] * d['dxy']), d['kw']['PA'], d['u'], d['v'],
                               check_central=False)
        assert np.abs(vis_h - ref).max() <= 1e-12 * np.abs(ref).max()


def test_device_out_of_grid_flag():
    from galario_amd import torch_api as ta
    d = workload(nvis=1000)
    dev = torch.device('cuda:0')
    u, v = d['u'].copy(), d['v'].copy()
    duv = 1 / (d['nxy'] * d['dxy'])
    u[17], v[17] = 0.8 * d['nxy'] * duv, 0.0       # radius 0.8 nxy cells: off the grid for every PA
    T = lambda a: torch.from_numpy(a).to(dev)  # noqa: E731
    chi = ta.chi2_profile(T(d['I']), d['Rmin'], d['dR'], d['nxy'], d['dxy'], T(u), T(v), T(d['re']), T(d['im']),
                          T(d['w']), check_central_pixel=False, **d['kw'])
    assert math.isnan(chi.item())
    assert ta.take_oob_flag() is True
    assert ta.take_oob_flag() is False


def test_stage_timer_records_calls():
    from galario_amd import torch_api as ta
    d = workload(nvis=10_000)
    dev = torch.device('cuda:0')
    ds = ta.Dataset(d['u'], d['v'], d['re'], d['im'], d['w'], device=dev)
    tI = torch.from_numpy(d['I']).to(dev)
    ta.stage_timer(True)
    for _ in range(3):
        ds.chi2_profile(tI, d['Rmin'], d['dR'], d['nxy'], d['dxy'], check_central_pixel=False, **d['kw'])
    ms, n = ta.stage_times()
    ta.stage_timer(False)
    assert n == 3 and all(ms[k] > 0 for k in ('row_pass', 'column_pass', 'sample'))
    assert ms['adjoint_column_pass'] == 0 and ms['adjoint_row_pass'] == 0


def test_stage_repeat_leaves_results_unchanged():
    """Repeated launches of the forward kernels (kernel timing) are idempotent."""
    from galario_amd import torch_api as ta
    d = workload(nvis=10_000)
    dev = torch.device('cuda:0')
    ds = ta.Dataset(d['u'], d['v'], d['re'], d['im'], d['w'], device=dev)
    tI = torch.from_numpy(d['I']).to(dev)
    args = (tI, d['Rmin'], d['dR'], d['nxy'], d['dxy'])
    ref = ds.chi2_profile(*args, check_central_pixel=False, **d['kw']).item()
    vref = ds.sample_profile(*args, check_central_pixel=False, **d['kw']).cpu().numpy()
    ta.stage_timer(True)
    try:
        for st in ('row_pass', 'column_pass', 'sample'):
            ta.stage_repeat(st, 3)
            assert ds.chi2_profile(*args, check_central_pixel=False, **d['kw']).item() == ref
            assert np.array_equal(ds.sample_profile(*args, check_central_pixel=False, **d['kw']).cpu().numpy(), vref)
        ms, n = ta.stage_times()
        assert n == 6 and ms['reduce'] > 0
    finally:
        ta.stage_repeat(None, 1)
        ta.stage_timer(False)
    with pytest.raises(ValueError):
        ta.stage_repeat('adjoint_row_pass', 2)


@pytest.mark.parametrize("grid", [1, 7, 1024])
def test_read_probe_reads_every_element(grid):
    """The read-rate reference visits each 16-byte element once: its per-workgroup
    sums add up to the buffer's sum (small integers: exact in FP64)."""
    from galario_amd import torch_api as ta
    dev = torch.device('cuda:0')
    n = 2 * 100_003
    buf = torch.arange(n, dtype=torch.float64, device=dev) % 1000
    out = torch.full((grid,), float('nan'), dtype=torch.float64, device=dev)
    ta.read_probe(buf, grid, out)
    assert out.sum().item() == buf.sum().item()
    with pytest.raises(ValueError):
        ta.read_probe(buf, grid + 1, out)
