"""GPU: the batched form (many walkers' parameter sets over several datasets
of uneven nxy and Nvis, one plan, graph replay).

Checked against the single-call device path on the same sorted datasets:
chi2 of every (walker, dataset) to 1e-12 relative in double (the kernels and
the summation order are the same; only the per-walker cos/sin are evaluated
on the device instead of the host), 1e-5 in float; gradients summed over
datasets to 1e-9 relative (atomic accumulation order); graph replay == eager
launches and every chunk size gives identical chi2 (fixed-order reduction).
One walker on the smallest dataset is also checked against the oracle."""
import math

import numpy as np
import pytest

from oracle import galario_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')
arcsec = O.ARCSEC
DEV = torch.device('cuda:0')


def make_sets(dtype=torch.float64, seed=5, specs=((256, 3000, 0.02), (512, 20000, 0.01), (1024, 40000, 0.005),
                                                    (512, 7000, 0.01))):
    from galario_amd import torch_api as ta
    rng = np.random.default_rng(seed)
    sets, raw = [], []
    for nxy, nvis, dxy_as in specs:
        dxy = dxy_as * arcsec
        duv = 1 / (nxy * dxy)
        b = np.exp(rng.uniform(np.log(duv), np.log(0.45 * nxy * duv), nvis))
        th = rng.uniform(0, 2 * np.pi, nvis)
        u, v = b * np.cos(th), b * np.sin(th)
        re, im = rng.standard_normal(nvis) * 1e-3, rng.standard_normal(nvis) * 1e-3
        w = rng.uniform(0.5, 2.0, nvis) * 1e6
        sets.append((ta.Dataset(u, v, re, im, w, device=DEV, dtype=dtype), nxy, dxy))
        raw.append(dict(u=u, v=v, re=re, im=im, w=w, nxy=nxy, dxy=dxy))
    return sets, raw


def walkers(W, nr=300, seed=9, dtype=torch.float64):
    rng = np.random.default_rng(seed)
    Rmin, dR = 1e-4 * arcsec, 0.004 * arcsec
    R = Rmin + dR * np.arange(nr)
    base = np.exp(-((R - 0.4 * arcsec) / (0.1 * arcsec)) ** 2) + 0.3 * np.exp(-R / (0.6 * arcsec))
    I = 1e10 * base[None, :] * np.exp(0.1 * rng.standard_normal((W, 1)))
    geom = np.stack([rng.uniform(0.1, 1.2, W), rng.uniform(-3, 3, W), rng.uniform(-0.05, 0.05, W) * arcsec,
                     rng.uniform(-0.05, 0.05, W) * arcsec], axis=1)
    geom[0, 1] = 0.0      # one walker with PA == 0 (rotation skipped)
    geom[1, 2:] = 0.0     # one without offsets (phase skipped)
    T = lambda a: torch.tensor(a, dtype=dtype, device=DEV)  # noqa: E731
    return T(I), T(geom), Rmin, dR


def plain(sets, I, geom, Rmin, dR, grad=False):
    out = []
    for w in range(I.shape[0]):
        g = geom[w].double().cpu().numpy()
        kw = dict(inc=float(g[0]), PA=float(g[1]), dRA=float(g[2]), dDec=float(g[3]), check_central_pixel=False)
        row = []
        for ds, nxy, dxy in sets:
            if grad:
                c, gI, gg = ds.chi2_profile_grad(I[w].contiguous(), Rmin, dR, nxy, dxy, **kw)
                row.append((c.item(), gI.cpu().numpy(), gg.cpu().numpy()))
            else:
                row.append(ds.chi2_profile(I[w].contiguous(), Rmin, dR, nxy, dxy, **kw).item())
        out.append(row)
    return out


def test_batch_forward_matches_single_calls_and_oracle():
    from galario_amd import torch_api as ta
    sets, raw = make_sets()
    W = 6
    I, geom, Rmin, dR = walkers(W)
    plan = ta.BatchedChi2Profile(sets, I.shape[1], Rmin, dR, max_walkers=8, check_central_pixel=False)
    chi2, per = plan(I, geom, return_sets=True)
    chunk, kw, ready = plan.info(W)
    assert chunk >= 1 and kw in (1, 2, 4, 8) and ready
    ref = np.array(plain(sets, I, geom, Rmin, dR))
    got = per.cpu().numpy()
    print(f"\nchunk {chunk} x {kw}; max rel diff vs single calls {np.max(np.abs(got - ref) / np.abs(ref)):.2e}")
    np.testing.assert_allclose(got, ref, rtol=1e-12)
    np.testing.assert_allclose(chi2.cpu().numpy(), got.sum(axis=1), rtol=1e-15)
    # oracle: walker 2 on dataset 0
    d, g = raw[0], geom[2].cpu().numpy()
    c_or = O.chi2_profile(I[2].cpu().numpy(), Rmin, dR, d['dxy'], d['nxy'], g[0], g[2], g[3], 1 / (d['nxy'] * d['dxy']),
                          g[1], d['u'], d['v'], d['re'], d['im'], d['w'], check_central=False)
    assert math.isclose(got[2, 0], c_or, rel_tol=1e-12)
    assert not plan.take_oob_flag()
    # graph replay with new parameters, eager launches, every chunk size: identical
    I2, geom2, _, _ = walkers(W, seed=10)
    c_graph = plan(I2, geom2).cpu().numpy()
    c_eager = plan(I2, geom2, eager=True).cpu().numpy()
    assert np.array_equal(c_graph, c_eager)
    for c, kw in ((1, 1), (3, 2), (W, 4), (W, 8), (5, 8)):
        plan.set_chunk(W, c, kw)
        assert np.array_equal(plan(I2, geom2).cpu().numpy(), c_graph), (c, kw)
    ref2 = np.array(plain(sets, I2, geom2, Rmin, dR)).sum(axis=1)
    np.testing.assert_allclose(c_graph, ref2, rtol=1e-12)
    # fewer walkers than the plan holds: its own tuning and graph
    c3 = plan(I2[:3].contiguous(), geom2[:3].contiguous()).cpu().numpy()
    assert np.array_equal(c3, c_graph[:3])
    plan.close()


def test_batch_gradient_matches_single_calls():
    from galario_amd import torch_api as ta
    sets, _ = make_sets(specs=((256, 3000, 0.02), (512, 20000, 0.01), (512, 5000, 0.01)))
    W = 4
    I, geom, Rmin, dR = walkers(W)
    plan = ta.BatchedChi2Profile(sets, I.shape[1], Rmin, dR, max_walkers=W, check_central_pixel=False)
    chi2, per, gI, gg = plan.grad(I, geom)
    ref = plain(sets, I, geom, Rmin, dR, grad=True)
    for w in range(W):
        c_ref = [r[0] for r in ref[w]]
        gI_ref = sum(r[1] for r in ref[w])
        gg_ref = sum(r[2] for r in ref[w])
        np.testing.assert_allclose(per[w].cpu().numpy(), c_ref, rtol=1e-12)
        gIw = gI[w].cpu().numpy()
        assert np.linalg.norm(gIw - gI_ref) <= 1e-9 * np.linalg.norm(gI_ref), w
        np.testing.assert_allclose(gg[w].cpu().numpy(), gg_ref, rtol=1e-9)
    # forward-only call of the same plan agrees with the chi2 of the gradient call
    assert np.array_equal(plan(I, geom).cpu().numpy(), chi2.cpu().numpy())
    # repeat (the adjoint grids are cleared between calls) and eager
    _, _, gI2, gg2 = plan.grad(I, geom, eager=True)
    assert torch.linalg.norm(gI2 - gI) <= 1e-12 * torch.linalg.norm(gI)
    np.testing.assert_allclose(gg2.cpu().numpy(), gg.cpu().numpy(), rtol=1e-12)
    # autograd: d(sum_w a_w chi2_w)
    a = torch.arange(1, W + 1, dtype=torch.float64, device=DEV)
    Ir = I.clone().requires_grad_(True)
    gr = geom.clone().requires_grad_(True)
    (a * plan.autograd(Ir, gr)).sum().backward()
    assert torch.allclose(Ir.grad, a[:, None] * gI, rtol=1e-12, atol=0)
    assert torch.allclose(gr.grad, a[:, None] * gg, rtol=1e-12, atol=0)
    plan.close()


def test_batch_every_tile_agrees():
    """The autotuner's tile dimensions change no output: every (walkers per sample workgroup, 256-point blocks
    per workgroup) of the forward sample kernel gives identical chi2 (each block's partial is the same
    fixed-order sum), every LDS window of the gradient sample kernel identical chi2 and gradients to 1e-12
    relative (only the order of the atomic additions into the half spectrum changes)."""
    from galario_amd import torch_api as ta
    sets, _ = make_sets(specs=((256, 3000, 0.02), (512, 20000, 0.01), (1024, 9000, 0.005)))
    W = 5
    I, geom, Rmin, dR = walkers(W)
    plan = ta.BatchedChi2Profile(sets, I.shape[1], Rmin, dR, max_walkers=W, check_central_pixel=False)
    c0 = plan(I, geom).cpu().numpy()
    assert plan.tile(W)[0] in (1, 2) and plan.tile(W)[1] == 0
    chunk = plan.info(W)[0]
    for kw in (1, 2, 4, 8):
        for pb in (1, 2):
            plan.set_chunk(W, chunk, kw)
            plan.set_tile(W, point_blocks=pb)
            assert plan.tile(W) == (pb, 0)
            assert np.array_equal(plan(I, geom).cpu().numpy(), c0), (kw, pb)
    chi2, _, gI, gg = plan.grad(I, geom)
    assert np.array_equal(chi2.cpu().numpy(), c0)
    assert plan.tile(W, grad=True)[1] in (2048, 3072, 4096)
    for win in (2048, 3072, 4096):
        plan.set_tile(W, grad_window=win, grad=True)
        assert plan.tile(W, grad=True) == (1, win)
        c2, _, gI2, gg2 = plan.grad(I, geom)
        assert np.array_equal(c2.cpu().numpy(), c0), win
        assert torch.linalg.norm(gI2 - gI) <= 1e-12 * torch.linalg.norm(gI), win
        np.testing.assert_allclose(gg2.cpu().numpy(), gg.cpu().numpy(), rtol=1e-12)
    # a tile fixed before the first call is kept by the timed choice of the chunk
    plan2 = ta.BatchedChi2Profile(sets, I.shape[1], Rmin, dR, max_walkers=W, check_central_pixel=False)
    plan2.set_tile(W, point_blocks=2)
    assert np.array_equal(plan2(I, geom).cpu().numpy(), c0)
    assert plan2.tile(W) == (2, 0) and plan2.info(W)[0] >= 1
    with pytest.raises(ValueError, match="point blocks"):
        plan2.set_tile(W, point_blocks=3)
    with pytest.raises(ValueError, match="gradient window"):
        plan2.set_tile(W, grad_window=1000, grad=True)
    plan.close()
    plan2.close()


def test_allocation_failure_raises_memory_error():
    """galario maps std::bad_alloc to MemoryError (src/galario.cpp:158-198, docs/py-api.rst:72-82): a plan
    whose parameter buffers exceed the GPU's memory (2^31 - 1 walkers x 300 bins: 5 TB) fails with
    MemoryError, and the next plan and call work normally (the failed allocation leaves no error behind)."""
    from galario_amd import torch_api as ta
    sets, _ = make_sets(specs=((256, 3000, 0.02),))
    I, geom, Rmin, dR = walkers(2)
    with pytest.raises(MemoryError):
        ta.BatchedChi2Profile(sets, I.shape[1], Rmin, dR, max_walkers=2 ** 31 - 1, check_central_pixel=False)
    plan = ta.BatchedChi2Profile(sets, I.shape[1], Rmin, dR, max_walkers=2, check_central_pixel=False)
    ref = np.array(plain(sets, I, geom, Rmin, dR)).sum(axis=1)
    np.testing.assert_allclose(plan(I, geom).cpu().numpy(), ref, rtol=1e-12)
    plan.close()


def test_batch_single_precision():
    from galario_amd import torch_api as ta
    sets, _ = make_sets(dtype=torch.float32, specs=((256, 3000, 0.02), (1024, 30000, 0.005)))
    W = 3
    I, geom, Rmin, dR = walkers(W, dtype=torch.float32)
    plan = ta.BatchedChi2Profile(sets, I.shape[1], Rmin, dR, max_walkers=W, check_central_pixel=False)
    _, per = plan(I, geom, return_sets=True)
    ref = np.array(plain(sets, I, geom, Rmin, dR))
    np.testing.assert_allclose(per.cpu().numpy(), ref, rtol=1e-5)
    sets64, _ = make_sets(dtype=torch.float64, specs=((256, 3000, 0.02), (1024, 30000, 0.005)))
    ref64 = np.array(plain(sets64, I.double(), geom.double(), Rmin, dR))
    print(f"\nfloat vs double chi2: max rel {np.max(np.abs(per.cpu().numpy() - ref64) / ref64):.2e}")
    np.testing.assert_allclose(per.cpu().numpy(), ref64, rtol=1e-3)
    c, _, gI, gg = plan.grad(I, geom)
    assert torch.isfinite(gI).all() and torch.isfinite(gg).all()
    plan.close()


def test_batch_out_of_grid_point():
    from galario_amd import torch_api as ta
    sets, raw = make_sets(specs=((256, 2000, 0.02),))
    d = raw[0]
    u, v = d['u'].copy(), d['v'].copy()
    u[5], v[5] = 0.8 * d['nxy'] / (d['nxy'] * d['dxy']), 0.0
    bad = ta.Dataset(u, v, d['re'], d['im'], d['w'], device=DEV)
    I, geom, Rmin, dR = walkers(2)
    plan = ta.BatchedChi2Profile([sets[0], (bad, d['nxy'], d['dxy'])], I.shape[1], Rmin, dR, max_walkers=2,
                                 check_central_pixel=False)
    _, per = plan(I, geom, return_sets=True)
    per = per.cpu().numpy()
    assert np.isfinite(per[:, 0]).all() and np.isnan(per[:, 1]).all()
    assert plan.take_oob_flag() is True and plan.take_oob_flag() is False
    with pytest.raises(ValueError, match="nwalkers"):
        plan(torch.cat([I, I]), torch.cat([geom, geom]))
    plan.close()


def test_batch_plan_survives_cleanup():
    """galario's _cleanup() (a public API) between two calls of one plan: the plan's captured graph keeps
    working with the same results (the FFT twiddle tables it reads live until the process ends)."""
    import galario_amd
    from galario_amd import torch_api as ta
    sets, _ = make_sets(specs=((256, 3000, 0.02), (512, 8000, 0.01)))
    I, geom, Rmin, dR = walkers(3)
    plan = ta.BatchedChi2Profile(sets, I.shape[1], Rmin, dR, max_walkers=3, check_central_pixel=False)
    c1 = plan(I, geom).cpu().numpy()
    _, _, ready = plan.info(3)
    assert ready
    torch.cuda.synchronize()
    galario_amd.double._cleanup()
    c2 = plan(I, geom).cpu().numpy()
    assert np.array_equal(c1, c2)
    # a plain call after the cleanup rebuilds its workspace and agrees with the plan
    ref = np.array(plain(sets, I, geom, Rmin, dR)).sum(axis=1)
    np.testing.assert_allclose(c2, ref, rtol=1e-12)
    plan.close()


def test_batch_float_bulk_with_f64_phase():
    """phase_f64: float datasets and spectra, double geometry and per-point arithmetic.  Its chi2 is
    within the float bar (1e-5) of the double single calls (what remains is the float spectrum's
    rounding; both forms' errors are printed), its profile gradient within 1e-4; gradient calls agree
    with forward calls bit for bit; graph == eager."""
    from galario_amd import torch_api as ta
    specs = ((256, 3000, 0.02), (1024, 30000, 0.005))
    sets32, _ = make_sets(dtype=torch.float32, specs=specs)
    sets64, _ = make_sets(dtype=torch.float64, specs=specs)
    W = 4
    I64, geom64, Rmin, dR = walkers(W)
    I32 = I64.float()
    mixed = ta.BatchedChi2Profile(sets32, I32.shape[1], Rmin, dR, max_walkers=W, check_central_pixel=False,
                                  phase_f64=True)
    pure = ta.BatchedChi2Profile(sets32, I32.shape[1], Rmin, dR, max_walkers=W, check_central_pixel=False)
    with pytest.raises(ValueError, match="float64"):
        mixed(I32, geom64.float())
    _, per_mix = mixed(I32, geom64, return_sets=True)
    _, per_f32 = pure(I32, geom64.float(), return_sets=True)
    ref = np.array(plain(sets64, I32.double(), geom64, Rmin, dR))
    e_mix = np.max(np.abs(per_mix.cpu().numpy() - ref) / ref)
    e_f32 = np.max(np.abs(per_f32.cpu().numpy() - ref) / ref)
    print(f"\nchi2 vs double single calls: float bulk + f64 phase {e_mix:.1e}, float {e_f32:.1e}")
    assert e_mix <= 1e-5
    assert mixed.info(W)[2]                  # the forward graph is captured
    c_eager = mixed(I32, geom64, eager=True).cpu().numpy()
    c, _, gI, gg = mixed.grad(I32, geom64)   # (grows the plan's slots: that drops the forward graph)
    assert np.array_equal(c.cpu().numpy(), c_eager)
    assert np.array_equal(c_eager, per_mix.cpu().numpy().sum(axis=1)) or \
        np.allclose(c_eager, per_mix.cpu().numpy().sum(axis=1), rtol=1e-15, atol=0)
    ref_g = plain(sets64, I32.double(), geom64, Rmin, dR, grad=True)
    for w in range(W):
        gI_ref = sum(r[1] for r in ref_g[w])
        assert np.linalg.norm(gI[w].cpu().numpy() - gI_ref) <= 1e-4 * np.linalg.norm(gI_ref), w
    assert mixed.info(W, grad=True)[2]
    mixed.close()
    pure.close()
