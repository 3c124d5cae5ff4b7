"""Seeded synthetic workloads for tests and benchmarks.

RNG: ``numpy.random.default_rng(seed)`` (PCG64).  Draw order per workload is
fixed and documented in each function, so every array is reproducible from
its seed.  Observed visibilities are a model plus Gaussian noise whose
per-point sigma is log-uniform in [0.5, 2] x median(|V_model|); the weights
are w = 1/sigma^2.  The model visibilities are supplied by the caller (a
callable), so the same generator serves the GPU benchmark and the CPU oracle.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

ARCSEC = 4.84813681109536e-06
DEG = 0.017453292519943295


def log_uniform_uv(rng, n: int, bmin: float, bmax: float):
    """u, v with |b| log-uniform in [bmin, bmax] wavelengths and angle uniform in [0, 2 pi).
    Draws: n log-radii, then n angles."""
    b = np.exp(rng.uniform(np.log(bmin), np.log(bmax), n))
    th = rng.uniform(0.0, 2.0 * np.pi, n)
    return b * np.cos(th), b * np.sin(th)


def noisy_observation(rng, vis_model):
    """(re, im, w): model + noise, sigma log-uniform in [0.5, 2] x median |V|.
    Draws: n log-sigma factors, n real-part normals, n imaginary-part normals."""
    n = len(vis_model)
    med = float(np.median(np.abs(vis_model))) if n else 1.0
    sigma = med * np.exp(rng.uniform(np.log(0.5), np.log(2.0), n))
    re = vis_model.real + sigma * rng.standard_normal(n)
    im = vis_model.imag + sigma * rng.standard_normal(n)
    return re, im, 1.0 / sigma ** 2


def ring_envelope_profile(nr: int, Rmin: float, dR: float, scale: float = 1e10):
    """Radial profile in Jy/sr: two Gaussian rings plus a power-law envelope.

    I(R) = scale * [ exp(-(R-0.40")^2 / (2 (0.05")^2))
                   + 0.5 exp(-(R-0.90")^2 / (2 (0.08")^2))
                   + 0.3 (R/1")^-0.8 exp(-R/1.2") ],   R = Rmin + k dR."""
    R = Rmin + dR * np.arange(nr)
    a = ARCSEC
    ring1 = np.exp(-((R - 0.40 * a) ** 2) / (2 * (0.05 * a) ** 2))
    ring2 = 0.5 * np.exp(-((R - 0.90 * a) ** 2) / (2 * (0.08 * a) ** 2))
    env = 0.3 * (R / a) ** -0.8 * np.exp(-R / (1.2 * a))
    return scale * (ring1 + ring2 + env)


def gaussian_ring_image(rng, nxy: int, dxy: float, r0: float = 0.3 * ARCSEC, sigma: float = 0.05 * ARCSEC,
                        noise: float = 1e-3):
    """nxy x nxy image of exp(-(r-r0)^2/(2 sigma^2)) + U[0, noise], r measured from
    pixel [nxy/2, nxy/2] with x_j = (nxy/2 - j) dxy, y_i = (nxy/2 - i) dxy.
    Draws: nxy*nxy uniforms (row-major)."""
    c = nxy // 2 - np.arange(nxy)
    x = c * dxy
    y = c * dxy
    r = np.hypot(x[None, :], y[:, None])
    img = np.exp(-((r - r0) ** 2) / (2 * sigma ** 2)) + rng.uniform(0.0, noise, (nxy, nxy))
    return img


@dataclass
class ProfileWorkload:
    """One chi2Profile evaluation's inputs (float64 arrays)."""
    intensity: np.ndarray
    Rmin: float
    dR: float
    nxy: int
    dxy: float
    inc: float
    PA: float
    dRA: float
    dDec: float
    u: np.ndarray
    v: np.ndarray
    re: np.ndarray = field(default=None)
    im: np.ndarray = field(default=None)
    w: np.ndarray = field(default=None)

    @property
    def duv(self) -> float:
        return 1.0 / (self.dxy * self.nxy)

    def args(self):
        """Positional arguments of chi2Profile."""
        return (self.intensity, self.Rmin, self.dR, self.nxy, self.dxy, self.u, self.v, self.re, self.im, self.w)

    def kwargs(self):
        return dict(dRA=self.dRA, dDec=self.dDec, PA=self.PA, inc=self.inc)


def ring_profile_workload(model, nvis: int = 1_000_000, nxy: int = 4096, nr: int = 1000, seed: int = 1,
                          dR: float = 0.002 * ARCSEC, dxy: float = 0.01 * ARCSEC, bmin: float = 5e3,
                          bmax: float = 1e7) -> ProfileWorkload:
    """Profile-fitting workload: nr-bin ring+envelope profile (Rmin = 1e-4", dR),
    nxy x nxy image of dxy pixels, inc = 40 deg, PA = 120 deg, dRA = dDec = 0.02",
    nvis log-uniform baselines in [bmin, bmax] wavelengths, noisy observation.

    ``model(workload) -> complex visibilities`` supplies the noiseless model.
    Draws (seed): uv (log_uniform_uv), then the noise (noisy_observation)."""
    rng = np.random.default_rng(seed)
    Rmin = 1e-4 * ARCSEC
    wl = ProfileWorkload(intensity=ring_envelope_profile(nr, Rmin, dR), Rmin=Rmin, dR=dR, nxy=nxy, dxy=dxy,
                         inc=40 * DEG, PA=120 * DEG, dRA=0.02 * ARCSEC, dDec=0.02 * ARCSEC,
                         u=None, v=None)
    wl.u, wl.v = log_uniform_uv(rng, nvis, bmin, bmax)
    wl.re, wl.im, wl.w = noisy_observation(rng, model(wl))
    return wl


@dataclass
class ImageWorkload:
    image: np.ndarray
    dxy: float
    PA: float
    dRA: float
    dDec: float
    u: np.ndarray
    v: np.ndarray
    re: np.ndarray = field(default=None)
    im: np.ndarray = field(default=None)
    w: np.ndarray = field(default=None)


def ring_image_workload(model, nvis: int = 100_000, nxy: int = 1024, seed: int = 0,
                        dxy: float = 0.01 * ARCSEC) -> ImageWorkload:
    """Image-fitting workload: Gaussian ring (r0 = 0.3", sigma = 0.05") plus
    U[0, 1e-3] noise on an nxy^2 image, nvis baselines log-uniform in
    [3e4, 3e6] wavelengths, dRA = 0.05", dDec = -0.03", PA = 30 deg.
    Draws (seed): image noise, uv, observation noise."""
    rng = np.random.default_rng(seed)
    img = gaussian_ring_image(rng, nxy, dxy)
    wl = ImageWorkload(image=img, dxy=dxy, PA=30 * DEG, dRA=0.05 * ARCSEC, dDec=-0.03 * ARCSEC, u=None, v=None)
    wl.u, wl.v = log_uniform_uv(rng, nvis, 3e4, 3e6)
    wl.re, wl.im, wl.w = noisy_observation(rng, model(wl))
    return wl


def walker_ensemble(rng, nwalkers: int, nr: int, Rmin: float, dR: float, inc: float, PA: float, dRA: float,
                    dDec: float, jitter: float = 0.1):
    """Parameter sets around ring_envelope_profile(nr, Rmin, dR) and (inc, PA, dRA, dDec):
    per-bin lognormal jitter of width `jitter` on the profile; inc + 0.05 N, PA + 0.2 N,
    dRA/dDec + 0.005" N.  Returns (intensity [W, nr], geom [W, 4]).
    Draws: nwalkers*nr normals (profile), then nwalkers*4 normals (geometry)."""
    base = ring_envelope_profile(nr, Rmin, dR)
    I = base[None, :] * np.exp(jitter * rng.standard_normal((nwalkers, nr)))
    z = rng.standard_normal((nwalkers, 4))
    geom = np.stack([inc + 0.05 * z[:, 0], PA + 0.2 * z[:, 1], dRA + 0.005 * ARCSEC * z[:, 2],
                     dDec + 0.005 * ARCSEC * z[:, 3]], axis=1)
    return I, geom


@dataclass
class BatchWorkload:
    """Many parameter sets over several datasets (float64 arrays).

    ``sets``: one dict per dataset with nxy, dxy, u, v, re, im, w;
    ``intensity`` [W, nr] and ``geom`` [W, 4] = (inc, PA, dRA, dDec)."""
    nr: int
    Rmin: float
    dR: float
    sets: list
    intensity: np.ndarray
    geom: np.ndarray


def _batch_sets(rng, model, base: ProfileWorkload, shapes, bmin, bmax):
    sets = []
    for nxy, dxy, nvis in shapes:
        base.nxy, base.dxy = nxy, dxy
        base.u, base.v = log_uniform_uv(rng, nvis, bmin, bmax)
        re, im, w = noisy_observation(rng, model(base))
        sets.append(dict(nxy=nxy, dxy=dxy, u=base.u, v=base.v, re=re, im=im, w=w))
    return sets


def walkers_uneven_grids_workload(model, nwalkers: int = 256, seed: int = 3) -> BatchWorkload:
    """Batched forward workload: nwalkers parameter sets (Nr = 300 bins of the ring+envelope
    shape, Rmin = 1e-4", dR = 2"/300, walker_ensemble jitter) against three datasets of uneven
    size, (nxy, Nvis) = (512, 1e4), (1024, 2e5), (2048, 1e6), all with dxy = 0.03" and
    baselines log-uniform in [3e4, 3e6] wavelengths.

    Draws (seed): per dataset in that order its uv then its noise (the model of each dataset
    is evaluated at the base parameters inc = 40 deg, PA = 120 deg, dRA = dDec = 0.02"),
    then the walker ensemble."""
    rng = np.random.default_rng(seed)
    nr, Rmin, dR, dxy = 300, 1e-4 * ARCSEC, 2.0 * ARCSEC / 300, 0.03 * ARCSEC
    base = ProfileWorkload(intensity=ring_envelope_profile(nr, Rmin, dR), Rmin=Rmin, dR=dR, nxy=0, dxy=dxy,
                           inc=40 * DEG, PA=120 * DEG, dRA=0.02 * ARCSEC, dDec=0.02 * ARCSEC, u=None, v=None)
    sets = _batch_sets(rng, model, base, ((512, dxy, 10_000), (1024, dxy, 200_000), (2048, dxy, 1_000_000)),
                       3e4, 3e6)
    I, geom = walker_ensemble(rng, nwalkers, nr, Rmin, dR, base.inc, base.PA, base.dRA, base.dDec)
    return BatchWorkload(nr=nr, Rmin=Rmin, dR=dR, sets=sets, intensity=I, geom=geom)


def sets_three_datasets_4096_workload(model, nsets: int = 64, seed: int = 4) -> BatchWorkload:
    """Batched forward+gradient workload: nsets parameter sets (Nr = 500 bins of the ring+envelope
    shape, Rmin = 1e-4", dR = 0.004") against three datasets on one nxy = 4096, dxy = 0.01" grid
    with Nvis = 5e5, 1.5e6 and 3e6 (5e6 in total), baselines log-uniform in [5e3, 1e7] wavelengths.

    Draws (seed): as walkers_uneven_grids_workload."""
    rng = np.random.default_rng(seed)
    nr, Rmin, dR, dxy = 500, 1e-4 * ARCSEC, 0.004 * ARCSEC, 0.01 * ARCSEC
    base = ProfileWorkload(intensity=ring_envelope_profile(nr, Rmin, dR), Rmin=Rmin, dR=dR, nxy=4096, dxy=dxy,
                           inc=40 * DEG, PA=120 * DEG, dRA=0.02 * ARCSEC, dDec=0.02 * ARCSEC, u=None, v=None)
    sets = _batch_sets(rng, model, base, ((4096, dxy, 500_000), (4096, dxy, 1_500_000), (4096, dxy, 3_000_000)),
                       5e3, 1e7)
    I, geom = walker_ensemble(rng, nsets, nr, Rmin, dR, base.inc, base.PA, base.dRA, base.dDec)
    return BatchWorkload(nr=nr, Rmin=Rmin, dR=dR, sets=sets, intensity=I, geom=geom)
