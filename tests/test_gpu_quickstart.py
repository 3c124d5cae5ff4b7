"""GPU: galario's quickstart (docs/quickstart.rst:84-190) through the ``galario`` import path.

The posterior below is the quickstart's ``GaussianProfile`` / ``lnpriorfn`` /
``lnpostfn``, statement for statement, importing ``galario.double`` exactly
as the quickstart does.  It runs on the first 2000 rows of galario's
docs/uvtable.txt (tests/golden/uvtable_head.npz, u and v converted with the
quickstart's 1 mm wavelength), with the image size from ``get_image_size``.

* With the quickstart's own radial grid (Rmin = 1e-4", dR = 0.01") galario
  v1.2.2 raises in its central-pixel guard ((dxy/2 - Rmin)/dR = 2.83 < 5,
  src/galario.cpp:113-121); so does this build, with galario's message.
* With dR = 0.005" (ratio 5.65) the log-posterior matches the oracle's chi2 to
  1e-12 relative at five parameter points around the quickstart's p0.
"""
import math
import os

import numpy as np
import pytest

from oracle import galario_oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def quickstart_posterior():
    from galario import deg, arcsec
    from galario.double import chi2Profile

    def GaussianProfile(f0, sigma, Rmin, dR, nR):
        """ Gaussian brightness profile. """
        R = np.linspace(Rmin, Rmin + dR*nR, nR, endpoint=False)
        return f0 * np.exp(-0.5*(R/sigma)**2)

    def lnpriorfn(p, par_ranges):
        """ Uniform prior probability function """
        for i in range(len(p)):
            if p[i] < par_ranges[i][0] or p[i] > par_ranges[i][1]:
                return -np.inf
        jacob = -p[0]       # jacobian of the log transformation
        return jacob

    def lnpostfn(p, p_ranges, Rmin, dR, nR, nxy, dxy, u, v, Re, Im, w):
        """ Log of posterior probability function """
        lnprior = lnpriorfn(p, p_ranges)  # apply prior
        if not np.isfinite(lnprior):
            return -np.inf
        f0, sigma, inc, PA, dRA, dDec = p
        f0 = 10.**f0        # convert from log to real space
        sigma *= arcsec
        Rmin *= arcsec
        dR *= arcsec
        inc *= deg
        PA *= deg
        dRA *= arcsec
        dDec *= arcsec
        f = GaussianProfile(f0, sigma, Rmin, dR, nR)
        chi2 = chi2Profile(f, Rmin, dR, nxy, dxy, u, v, Re, Im, w,
                           inc=inc, PA=PA, dRA=dRA, dDec=dDec)
        return -0.5 * chi2 + lnprior

    return lnpostfn, GaussianProfile


def test_quickstart_posterior_through_galario_import_path():
    import galario
    from galario.double import get_image_size
    g = np.load(os.path.join(HERE, 'golden', 'uvtable_head.npz'))
    u, v, Re, Im, w = (np.require(g[k], requirements='C') for k in ('u', 'v', 're', 'im', 'w'))
    nxy, dxy = get_image_size(u, v)
    lnpostfn, GaussianProfile = quickstart_posterior()
    p_ranges = [[1, 20], [0., 8.], [0., 90.], [0., 180.], [-2., 2.], [-2., 2.]]
    p0 = [10, 0.5, 70., 60., 0., 0.]
    # the quickstart's own grid trips galario's central-pixel guard
    with pytest.raises(ValueError, match=r"Expect \(dxy/2-Rmin\)/dR > 5"):
        lnpostfn(list(p0), p_ranges, 1e-4, 0.01, 2000, nxy, dxy, u, v, Re, Im, w)
    Rmin, dR, nR = 1e-4, 0.005, 4000
    rng = np.random.default_rng(12)
    duv = 1 / (nxy * dxy)
    a, dg = galario.arcsec, galario.deg
    for k in range(5):
        p = list(np.array(p0) + rng.standard_normal(6) * np.array([0.1, 0.05, 5., 10., 0.05, 0.05]))
        got = lnpostfn(list(p), p_ranges, Rmin, dR, nR, nxy, dxy, u, v, Re, Im, w)
        f = GaussianProfile(10. ** p[0], p[1] * a, Rmin * a, dR * a, nR)
        vis = O.sample_profile(f, Rmin * a, dR * a, dxy, nxy, p[2] * dg, p[4] * a, p[5] * a, duv, p[3] * dg, u, v)
        ref = -0.5 * O.reduce_chi2(Re, Im, vis, w) - p[0]
        print(f"\np{k}: lnpost {got:.12e} oracle {ref:.12e}")
        assert math.isclose(got, ref, rel_tol=1e-12), (k, got, ref)
    # outside the prior box: -inf without calling the library
    assert lnpostfn([25, 0.5, 70., 60., 0., 0.], p_ranges, Rmin, dR, nR, nxy, dxy, u, v, Re, Im, w) == -np.inf
