"""Writes the golden fixtures of tests/golden/ (run from the repository root):

    python tests/golden/make_golden.py

galario ships no golden vectors, so these are computed by methods independent
of both the HIP library and the FFT path of the oracle:

* ``dft_nodes.npz`` -- a 64 x 64 image swept from a radial profile (the
  oracle's sweep, itself pinned against np.interp on a meshgrid), (u, v) on
  grid nodes, rotation, offsets, and the visibilities by a long-double direct
  Fourier sum  V = sum_ij I_ij exp(+2 pi i (urot x_j + vrot y_i)) exp(+2 pi i (urot dRA' + vrot dDec'))
  (docs/tech-specs.rst:76-91 axes).  At grid nodes galario's FFT +
  interpolation path equals this sum to rounding.
* ``uvtable_head.npz`` -- the first 2000 rows of galario's docs/uvtable.txt
  (u, v in metres, Re, Im, w), converted to wavelengths with the quickstart's
  wle = 1e-3 m (docs/quickstart.rst:56-63): realistic input data only.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import galario_oracle as O  # noqa: E402

REF_UVTABLE = '/root/reference/docs/uvtable.txt'


def dft_longdouble(img, dxy, urot, vrot, dRAr, dDecr, v_origin=1.0):
    img = np.asarray(img, dtype=np.longdouble)
    N = img.shape[0]
    idx = np.arange(N, dtype=np.longdouble)
    x = (N / 2 - idx) * np.longdouble(dxy)
    y = v_origin * (N / 2 - idx) * np.longdouble(dxy)
    two_pi = 2 * np.longdouble(np.pi)
    out = np.empty(len(urot), dtype=np.clongdouble)
    for p in range(len(urot)):
        ax = two_pi * np.longdouble(urot[p]) * x
        ay = two_pi * np.longdouble(vrot[p]) * y
        ex = np.cos(ax) + 1j * np.sin(ax)
        ey = np.cos(ay) + 1j * np.sin(ay)
        out[p] = ey @ (img.astype(np.clongdouble) @ ex)
    th = two_pi * (np.longdouble(1) * urot * dRAr + vrot * dDecr)
    return out * (np.cos(th) + 1j * np.sin(th))


def make_dft_nodes():
    rng = np.random.default_rng(2024)
    N, dxy = 64, 0.05 * O.ARCSEC
    Rmin, dR, nr = 1e-4 * O.ARCSEC, 0.002 * O.ARCSEC, 900
    R = Rmin + dR * np.arange(nr)
    I = 1e10 * (np.exp(-((R - 0.6 * O.ARCSEC) / (0.2 * O.ARCSEC)) ** 2) + 0.3 * np.exp(-R / (0.4 * O.ARCSEC)))
    inc = 0.6
    img = O.sweep(I, Rmin, dR, N, dxy, inc)
    duv = 1 / (N * dxy)
    n = 256
    cols = rng.integers(-N // 2 + 2, N // 2 - 1, n).astype(np.float64)
    rows = rng.integers(-N // 2 + 2, N // 2 - 1, n).astype(np.float64)
    dRA, dDec = 0.031 * O.ARCSEC, -0.017 * O.ARCSEC
    u, v = cols * duv, rows * duv           # PA = 0: rotation is the identity
    V = dft_longdouble(img, dxy, u, v, dRA, dDec)
    np.savez_compressed(os.path.join(HERE, 'dft_nodes.npz'), intensity=I, Rmin=Rmin, dR=dR, nxy=N, dxy=dxy,
                        inc=inc, dRA=dRA, dDec=dDec, image=img, u=u, v=v,
                        vis_re=V.real.astype(np.float64), vis_im=V.imag.astype(np.float64))


def make_uvtable_head():
    if not os.path.exists(REF_UVTABLE):
        print('reference uvtable not present; uvtable_head.npz not regenerated')
        return
    tab = np.loadtxt(REF_UVTABLE, max_rows=2000)
    wle = 1e-3
    np.savez_compressed(os.path.join(HERE, 'uvtable_head.npz'), u=tab[:, 0] / wle, v=tab[:, 1] / wle,
                        re=tab[:, 2], im=tab[:, 3], w=tab[:, 4])


if __name__ == '__main__':
    make_dft_nodes()
    make_uvtable_head()
    print('written', sorted(f for f in os.listdir(HERE) if f.endswith('.npz')))
