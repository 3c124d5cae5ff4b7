"""numpy restatement of galario's CPU model-to-visibility path (the oracle).

TEST INFRASTRUCTURE ONLY -- never imported by the shipped library.

Every function follows the arithmetic of galario v1.2.2's OpenMP/FFTW build
(``src/galario.cpp``, compiled -std=c++11 without -march, i.e. IEEE double or
float with no FMA contraction), in the same evaluation order, so that the
image it builds is bit-identical to the reference's and the visibilities
differ from it only through the FFT's rounding (``numpy.fft`` / pocketfft in
place of FFTW).  Line numbers cite ``src/galario.cpp`` unless stated.

How each stage is pinned (``tests/test_oracle.py``):

=================  ==========================================================
stage              independent check
=================  ==========================================================
sweep              loop restatement of ``python/utils.py:209-242``
                   (``g_sweep_prototype``) and ``np.interp`` on the meshgrid
                   (``python/utils.py:245-321``, ``sweep_ref``), rtol 1e-12
central pixel      closed form on a linear profile (trapezoid rule is exact)
fftshift(s)        index-permutation definition, bitwise
rfft2              long-double brute-force DFT at N <= 32; delta / constant /
                   cosine known answers; Parseval
uv_rotate          closed form, PA=0 copy path
interpolate        reduces to the grid value at nodes (conjugate for u<0);
                   loop restatement of ``python/utils.py:390-407``
apply_phase        closed form (``python/utils.py:427-457``)
reduce_chi2        ``math.fsum`` of ``w*|V-O|^2``
whole pipeline     direct DFT (``direct_dft_vis``) at grid-node uv points
=================  ==========================================================

The single-precision variants run the same code on float32 arrays; the
reference's mixed float/double promotions in the sweep and central pixel are
reproduced where the C++ promotes to double.
"""
from __future__ import annotations

import math

import numpy as np

ARCSEC = 4.84813681109536e-06   # python/libcommon.pyx:43
DEG = 0.017453292519943295      # python/libcommon.pyx:44


class OracleError(ValueError):
    """Raised where galario raises std::invalid_argument (-> ValueError)."""


# --------------------------------------------------------------------------- #
# argument checks
# --------------------------------------------------------------------------- #
def check_input(nx: int) -> None:
    """CHECK_INPUT, galario.cpp:101-105."""
    if nx < 2:
        raise OracleError(f"x dimension = {nx} is less than 2")
    if nx % 2 != 0:
        raise OracleError(f"x dimension = {nx} is odd")


def check_input_xy(nx: int, ny: int) -> None:
    """CHECK_INPUTXY, galario.cpp:107-111."""
    if nx != ny:
        raise OracleError(f"Expect a square image but got shape ({nx}, {ny})")
    check_input(nx)


def central_pixel_ratio(dxy, Rmin, dR, dtype=np.float64):
    """(dxy/2 - Rmin)/dR as CHECK_CENTRAL_PIXEL computes it, galario.cpp:113-121.

    ``dxy / 2.`` promotes to double; the result is stored in a dreal."""
    f = np.dtype(dtype).type
    return f((float(f(dxy)) / 2.0 - float(f(Rmin))) / float(f(dR)))


def check_central_pixel(dxy, Rmin, dR, dtype=np.float64) -> None:
    r = central_pixel_ratio(dxy, Rmin, dR, dtype)
    if r < 5:
        raise OracleError("Expect (dxy/2-Rmin)/dR > 5, i.e. dR small enough to ensure reliable "
                          "interpolation the inside central pixel but got (dxy/2-Rmin)/dR = "
                          f"{float(r):f}. Try reducing dR.")


def v_origin_of(origin: str) -> float:
    """set_v_origin, python/libcommon.pyx:107-116."""
    if origin == 'upper':
        return 1.0
    if origin == 'lower':
        return -1.0
    raise AssertionError(f"Expect origin='upper' or 'lower', got {origin}")


# --------------------------------------------------------------------------- #
# sweep: radial profile -> image (create_image_h, galario.cpp:1186-1235)
# --------------------------------------------------------------------------- #
def sweep_rmax(nr, Rmin, dR, nxy, dxy, dtype=np.float64) -> int:
    """rmax = min(ceil((Rmin + nr*dR)/dxy), nxy/2), galario.cpp:1200 (dreal arithmetic)."""
    f = np.dtype(dtype).type
    val = (f(Rmin) + f(nr) * f(dR)) / f(dxy)
    return int(min(int(math.ceil(float(val))), nxy // 2))


def central_pixel_value(intensity, Rmin, dR, dxy, dtype=np.float64):
    """Flux-conserving mean intensity of the central pixel, galario.cpp:1214-1232.

    Trapezoid rule on 2*I(R)*R over [Rmin, dxy/2] divided by (dxy/2)^2-Rmin^2,
    times sr_to_px = dxy^2.  In the float build ``dxy / 2.`` and ``pow`` are
    double, so parts of the sum are carried in double exactly as the C++
    promotions do; ``flux``, ``I_interp``, ``area`` are dreal variables."""
    f = np.dtype(dtype).type
    I = np.asarray(intensity, dtype=dtype)
    Rmin, dR, dxy = f(Rmin), f(dR), f(dxy)
    iIN = int(math.floor(float(f(f(dxy / f(2)) - Rmin) / dR)))      # dxy / 2: int literal
    if iIN < 0 or iIN + 1 >= len(I):
        raise OracleError(f"central pixel needs intensity[{iIN}] and intensity[{iIN + 1}] "
                          f"but the profile has {len(I)} points")
    flux = f(0)
    for i in range(1, iIN):
        flux = f(flux + (Rmin + dR * f(i)) * I[i])
    flux = f(float(flux) * 2.0)
    flux = f(flux + (Rmin * I[0] + (Rmin + f(iIN) * dR) * I[iIN]))
    flux = f(flux * dR)
    half = float(dxy) / 2.0                                  # dxy / 2. -> double
    if dtype == np.float32:
        edge = float(Rmin + dR * f(iIN))                         # dreal
        slope = float(f((I[iIN + 1] - I[iIN]) / dR))             # dreal
        I_interp = f(slope * (half - edge) + float(I[iIN]))      # double -> dreal
        edge2 = float(Rmin + f(iIN) * dR)
        term = (float(f(f(edge2) * I[iIN])) + half * float(I_interp)) * (half - edge2)
        flux = f(float(flux) + term)
        area = f(half ** 2 - float(Rmin) ** 2)
        sr_to_px = f(dxy * dxy)
        return f(f(sr_to_px * flux) / area)
    I_interp = (I[iIN + 1] - I[iIN]) / dR * (half - (Rmin + dR * iIN)) + I[iIN]
    flux += ((Rmin + iIN * dR) * I[iIN] + half * I_interp) * (half - (Rmin + iIN * dR))
    area = half * half - Rmin * Rmin
    sr_to_px = dxy * dxy
    return sr_to_px * flux / area


def sweep_box(intensity, Rmin, dR, nxy, dxy, inc, dtype=np.float64):
    """Values of the swept (2*rmax)^2 box, sweep_core galario.cpp:1048-1071.

    Returns (box, rmax); box[i, j] is image pixel (i + nxy/2 - rmax, j + nxy/2 - rmax)."""
    f = np.dtype(dtype).type
    I = np.asarray(intensity, dtype=dtype)
    nr = len(I)
    Rmin, dR, dxy, inc = f(Rmin), f(dR), f(dxy), f(inc)
    cos_inc = f(math.cos(float(inc)))                         # libm cos in double -> dreal
    rmax = sweep_rmax(nr, Rmin, dR, nxy, dxy, dtype)
    sr_to_px = f(dxy * dxy)
    idx = np.arange(2 * rmax)
    x = (rmax - idx).astype(dtype) * dxy                      # (rmax - j) * dxy
    y = x.copy()
    xs = (x / cos_inc).astype(np.float64)                     # pow(.., 2) is double
    r2 = xs[None, :] ** 2 + (y.astype(np.float64) ** 2)[:, None]
    r = np.sqrt(r2).astype(dtype)                             # stored in a dreal
    iR = np.floor((r - Rmin) / dR).astype(np.int64)
    iR = np.maximum(iR, 0)
    valid = iR <= nr - 2
    iRc = np.where(valid, iR, 0)
    I0 = I[iRc]
    I1 = I[np.minimum(iRc + 1, nr - 1)]
    val = sr_to_px * (I0 + (r - iRc.astype(dtype) * dR - Rmin) * (I1 - I0) / dR)
    box = np.where(valid, val, f(0)).astype(dtype)
    return box, rmax


def sweep(intensity, Rmin, dR, nxy, dxy, inc, dtype=np.float64, check_central=True):
    """Full nxy x nxy image from a radial profile (create_image_h, galario.cpp:1186-1235).

    Zero image, the swept box centred on pixel [nxy/2, nxy/2], then the central
    pixel overwritten with the flux-conserving value."""
    check_input(nxy)
    if check_central:
        check_central_pixel(dxy, Rmin, dR, dtype)
    box, rmax = sweep_box(intensity, Rmin, dR, nxy, dxy, inc, dtype)
    img = np.zeros((nxy, nxy), dtype=dtype)
    o = nxy // 2 - rmax
    img[o:o + 2 * rmax, o:o + 2 * rmax] = box
    img[nxy // 2, nxy // 2] = central_pixel_value(intensity, Rmin, dR, dxy, dtype)
    return img


# --------------------------------------------------------------------------- #
# shifts and FFT
# --------------------------------------------------------------------------- #
def fftshift2(img):
    """Quadrant swap, shift_core galario.cpp:565-589 (== np.fft.fftshift on both axes)."""
    n0, n1 = img.shape
    return np.concatenate([np.concatenate([img[n0 // 2:, n1 // 2:], img[n0 // 2:, :n1 // 2]], 1),
                           np.concatenate([img[:n0 // 2, n1 // 2:], img[:n0 // 2, :n1 // 2]], 1)], 0)


def fftshift_axis0(mat):
    """Swap of the upper and lower row halves, shift_axis0_core galario.cpp:657-668."""
    n0 = mat.shape[0]
    return np.concatenate([mat[n0 // 2:], mat[:n0 // 2]], 0)


def rfft2(img):
    """Unnormalised forward 2-D R2C transform, e^{-2 pi i kn/N} (fft_h, galario.cpp:514-525)."""
    return np.fft.rfft2(img)


def spectrum_of_image(img):
    """shift -> R2C FFT -> shift_axis0 (sample_h, galario.cpp:1329-1333); shape (N, N/2+1)."""
    return fftshift_axis0(rfft2(fftshift2(img)))


# --------------------------------------------------------------------------- #
# rotation, interpolation, phase, chi2
# --------------------------------------------------------------------------- #
def uv_rotate(PA, dRA, dDec, u, v, dtype=np.float64):
    """uv_rotate_h, galario.cpp:976-999: PA == 0 copies, otherwise rotates u, v and dRA, dDec."""
    f = np.dtype(dtype).type
    u = np.asarray(u, dtype=dtype)
    v = np.asarray(v, dtype=dtype)
    PA, dRA, dDec = f(PA), f(dRA), f(dDec)
    if PA == 0:
        return dRA, dDec, u.copy(), v.copy()
    c = f(math.cos(float(PA)))
    s = f(math.sin(float(PA)))
    urot = u * c - v * s
    vrot = u * s + v * c
    return f(dRA * c - dDec * s), f(dRA * s + dDec * c), urot, vrot


def interp_indices(nrow, u, v, v_origin, duv, dtype=np.float64):
    """Fractional grid coordinates, interpolate_core galario.cpp:755-768."""
    f = np.dtype(dtype).type
    u = np.asarray(u, dtype=dtype)
    v = np.asarray(v, dtype=dtype)
    duv = f(duv)
    indu = np.abs(u) / duv
    sign_u = np.copysign(f(1), u)
    indv = f(nrow // 2) + f(v_origin) * sign_u * v / duv
    fl_u = np.floor(indu).astype(np.int64)
    fl_v = np.floor(indv).astype(np.int64)
    q = indu - fl_u.astype(dtype)
    t = indv - fl_v.astype(dtype)
    return fl_u, fl_v, q, t, sign_u


def interpolate(S, v_origin, u, v, duv, dtype=np.float64):
    """Amplitude/phase bilinear sampling of the shifted R2C spectrum, galario.cpp:749-815.

    Corners y0=[fl_v, fl_u], y1=[fl_v+1, fl_u], y2=[fl_v+1, fl_u+1],
    y3=[fl_v, fl_u+1]; the complex bilinear Z (reference evaluation order)
    gives the phase, sign-flipped for u<0; the amplitude is the bilinear of
    the four moduli."""
    S = np.asarray(S)
    nrow, ncol = S.shape
    fl_u, fl_v, q, t, sign_u = interp_indices(nrow, u, v, v_origin, duv, dtype)
    flat = S.reshape(-1)
    base = fl_u + fl_v * ncol            # linear index of y0, as in the reference
    y0 = flat[base]
    y1 = flat[base + ncol]
    y2 = flat[base + ncol + 1]
    y3 = flat[base + 1]
    term1 = t * (q * ((y0 + y2) - (y1 + y3)))
    term2 = t * (y1 - y0)
    term3 = q * (y3 - y0)
    Z = (term1 + (term2 + term3)) + y0
    phase = np.angle(Z) * sign_u
    a0, a1, a2, a3 = np.abs(y0), np.abs(y1), np.abs(y2), np.abs(y3)
    amp = a0
    amp = amp + (a3 - a0) * q
    amp = amp + (a1 - a0) * t
    amp = amp + (((a0 - a1) + a2) - a3) * t * q
    cdt = np.complex64 if np.dtype(dtype) == np.float32 else np.complex128
    out = np.empty(len(amp), dtype=cdt)
    # cos/sin of a dreal go through the double libm call, then back to dreal
    out.real = amp * np.cos(phase.astype(np.float64)).astype(dtype)
    out.imag = amp * np.sin(phase.astype(np.float64)).astype(dtype)
    return out


def apply_phase(dRA, dDec, u, v, vis, dtype=np.float64):
    """apply_phase_sampled_h, galario.cpp:908-921 / core 877-884: V *= e^{i 2 pi (u dRA + v dDec)}."""
    f = np.dtype(dtype).type
    vis = np.array(vis, copy=True)
    dRA, dDec = f(dRA), f(dDec)
    if dRA == 0 and dDec == 0:
        return vis
    dRA = f(float(dRA) * (2.0 * float(f(np.pi))))
    dDec = f(float(dDec) * (2.0 * float(f(np.pi))))
    u = np.asarray(u, dtype=dtype)
    v = np.asarray(v, dtype=dtype)
    ang = u * dRA + v * dDec
    c = np.cos(ang.astype(np.float64)).astype(dtype)
    s = np.sin(ang.astype(np.float64)).astype(dtype)
    re = vis.real * c - vis.imag * s
    im = vis.real * s + vis.imag * c
    vis.real = re
    vis.imag = im
    return vis


def chi2_terms(vis_obs_re, vis_obs_im, vis, w, dtype=np.float64):
    """Per-point (V-O)*sqrt(w) squared modulus, diff_weighted_core galario.cpp:1453-1461."""
    sw = np.sqrt(np.asarray(w, dtype=dtype))
    a = (vis.real - np.asarray(vis_obs_re, dtype=dtype)) * sw
    b = (vis.imag - np.asarray(vis_obs_im, dtype=dtype)) * sw
    return a.astype(np.float64) ** 2 + b.astype(np.float64) ** 2


def reduce_chi2(vis_obs_re, vis_obs_im, vis, w, dtype=np.float64):
    """chi2 = sum |(V-O) sqrt(w)|^2 (reduce_chi2, galario.cpp:1506-1538), summed exactly with fsum."""
    return math.fsum(chi2_terms(vis_obs_re, vis_obs_im, vis, w, dtype).tolist())


# --------------------------------------------------------------------------- #
# whole pipelines
# --------------------------------------------------------------------------- #
def sample_spectrum(S, v_origin, dRA, dDec, duv, PA, u, v, dtype=np.float64):
    """rotate -> interpolate -> phase on a precomputed shifted spectrum (sample_h 1335-1345)."""
    dRArot, dDecrot, urot, vrot = uv_rotate(PA, dRA, dDec, u, v, dtype)
    vis = interpolate(S, v_origin, urot, vrot, duv, dtype)
    return apply_phase(dRArot, dDecrot, urot, vrot, vis, dtype)


def sample_image(image, v_origin, dRA, dDec, duv, PA, u, v, dtype=np.float64):
    """sample_image, galario.cpp:1359-1390."""
    image = np.asarray(image, dtype=dtype)
    check_input(image.shape[0])
    S = spectrum_of_image(image)
    return sample_spectrum(S, v_origin, dRA, dDec, duv, PA, u, v, dtype)


def sample_profile(intensity, Rmin, dR, dxy, nxy, inc, dRA, dDec, duv, PA, u, v,
                   dtype=np.float64, check_central=True):
    """sample_profile, galario.cpp:1404-1436 (v_origin = +1)."""
    img = sweep(intensity, Rmin, dR, nxy, dxy, inc, dtype, check_central)
    return sample_image(img, 1.0, dRA, dDec, duv, PA, u, v, dtype)


def chi2_image(image, v_origin, dRA, dDec, duv, PA, u, v, re, im, w, dtype=np.float64):
    """chi2_image, galario.cpp:1556-1606."""
    check_input_xy(*np.asarray(image).shape)
    vis = sample_image(image, v_origin, dRA, dDec, duv, PA, u, v, dtype)
    return reduce_chi2(re, im, vis, w, dtype)


def chi2_profile(intensity, Rmin, dR, dxy, nxy, inc, dRA, dDec, duv, PA, u, v, re, im, w,
                 dtype=np.float64, check_central=True):
    """chi2_profile, galario.cpp:1615-1656."""
    vis = sample_profile(intensity, Rmin, dR, dxy, nxy, inc, dRA, dDec, duv, PA, u, v,
                         dtype, check_central)
    return reduce_chi2(re, im, vis, w, dtype)


# --------------------------------------------------------------------------- #
# independent evaluators used to pin the restatement
# --------------------------------------------------------------------------- #
def direct_dft_vis(image, dxy, u, v, dRA=0.0, dDec=0.0, PA=0.0, v_origin=1.0, chunk=4096):
    """Direct Fourier sum of an image at (u, v), no FFT and no interpolation.

    V(u,v) = sum_ij I[i,j] exp(+2 pi i (urot x_j + vrot y_i)) exp(+2 pi i (urot dRArot + vrot dDecrot))
    with x_j = (N/2 - j) dxy and y_i = v_origin (N/2 - i) dxy (pixel centres;
    docs/tech-specs.rst:76-91).  Only the rows/columns holding non-zero pixels
    are summed.  Separable evaluation: (E_y * (img @ E_x^T)).sum(0)."""
    img = np.asarray(image, dtype=np.float64)
    N = img.shape[0]
    rows = np.nonzero(np.any(img != 0, axis=1))[0]
    cols = np.nonzero(np.any(img != 0, axis=0))[0]
    if rows.size == 0:
        return np.zeros(len(u), dtype=np.complex128)
    r0, r1 = rows.min(), rows.max() + 1
    c0, c1 = cols.min(), cols.max() + 1
    sub = img[r0:r1, c0:c1]
    xj = (N / 2 - np.arange(c0, c1)) * dxy
    yi = v_origin * (N / 2 - np.arange(r0, r1)) * dxy
    cPA, sPA = math.cos(PA), math.sin(PA)
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    if PA == 0:
        urot, vrot, dRAr, dDecr = u, v, dRA, dDec
    else:
        urot = u * cPA - v * sPA
        vrot = u * sPA + v * cPA
        dRAr = dRA * cPA - dDec * sPA
        dDecr = dRA * sPA + dDec * cPA
    out = np.empty(len(u), dtype=np.complex128)
    for s in range(0, len(u), chunk):
        ur, vr = urot[s:s + chunk], vrot[s:s + chunk]
        Ex = np.exp(2j * np.pi * np.outer(xj, ur))      # (ncols, P)
        Ey = np.exp(2j * np.pi * np.outer(yi, vr))      # (nrows, P)
        out[s:s + chunk] = np.einsum('ip,ip->p', Ey, sub @ Ex)
    theta = 2 * np.pi * (urot * dRAr + vrot * dDecr)
    return out * np.exp(1j * theta)


def brute_rfft2(img):
    """Long-double O(N^4) R2C DFT for tiny images (pins rfft2)."""
    img = np.asarray(img, dtype=np.longdouble)
    N0, N1 = img.shape
    C = N1 // 2 + 1
    n0 = np.arange(N0, dtype=np.longdouble)
    n1 = np.arange(N1, dtype=np.longdouble)
    two_pi = 2 * np.longdouble(np.pi)
    out = np.zeros((N0, C), dtype=np.clongdouble)
    for k0 in range(N0):
        a0 = -two_pi * np.longdouble(k0) * n0 / N0
        for k1 in range(C):
            a = a0[:, None] - two_pi * np.longdouble(k1) * n1[None, :] / N1
            out[k0, k1] = np.sum(img * np.cos(a)) + 1j * np.sum(img * np.sin(a))
    return out
