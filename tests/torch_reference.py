"""torch composition of galario's chi2Profile for reference gradients (CPU, float64).

Differentiable restatement of the stages with torch ops (sweep
src/galario.cpp:1048-1071 and central pixel 1214-1232, quadrant shift,
torch.fft.rfft2, row shift, rotation 955-960, amplitude/phase interpolation
749-815 on the complex spectrum, phase 877-884, chi2 1453-1461); autograd
gives the gradient with respect to the profile and inc, PA, dRA, dDec.  It
shares no code with the HIP adjoint and works on the complex spectrum (the
HIP path uses the real, even one).
"""
import math

import torch


def sweep_image(I, Rmin, dR, nxy, dxy, inc):
    nr = I.shape[0]
    rmax = min(int(math.ceil((Rmin + nr * dR) / dxy)), nxy // 2)
    idx = torch.arange(nxy, dtype=torch.float64)
    xint = (nxy // 2 - idx)                      # x = (N/2 - j) dxy, y = (N/2 - i) dxy
    x = xint * dxy
    y = xint * dxy
    xs = x[None, :] / torch.cos(inc)
    r2 = xs ** 2 + (y ** 2)[:, None]
    r = torch.sqrt(torch.clamp(r2, min=1e-300))
    iR = torch.clamp(torch.floor((r.detach() - Rmin) / dR), min=0).long()
    valid = iR <= nr - 2
    iRc = torch.where(valid, iR, torch.zeros_like(iR))
    I0 = I[iRc]
    I1 = I[torch.clamp(iRc + 1, max=nr - 1)]
    val = dxy * dxy * (I0 + (r - iRc.double() * dR - Rmin) * (I1 - I0) / dR)
    inbox = ((idx >= nxy // 2 - rmax) & (idx < nxy // 2 + rmax))
    box = inbox[:, None] & inbox[None, :]
    img = torch.where(valid & box, val, torch.zeros_like(val))
    # central pixel
    iIN = int(math.floor((dxy / 2 - Rmin) / dR))
    Ri = Rmin + dR * torch.arange(iIN + 2, dtype=torch.float64)
    flux = dR * (Rmin * I[0] + 2 * (Ri[1:iIN] * I[1:iIN]).sum() + (Rmin + iIN * dR) * I[iIN])
    h = dxy / 2
    I_interp = (I[iIN + 1] - I[iIN]) / dR * (h - (Rmin + dR * iIN)) + I[iIN]
    flux = flux + ((Rmin + iIN * dR) * I[iIN] + h * I_interp) * (h - (Rmin + iIN * dR))
    cval = dxy * dxy * flux / (h * h - Rmin * Rmin)
    center = torch.zeros(nxy, nxy, dtype=torch.bool)
    center[nxy // 2, nxy // 2] = True
    return torch.where(center, cval, img)


def chi2_profile(I, geom, Rmin, dR, nxy, dxy, u, v, re, im, w):
    inc, PA, dRA, dDec = geom[0], geom[1], geom[2], geom[3]
    img = sweep_image(I, Rmin, dR, nxy, dxy, inc)
    S = torch.fft.fftshift(torch.fft.rfft2(torch.fft.fftshift(img)), dim=0)
    c, s = torch.cos(PA), torch.sin(PA)
    urot = u * c - v * s
    vrot = u * s + v * c
    dRAr = dRA * c - dDec * s
    dDecr = dRA * s + dDec * c
    duv = 1.0 / (nxy * dxy)
    sg = torch.where(urot.detach() < 0, -1.0, 1.0).double()
    indu = torch.abs(urot) / duv
    indv = nxy // 2 + sg * vrot / duv
    fu = torch.floor(indu.detach()).long()
    fv = torch.floor(indv.detach()).long()
    q = indu - fu.double()
    t = indv - fv.double()
    y0, y1, y2, y3 = S[fv, fu], S[fv + 1, fu], S[fv + 1, fu + 1], S[fv, fu + 1]
    Z = y0 * (1 - t) * (1 - q) + y1 * t * (1 - q) + y2 * t * q + y3 * (1 - t) * q
    A = (torch.abs(y0) * (1 - t) * (1 - q) + torch.abs(y1) * t * (1 - q) + torch.abs(y2) * t * q
         + torch.abs(y3) * (1 - t) * q)
    ph = sg * torch.atan2(Z.imag, Z.real)
    theta = 2 * math.pi * (urot * dRAr + vrot * dDecr)
    Vre = A * torch.cos(ph + theta)
    Vim = A * torch.sin(ph + theta)
    return (((Vre - re) ** 2 + (Vim - im) ** 2) * w).sum()
