"""Binned power spectrum of a Fourier field (extractPowerSpectrum,
tools.hpp:65-342, KTYPE_LINEAR), evaluated on the GPU with torch reductions
in FP64.  Measurement support for the P(k) error of the mixed-precision form.
"""
import math

import torch


def power_spectrum(ctx, F, numbins, deconvolve=True):
    """F: (N, N, P) or (1, N, N, P) complex device tensor in the library layout.
    Returns (kbin, power, kscatter, pscatter, occupation) as float64/int64 CPU tensors,
    power = <|F|^2 k^3 / sinc^2> per bin exactly as the reference defines it."""
    N = ctx.N
    H = N // 2
    F = F.reshape(N, N, ctx.P)[:, :, : H + 1]
    dev = F.device
    i = torch.arange(N, device=dev, dtype=torch.float64)
    lin = torch.where(i <= H, i, N - i)                      # |signed frequency|
    typek2 = (2 * math.pi * lin) ** 2
    s = torch.ones(N, device=dev, dtype=torch.float64)
    if deconvolve:
        nz = lin > 0
        s[nz] = torch.sin(math.pi * lin[nz] / N) * N / (math.pi * lin[nz])
    kmax = math.sqrt(3.0 * (2 * math.pi * H) ** 2)
    kx = torch.arange(H + 1, device=dev)
    k2 = typek2[:, None, None] + typek2[None, :, None] + typek2[kx][None, None, :]
    sinc2 = (s[:, None, None] * s[None, :, None] * s[kx][None, None, :]) ** 2
    w = torch.where((kx == 0) | (kx == H), 1.0, 2.0).to(torch.float64)[None, None, :].expand_as(k2)
    p = F.real.double() ** 2 + F.imag.double() ** 2
    sk = torch.sqrt(k2)
    b = torch.floor(numbins * sk / kmax).to(torch.int64)
    valid = (b < numbins)
    valid[0, 0, 0] = False
    b = b[valid]; w = w[valid]; sk = sk[valid]; k2v = k2[valid]; pv = p[valid]; sv = sinc2[valid]
    occ = torch.bincount(b, weights=w, minlength=numbins)
    ksum = torch.bincount(b, weights=w * sk, minlength=numbins)
    k2sum = torch.bincount(b, weights=w * k2v, minlength=numbins)
    psum = torch.bincount(b, weights=w * pv * k2v * sk / sv, minlength=numbins)
    p2sum = torch.bincount(b, weights=w * pv * pv * k2v ** 3 / sv / sv, minlength=numbins)
    occ_c = occ.clamp(min=1)
    kbin = torch.where(occ > 0, ksum / occ_c, torch.zeros_like(ksum))
    kscat = torch.sqrt(torch.clamp(k2sum * occ - ksum * ksum, min=0)) / occ_c
    power = torch.where(occ > 0, psum / occ_c, torch.zeros_like(psum))
    pscat = torch.sqrt(torch.clamp(p2sum / occ_c - power * power, min=0))
    return (kbin.cpu(), power.cpu(), torch.where(occ > 0, kscat, 0 * kscat).cpu(),
            torch.where(occ > 0, pscat, 0 * pscat).cpu(), occ.round().to(torch.int64).cpu())


def density_power(ctx, pcls, numbins, work=None, F=None):
    """P(k) of the CIC mass density of a particle container (deconvolved)."""
    rho = ctx.real(1) if work is None else work
    pcls.project_mass(rho)
    F = ctx.fft_r2c(rho, out=F)
    return power_spectrum(ctx, F, numbins, deconvolve=True)
