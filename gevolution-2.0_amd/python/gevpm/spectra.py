"""Binned power spectrum of a Fourier field (extractPowerSpectrum,
tools.hpp:65-342, KTYPE_LINEAR), computed on the device by gev_power_spectrum
(FP64 sums), and the deconvolved CIC density spectrum that writeSpectra
(output.hpp:3006-3033) reports for the particles.
"""
import ctypes as C

import numpy as np
import torch

from ._lib import GevError, check, lib


def power_spectrum(ctx, F, numbins, deconvolve=True, noalloc=False):
    """F: (N, N, P) or (1, N, N, P) complex device tensor in the library layout.
    Returns (kbin, power, kscatter, pscatter, occupation) as float64/int64 CPU
    tensors, power = <|F|^2 k^3 / W^2> per bin as extractPowerSpectrum defines it.
    noalloc=True calls gev_power_spectrum_noalloc (the library's static device
    workspace instead of a per-call allocation)."""
    N = ctx.N
    want = torch.complex128 if ctx.rdtype == torch.float64 else torch.complex64
    if F.dtype != want or not F.is_cuda or not F.is_contiguous() or F.numel() != N * N * ctx.P:
        raise GevError(f"power_spectrum: F must be a contiguous {want} device tensor of {N}*{N}*{ctx.P} modes")
    out = [np.zeros(numbins) for _ in range(4)]
    occ = np.zeros(numbins, np.int64)
    dp = [o.ctypes.data_as(C.POINTER(C.c_double)) for o in out]
    ctx.bind_stream()
    name = "gev_power_spectrum_noalloc" if noalloc else "gev_power_spectrum"
    check(getattr(lib, name)(ctx._h, C.c_void_p(F.data_ptr()), int(numbins), int(bool(deconvolve)), *dp,
                             occ.ctypes.data_as(C.POINTER(C.c_int64))), name)
    return tuple(torch.from_numpy(o) for o in out) + (torch.from_numpy(occ),)


def density_power(ctx, pcls, numbins, work=None, F=None):
    """P(k) of the CIC mass density of a particle container (deconvolved)."""
    rho = ctx.real(1) if work is None else work
    pcls.project_mass(rho)
    F = ctx.fft_r2c(rho, out=F)
    return power_spectrum(ctx, F, numbins, deconvolve=True)
