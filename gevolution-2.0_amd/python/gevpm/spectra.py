"""Binned power and cross spectra of Fourier fields (extractPowerSpectrum /
extractCrossSpectrum, tools.hpp:65-342), computed on the device by
gev_power_spectrum and gev_cross_spectrum (FP64 sums), and the deconvolved
CIC density spectrum that writeSpectra (output.hpp:3006-3033) reports for the
particles.
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


KTYPE_GRID, KTYPE_LINEAR = 0, 1  # tools.hpp:27-28


def cross_spectrum(ctx, F1, F2, numbins, deconvolve=True, ktype=KTYPE_LINEAR, comp1=-1, comp2=-1, symmetric=False):
    """extractCrossSpectrum (tools.hpp:65-330); extractPowerSpectrum with F2 = F1.
    F1, F2: (ncomp, N, N, P) complex device tensors of the context (ncomp 1, 3,
    or 6 for a symmetric tensor xx xy xz yy yz zz with symmetric=True).
    Returns (kbin, power, kscatter, pscatter, occupation) as gev_power_spectrum."""
    N = ctx.N
    want = torch.complex128 if ctx.rdtype == torch.float64 else torch.complex64
    m = N * N * ctx.P
    for nm, F in (("F1", F1), ("F2", F2)):
        if F.dtype != want or not F.is_cuda or not F.is_contiguous() or F.numel() % m:
            raise GevError(f"cross_spectrum: {nm} must be a contiguous {want} device tensor of components of "
                           f"{N}*{N}*{ctx.P} modes")
    if F1.numel() != F2.numel():
        raise GevError("cross_spectrum: F1 and F2 hold different numbers of components")
    ncomp = F1.numel() // m
    out = [np.zeros(numbins) for _ in range(4)]
    occ = np.zeros(numbins, np.int64)
    dp = [o.ctypes.data_as(C.POINTER(C.c_double)) for o in out]
    ctx.bind_stream()
    check(lib.gev_cross_spectrum(ctx._h, C.c_void_p(F1.data_ptr()), C.c_void_p(F2.data_ptr()), int(ncomp),
                                 int(bool(symmetric)), int(comp1), int(comp2), int(numbins), int(bool(deconvolve)),
                                 int(ktype), *dp, occ.ctypes.data_as(C.POINTER(C.c_int64))), "gev_cross_spectrum")
    return tuple(torch.from_numpy(o) for o in out) + (torch.from_numpy(occ),)


def density_power(ctx, pcls, numbins, work=None, F=None):
    """P(k) of the CIC mass density of a particle container (deconvolved)."""
    rho = ctx.real(1) if work is None else work
    pcls.project_mass(rho)
    F = ctx.fft_r2c(rho, out=F)
    return power_spectrum(ctx, F, numbins, deconvolve=True)
