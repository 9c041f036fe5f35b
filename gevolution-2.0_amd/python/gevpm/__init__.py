"""gevpm -- Python host layer of the MI355X-native gevolution particle-mesh step.

Thin, allocation-owning wrappers over the C ABI (include/gev_pm.h): torch
tensors provide device memory and the stream, every computation runs in the
HIP kernels of libgevpm.so.  Names follow the reference interfaces
(gevolution.hpp projection_* / prepareFTsource / solveModifiedPoissonFT /
projectFT*, perfParticles updateVel / moveParticles).

Tensor layouts:
  real field     (ncomp, N, N, N) indexed [c, z, y, x]
  Fourier field  (ncomp, N, N, P) complex, indexed [c, kz, ky, kx], P = ctx.P
"""
import ctypes as C
import os
import math

import torch

from ._lib import (GEV_FP64, GEV_MIXED, GevError, check, gev_fields, gev_gadget2_header, gev_pcls_view,
                   gev_step_state, lib)

__all__ = ["Context", "Particles", "Fields", "StepState", "GEV_FP64", "GEV_MIXED", "GevError"]


def _checked(t, name, dtype, numel, device, optional=False, at_least=False):
    """device pointer of t after checking everything the kernels assume: a
    contiguous tensor of `dtype` on `device` with exactly `numel` elements
    (at least `numel` when at_least).  None is passed through only when
    optional."""
    if t is None:
        if optional:
            return None
        raise GevError(f"{name}: a tensor is required")
    if not isinstance(t, torch.Tensor):
        raise GevError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if t.dtype != dtype:
        raise GevError(f"{name}: expected {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise GevError(f"{name}: the tensor must be contiguous")
    if (t.numel() < numel) if at_least else (t.numel() != numel):
        raise GevError(f"{name}: expected {'at least ' if at_least else ''}{numel} elements, got {t.numel()}")
    if not t.is_cuda or (device is not None and t.device != device):
        raise GevError(f"{name}: the tensor must live on {device or 'the GPU'} (got {t.device})")
    return C.c_void_p(t.data_ptr())


class Context:
    """One N^3 periodic lattice in one precision form (LATfield2 Lattice + PlanFFT)."""

    def __init__(self, N, precision=GEV_FP64, device="cuda"):
        if not torch.cuda.is_available():
            raise GevError("gevpm needs a GPU (HIP device) -- there is no CPU path")
        self.N = int(N)
        self.precision = precision
        dev = torch.device(device)
        if dev.type != "cuda":
            raise GevError(f"gevpm: device must be a GPU, got {dev}")
        self.device = dev if dev.index is not None else torch.device("cuda", torch.cuda.current_device())
        self.rdtype = torch.float64 if precision == GEV_FP64 else torch.float32
        self.cdtype = torch.complex128 if precision == GEV_FP64 else torch.complex64
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            self.stream = torch.cuda.current_stream(self.device)
            check(lib.gev_ctx_create(self.N, precision, C.c_void_p(self.stream.cuda_stream), C.byref(h)),
                  "gev_ctx_create")
        self._h = h
        # follow torch's stream exactly, including the legacy default (NULL) stream
        check(lib.gev_ctx_set_stream(h, C.c_void_p(self.stream.cuda_stream)), "gev_ctx_set_stream")
        self.P = lib.gev_ctx_kpitch(h)
        self.M = self.N ** 3

    @property
    def handle(self):
        return self._h

    # -------------------------------------------------- checked pointers
    def r(self, t, ncomp, name, optional=False):
        """pointer to a real field of ncomp components of this context"""
        return _checked(t, name, self.rdtype, ncomp * self.M, self.device, optional)

    def k(self, t, ncomp, name, optional=False, at_least=False):
        """pointer to a Fourier field of ncomp components of this context"""
        return _checked(t, name, self.cdtype, ncomp * self.N * self.N * self.P, self.device, optional, at_least)

    def ncomp_real(self, t, name):
        if not isinstance(t, torch.Tensor) or t.numel() == 0 or t.numel() % self.M:
            raise GevError(f"{name}: a real field of this context holds a multiple of {self.M} elements")
        return t.numel() // self.M

    def ncomp_fourier(self, t, name):
        m = self.N * self.N * self.P
        if not isinstance(t, torch.Tensor) or t.numel() == 0 or t.numel() % m:
            raise GevError(f"{name}: a Fourier field of this context holds a multiple of {m} elements")
        return t.numel() // m

    def bind_stream(self):
        """Make the library follow torch's current stream."""
        s = torch.cuda.current_stream(self.device)
        if s.cuda_stream != self.stream.cuda_stream:
            self.stream = s
            check(lib.gev_ctx_set_stream(self._h, C.c_void_p(s.cuda_stream)), "gev_ctx_set_stream")

    def sync(self):
        check(lib.gev_ctx_sync(self._h), "gev_ctx_sync")

    def close(self):
        if self._h:
            lib.gev_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------- allocation
    def real(self, ncomp=1):
        return torch.zeros((ncomp, self.N, self.N, self.N), dtype=self.rdtype, device=self.device)

    def real_empty(self, ncomp=1):
        """an uninitialised real field, for outputs that a call overwrites whole"""
        return torch.empty((ncomp, self.N, self.N, self.N), dtype=self.rdtype, device=self.device)

    def fourier(self, ncomp=1):
        return torch.zeros((ncomp, self.N, self.N, self.P), dtype=self.cdtype, device=self.device)

    def unpad(self, F):
        """(c, N, N, P) -> (c, N, N, N//2+1): the numpy rfftn layout."""
        return F[..., : self.N // 2 + 1]

    # -------------------------------------------------------------- FFT
    def fft_r2c(self, f, out=None):
        f = f.contiguous()
        nc = self.ncomp_real(f, "fft_r2c")
        out = self.fourier(nc) if out is None else out
        pf, po = self.r(f, nc, "fft_r2c(in)"), self.k(out, nc, "fft_r2c(out)")
        self.bind_stream()
        check(lib.gev_fft_r2c(self._h, pf, po, nc), "gev_fft_r2c")
        return out

    def fft_c2r(self, F, out=None, preserve=True):
        nc = self.ncomp_fourier(F, "fft_c2r")
        src = F.clone() if preserve else F
        out = self.real_empty(nc) if out is None else out
        pf, po = self.k(src, nc, "fft_c2r(in)"), self.r(out, nc, "fft_c2r(out)")
        self.bind_stream()
        check(lib.gev_fft_c2r(self._h, pf, po, nc), "gev_fft_c2r")
        return out

    # ----------------------------------------------------------- k-space
    def poisson(self, src, coeff, modif=0.0, out=None):
        out = torch.empty_like(src) if out is None else out
        self.bind_stream()
        check(lib.gev_poisson(self._h, self.k(src, 1, "poisson(src)"), self.k(out, 1, "poisson(out)"), coeff, modif),
              "gev_poisson")
        return out

    def project_ft_scalar(self, S, chi=None):
        add = chi is not None
        out = self.fourier(1) if chi is None else chi
        self.bind_stream()
        check(lib.gev_project_ft_scalar(self._h, self.k(S, 6, "project_ft_scalar(S)"),
                                        self.k(out, 1, "project_ft_scalar(chi)"), int(add)), "gev_project_ft_scalar")
        return out

    def project_ft_vector(self, S, coeff=1.0, modif=0.0, out=None):
        out = torch.empty_like(S) if out is None else out
        self.bind_stream()
        check(lib.gev_project_ft_vector(self._h, self.k(S, 3, "project_ft_vector(S)"),
                                        self.k(out, 3, "project_ft_vector(B)"), coeff, modif), "gev_project_ft_vector")
        return out

    def evolve_ft_vector(self, S, B, a2dtau):
        self.bind_stream()
        check(lib.gev_evolve_ft_vector(self._h, self.k(S, 6, "evolve_ft_vector(S)"), self.k(B, 3, "evolve_ft_vector(B)"),
                                       a2dtau), "gev_evolve_ft_vector")
        return B

    def set_k0(self, F, re, im=0.0):
        self.bind_stream()
        check(lib.gev_set_k0(self._h, self.k(F, 1, "set_k0", at_least=True), re, im), "gev_set_k0")

    def adj_divergence(self, u, lphi, K):
        self.bind_stream()
        check(lib.gev_adj_divergence(self._h, self.r(u, 3, "adj_divergence(u)"), self.r(lphi, 1, "adj_divergence(lphi)"),
                                     K), "gev_adj_divergence")
        return lphi

    def adj_ic_lattice(self, lx, lq, cx, cq, lpsi):
        self.bind_stream()
        check(lib.gev_adj_ic_lattice(self._h, self.r(lx, 3, "adj_ic_lattice(lx)"), self.r(lq, 3, "adj_ic_lattice(lq)"), cx,
                                     cq, self.r(lpsi, 1, "adj_ic_lattice(lpsi)")), "gev_adj_ic_lattice")
        return lpsi

    # ------------------------------------------ adjoint of the GR cycle
    def adj_source_T00(self, phi, T00, lsrc, lT00, lphi, lchi, bg, c1, c2, c3):
        """reverse of prepare_source_T00: lT00 = , lphi += , lchi += """
        self.bind_stream()
        nm = "adj_source_T00"
        check(lib.gev_adj_source_T00(self._h, self.r(phi, 1, nm + "(phi)"), self.r(T00, 1, nm + "(T00)"),
                                     self.r(lsrc, 1, nm + "(lsrc)"), self.r(lT00, 1, nm + "(lT00)"),
                                     self.r(lphi, 1, nm + "(lphi)"), self.r(lchi, 1, nm + "(lchi)"), bg, c1, c2, c3),
              "gev_adj_source_T00")
        return lT00

    def adj_source_Tij(self, phi, lS, lphi):
        """lphi += the transpose of prepare_source_Tij's phi stencil applied to lS (6)"""
        self.bind_stream()
        check(lib.gev_adj_source_Tij(self._h, self.r(phi, 1, "adj_source_Tij(phi)"), self.r(lS, 6, "adj_source_Tij(lS)"),
                                     self.r(lphi, 1, "adj_source_Tij(lphi)")), "gev_adj_source_Tij")
        return lphi

    def adj_link_factor(self, phi, L, u, lphi):
        """u (= the B solve of lB) <- lL; lphi += the link factor's phi adjoint"""
        self.bind_stream()
        check(lib.gev_adj_link_factor(self._h, self.r(phi, 1, "adj_link_factor(phi)"), self.r(L, 3, "adj_link_factor(L)"),
                                      self.r(u, 3, "adj_link_factor(u)"), self.r(lphi, 1, "adj_link_factor(lphi)")),
              "gev_adj_link_factor")
        return u

    def adj_solve_chi(self, lchi, out=None, work1=None, work6=None):
        """the adjoint of solve_chi: lS (6) = c2r(conj(weights) r2c(lchi) / (2 k^4 N))"""
        out = self.real_empty(6) if out is None else out
        work1 = self.fourier(1) if work1 is None else work1
        work6 = self.fourier(6) if work6 is None else work6
        self.fft_r2c(lchi, out=work1)
        self.bind_stream()
        check(lib.gev_adj_project_ft_scalar(self._h, self.k(work1, 1, "adj_solve_chi(work1)"),
                                            self.k(work6, 6, "adj_solve_chi(work6)")), "gev_adj_project_ft_scalar")
        return self.fft_c2r(work6, out=out, preserve=False)

    # ------------------------------------ fused real -> k -> real solves
    def solve_poisson(self, src, coeff, modif=0.0, k0=None, out=None, work=None):
        """c2r(poisson(r2c(src))), optionally with the k=0 mode set to k0 first"""
        out = torch.empty_like(src) if out is None else out
        work = self.fourier(1) if work is None else work
        self.bind_stream()
        check(lib.gev_solve_poisson(self._h, self.r(src, 1, "solve_poisson(src)"), self.r(out, 1, "solve_poisson(out)"),
                                    self.k(work, 1, "solve_poisson(work)", at_least=True), coeff, modif,
                                    int(k0 is not None), 0.0 if k0 is None else float(k0)), "gev_solve_poisson")
        return out

    def solve_poisson_T00(self, phi, chi, source, bg, c1, c2, c3, coeff, modif=0.0, k0set=False, out=None,
                          work=None):
        """prepareFTsource (T00 form) and the Poisson solve in one chain; returns
        (potential, (sum of source, sum of the source terms)).  source is only
        read for 128 <= N <= 1024 and overwritten by the source terms below."""
        out = torch.empty_like(phi) if out is None else out
        work = self.fourier(1) if work is None else work
        sums = (C.c_double * 2)()
        self.bind_stream()
        nm = "solve_poisson_T00"
        check(lib.gev_solve_poisson_T00(self._h, self.r(phi, 1, nm + "(phi)"), self.r(chi, 1, nm + "(chi)"),
                                        self.r(source, 1, nm + "(source)"), bg, c1, c2, c3, self.r(out, 1, nm + "(out)"),
                                        self.k(work, 1, nm + "(work)", at_least=True), coeff, modif, int(k0set), sums),
              "gev_solve_poisson_T00")
        return out, (sums[0], sums[1])

    def solve_chi(self, Sij, out=None, work6=None, work1=None):
        """c2r(projectFTscalar(r2c(Sij)))"""
        out = self.real_empty(1) if out is None else out
        work6 = self.fourier(6) if work6 is None else work6
        work1 = self.fourier(1) if work1 is None else work1
        self.bind_stream()
        check(lib.gev_solve_chi(self._h, self.r(Sij, 6, "solve_chi(Sij)"), self.r(out, 1, "solve_chi(out)"),
                                self.k(work6, 6, "solve_chi(work6)", at_least=True),
                                self.k(work1, 1, "solve_chi(work1)", at_least=True)), "gev_solve_chi")
        return out

    def solve_chi_Tij(self, phi, Tij, coeff, out=None, work6=None, work1=None):
        """c2r(projectFTscalar(r2c(prepareFTsource(phi, Tij, coeff)))); Tij holds Sij afterwards for N < 128"""
        out = self.real_empty(1) if out is None else out
        work6 = self.fourier(6) if work6 is None else work6
        work1 = self.fourier(1) if work1 is None else work1
        self.bind_stream()
        nm = "solve_chi_Tij"
        check(lib.gev_solve_chi_Tij(self._h, self.r(phi, 1, nm + "(phi)"), self.r(Tij, 6, nm + "(Tij)"), coeff,
                                    self.r(out, 1, nm + "(out)"), self.k(work6, 6, nm + "(work6)", at_least=True),
                                    self.k(work1, 1, nm + "(work1)", at_least=True)), "gev_solve_chi_Tij")
        return out

    def solve_vector(self, Bi, coeff=1.0, modif=0.0, work=None):
        """Bi <- c2r(projectFTvector(r2c(Bi))), in place"""
        work = self.fourier(3) if work is None else work
        self.bind_stream()
        check(lib.gev_solve_vector(self._h, self.r(Bi, 3, "solve_vector(Bi)"),
                                   self.k(work, 3, "solve_vector(work)", at_least=True), coeff, modif),
              "gev_solve_vector")
        return Bi

    def solve_vector_T0i_links(self, Bi, phi, coeff=1.0, modif=0.0, work=None):
        """Bi (T0i link sums) <- B: the link factor 1 + phi_a + phi_b, then solve_vector"""
        work = self.fourier(3) if work is None else work
        self.bind_stream()
        nm = "solve_vector_T0i_links"
        check(lib.gev_solve_vector_T0i_links(self._h, self.r(Bi, 3, nm + "(Bi)"), self.r(phi, 1, nm + "(phi)"),
                                             self.k(work, 3, nm + "(work)", at_least=True), coeff, modif),
              "gev_solve_vector_T0i_links")
        return Bi

    # ----------------------------------------------------------- sources
    def prepare_source_T00(self, phi, chi, source, bg, c1, c2, c3, result=None):
        result = torch.empty_like(source) if result is None else result
        s1, s2 = C.c_double(), C.c_double()
        self.bind_stream()
        nm = "prepare_source_T00"
        check(lib.gev_prepare_source_T00(self._h, self.r(phi, 1, nm + "(phi)"), self.r(chi, 1, nm + "(chi)"),
                                         self.r(source, 1, nm + "(source)"), bg, self.r(result, 1, nm + "(result)"),
                                         c1, c2, c3, C.byref(s1), C.byref(s2)), "gev_prepare_source_T00")
        return result, s1.value, s2.value

    def prepare_source_Tij(self, phi, Tij, coeff, Sij=None):
        Sij = torch.empty_like(Tij) if Sij is None else Sij
        self.bind_stream()
        nm = "prepare_source_Tij"
        check(lib.gev_prepare_source_Tij(self._h, self.r(phi, 1, nm + "(phi)"), self.r(Tij, 6, nm + "(Tij)"),
                                         self.r(Sij, 6, nm + "(Sij)"), coeff), "gev_prepare_source_Tij")
        return Sij

    def field_sum(self, f):
        s = C.c_double()
        f = f.contiguous()
        nc = self.ncomp_real(f, "field_sum")
        pf = self.r(f, nc, "field_sum")
        self.bind_stream()
        check(lib.gev_field_sum(self._h, pf, nc, C.byref(s)), "gev_field_sum")
        return s.value

    def ic_displacement_ft(self, seed, kL, amp, kerFT, coeff=0.0, ksphere=1, out=None):
        import numpy as np
        kL = np.ascontiguousarray(kL, np.float64)
        amp = np.ascontiguousarray(amp, np.float64)
        out = self.fourier(1) if out is None else out
        self.bind_stream()
        check(lib.gev_ic_displacement_ft(self._h, seed, ksphere, coeff, len(kL),
                                         kL.ctypes.data_as(C.POINTER(C.c_double)),
                                         amp.ctypes.data_as(C.POINTER(C.c_double)),
                                         self.k(kerFT, 1, "ic_displacement_ft(kerFT)"),
                                         self.k(out, 1, "ic_displacement_ft(out)")),
              "gev_ic_displacement_ft")
        return out


class Particles:
    """Cell-sorted particle container (perfParticles / perfParticles_gevolution)."""

    def __init__(self, ctx, capacity, mass):
        self.ctx = ctx
        self.mass = float(mass)
        h = C.c_void_p()
        check(lib.gev_pcls_create(ctx.handle, int(capacity), self.mass, C.byref(h)), "gev_pcls_create")
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib.gev_pcls_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def f(self, t, ncomp, name, optional=False):
        """pointer to a real field of ncomp components of this container's context"""
        return self.ctx.r(t, ncomp, name, optional)

    def per_particle(self, t, name):
        """pointer to a [3][n] per-particle array in the context precision"""
        return _checked(t, name, self.ctx.rdtype, 3 * self.n, self.ctx.device)

    def view(self):
        v = gev_pcls_view()
        check(lib.gev_pcls_view_get(self._h, C.byref(v)), "gev_pcls_view_get")
        return v

    @property
    def n(self):
        return self.view().n

    def load(self, pos, q, pid):
        """pos, q: (3, n) float64; pid: (n,) int64 -- host (numpy/CPU torch) or device tensors."""
        def shapes(pos, q, pid):
            if tuple(pos.shape) != (3, pos.shape[-1]) or len(pos.shape) != 2 or tuple(q.shape) != tuple(pos.shape) \
                    or tuple(pid.shape) != (pos.shape[1],):
                raise GevError(f"load: expected pos (3, n), q (3, n), pid (n,); got {tuple(pos.shape)}, "
                               f"{tuple(q.shape)}, {tuple(pid.shape)}")
            return pos.shape[1]
        if isinstance(pos, torch.Tensor) and pos.is_cuda:
            pos = pos.to(torch.float64).contiguous(); q = q.to(torch.float64).contiguous()
            pid = pid.to(torch.int64).contiguous()
            n = shapes(pos, q, pid)
            dev = self.ctx.device
            pp = _checked(pos, "load(pos)", torch.float64, 3 * n, dev)
            pq = _checked(q, "load(q)", torch.float64, 3 * n, dev)
            pi = _checked(pid, "load(pid)", torch.int64, n, dev)
            self.ctx.bind_stream()
            check(lib.gev_pcls_load(self._h, n, pp, pq, pi, 1), "gev_pcls_load")
        else:
            import numpy as np
            pos = np.ascontiguousarray(pos, np.float64); q = np.ascontiguousarray(q, np.float64)
            pid = np.ascontiguousarray(pid, np.int64)
            shapes(pos, q, pid)
            self.ctx.bind_stream()
            check(lib.gev_pcls_load(self._h, pos.shape[1], C.c_void_p(pos.ctypes.data), C.c_void_p(q.ctypes.data),
                                    C.c_void_p(pid.ctypes.data), 0), "gev_pcls_load")

    def store(self):
        """-> (pos (3,n) float64, q (3,n) float64, id (n,) int64) device tensors, in cell order."""
        n = self.n
        dev = self.ctx.device
        pos = torch.empty((3, n), dtype=torch.float64, device=dev)
        q = torch.empty((3, n), dtype=torch.float64, device=dev)
        pid = torch.empty((n,), dtype=torch.int64, device=dev)
        self.ctx.bind_stream()
        check(lib.gev_pcls_store(self._h, C.c_void_p(pos.data_ptr()), C.c_void_p(q.data_ptr()),
                                 C.c_void_p(pid.data_ptr()), 1), "gev_pcls_store")
        return pos, q, pid

    def cells(self):
        """cell index (x + N(y + N z)) of every particle, in container order"""
        out = torch.empty((max(self.n, 1),), dtype=torch.int32, device=self.ctx.device)
        self.ctx.bind_stream()
        check(lib.gev_pcls_copy_index(self._h, C.c_void_p(out.data_ptr()), None), "gev_pcls_copy_index")
        return out[: self.n]

    def cell_start(self):
        out = torch.empty((self.ctx.M + 1,), dtype=torch.int32, device=self.ctx.device)
        self.ctx.bind_stream()
        check(lib.gev_pcls_copy_index(self._h, None, C.c_void_p(out.data_ptr())), "gev_pcls_copy_index")
        return out

    def sort(self):
        self.ctx.bind_stream()
        check(lib.gev_pcls_sort(self._h), "gev_pcls_sort")

    def init_lattice(self):
        self.ctx.bind_stream()
        check(lib.gev_pcls_init_lattice(self._h), "gev_pcls_init_lattice")

    # ---------------------------------------------------------- deposits
    def project_T00(self, out, a=1.0, phi=None, coeff=1.0, accumulate=False):
        self.ctx.bind_stream()
        check(lib.gev_project_T00(self._h, self.f(out, 1, "project_T00(out)"), a,
                                  self.f(phi, 1, "project_T00(phi)", True), coeff, int(accumulate)), "gev_project_T00")
        return out

    def project_mass(self, out, coeff=1.0, accumulate=False):
        self.ctx.bind_stream()
        check(lib.gev_project_mass(self._h, self.f(out, 1, "project_mass(out)"), coeff, int(accumulate)),
              "gev_project_mass")
        return out

    def project_T0i(self, out, phi=None, coeff=1.0, accumulate=False):
        self.ctx.bind_stream()
        check(lib.gev_project_T0i(self._h, self.f(out, 3, "project_T0i(out)"), self.f(phi, 1, "project_T0i(phi)", True),
                                  coeff, int(accumulate)), "gev_project_T0i")
        return out

    def project_Tij(self, out, a=1.0, phi=None, coeff=1.0, accumulate=False):
        self.ctx.bind_stream()
        check(lib.gev_project_Tij(self._h, self.f(out, 6, "project_Tij(out)"), a,
                                  self.f(phi, 1, "project_Tij(phi)", True), coeff, int(accumulate)), "gev_project_Tij")
        return out

    def project_T00_Tij(self, T00, Tij, a=1.0, phi=None, coeff=1.0, accumulate=False):
        """project_T00 and project_Tij of the same particles in one pass"""
        self.ctx.bind_stream()
        nm = "project_T00_Tij"
        check(lib.gev_project_T00_Tij(self._h, self.f(T00, 1, nm + "(T00)"), self.f(Tij, 6, nm + "(Tij)"), a,
                                      self.f(phi, 1, nm + "(phi)", True), coeff, int(accumulate)), "gev_project_T00_Tij")
        return T00, Tij

    def project_T00_Tij_T0i_links(self, T00, Tij, links, a=1.0, phi=None, coeff=1.0):
        """project_T00 and project_Tij, and the T0i link sums without their
        1 + phi_a + phi_b factor (see Context.solve_vector_T0i_links)"""
        self.ctx.bind_stream()
        nm = "project_T00_Tij_T0i_links"
        check(lib.gev_project_T00_Tij_T0i_links(self._h, self.f(T00, 1, nm + "(T00)"), self.f(Tij, 6, nm + "(Tij)"),
                                                self.f(links, 3, nm + "(links)"), a, self.f(phi, 1, nm + "(phi)", True),
                                                coeff), "gev_project_T00_Tij_T0i_links")
        return T00, Tij, links

    # ------------------------------------------------------------ updates
    def update_q(self, dtau, phi, chi, Bi, nfield, a, bcoeff):
        r = C.c_double()
        self.ctx.bind_stream()
        check(lib.gev_update_q(self._h, dtau, self.f(phi, 1, "update_q(phi)"), self.f(chi, 1, "update_q(chi)", True),
                               self.f(Bi, 3, "update_q(Bi)", True), nfield, a, bcoeff, C.byref(r)), "gev_update_q")
        return r.value

    def update_q_newton(self, dtau, psi, a, chi=None, nfield=1):
        r = C.c_double()
        self.ctx.bind_stream()
        check(lib.gev_update_q_newton(self._h, dtau, self.f(psi, 1, "update_q_newton(psi)"),
                                      self.f(chi, 1, "update_q_newton(chi)", True), nfield, a, C.byref(r)),
              "gev_update_q_newton")
        return r.value

    def update_pos(self, dtau, phi, chi, Bi, nfield, a, bcoeff):
        self.ctx.bind_stream()
        check(lib.gev_update_pos(self._h, dtau, self.f(phi, 1, "update_pos(phi)", True),
                                 self.f(chi, 1, "update_pos(chi)", True), self.f(Bi, 3, "update_pos(Bi)", True), nfield,
                                 a, bcoeff), "gev_update_pos")

    def update_pos_newton(self, dtau, a):
        self.ctx.bind_stream()
        check(lib.gev_update_pos_newton(self._h, dtau, a), "gev_update_pos_newton")

    def kick_drift(self, gr, dtau_kick, dtau_drift, phi, chi, Bi, nf_kick, nf_drift, a_kick, a_drift, bc_kick,
                   bc_drift, want_max=True):
        """-> the kick's maximum functor value (reading it back waits for the
        stream), or None with want_max=False (no wait)"""
        r = C.c_double()
        self.ctx.bind_stream()
        nm = "kick_drift"
        check(lib.gev_kick_drift(self._h, int(gr), dtau_kick, dtau_drift, self.f(phi, 1, nm + "(phi)"),
                                 self.f(chi, 1, nm + "(chi)", True), self.f(Bi, 3, nm + "(Bi)", True), nf_kick,
                                 nf_drift, a_kick, a_drift, bc_kick, bc_drift, C.byref(r) if want_max else None, 1),
              "gev_kick_drift")
        return r.value if want_max else None

    def displace(self, xi, coeff=1.0):
        r = C.c_double()
        self.ctx.bind_stream()
        check(lib.gev_displace_particles(self._h, self.f(xi, 1, "displace(xi)"), coeff, C.byref(r)),
              "gev_displace_particles")
        return r.value

    def initialize_q(self, pot, coeff):
        r = C.c_double()
        self.ctx.bind_stream()
        check(lib.gev_initialize_q(self._h, self.f(pot, 1, "initialize_q(pot)"), coeff, C.byref(r)), "gev_initialize_q")
        return r.value

    # ------------------------------------------------------ Gadget-2 files
    def save_gadget2(self, path, header, tracer_factor=1, dtau_pos=0.0, dtau_vel=0.0, phi=None):
        """saveGadget2: the particles with ID % tracer_factor == 0; `header` is a
        gev_gadget2_header whose npart[1] equals their number"""
        self.ctx.bind_stream()
        check(lib.gev_pcls_save_gadget2(self._h, os.fsencode(path), C.byref(header), int(tracer_factor), dtau_pos,
                                        dtau_vel, self.f(phi, 1, "save_gadget2(phi)", True)), "gev_pcls_save_gadget2")

    def load_gadget2(self, path):
        """loadGadget2: replaces the particles with the snapshot's; returns its header"""
        hdr = gev_gadget2_header()
        self.ctx.bind_stream()
        check(lib.gev_pcls_load_gadget2(self._h, os.fsencode(path), C.byref(hdr)), "gev_pcls_load_gadget2")
        return hdr

    def save_express(self, path, header):
        """saveExpress: the raw particle state to "<path>.0"; fills header's particle counts"""
        self.ctx.bind_stream()
        check(lib.gev_pcls_save_express(self._h, os.fsencode(path), C.byref(header)), "gev_pcls_save_express")
        return header

    def save_gadget2_lightcone(self, path, header, lightcone, dist, dtau, dtau_old, dadtau, vertices, idlog=None,
                               phi=None, tracer_factor=1):
        """saveGadget2, lightcone form: the particles crossing the light cone of
        `lightcone` (a gevpm.lightcone.Geometry) seen from `vertices` during this
        cycle, moved to the crossing; fills header's particle counts and returns
        the number written (no file when zero).  `idlog` (gevpm.lightcone.IDLog)
        supplies the backlog and receives the prelog."""
        import numpy as np
        from .lightcone import Geometry
        geo = lightcone.c() if isinstance(lightcone, Geometry) else lightcone
        v = np.ascontiguousarray(np.asarray(vertices, np.float64).reshape(-1, 3))
        written = C.c_int64(0)
        self.ctx.bind_stream()
        check(lib.gev_pcls_save_gadget2_lightcone(
            self._h, os.fsencode(path), C.byref(header), C.byref(geo), float(dist), float(dtau), float(dtau_old),
            float(dadtau), v.ctypes.data_as(C.POINTER(C.c_double)), len(v), None if idlog is None else idlog._h,
            self.f(phi, 1, "save_gadget2_lightcone(phi)", True), int(tracer_factor), C.byref(written)),
            "gev_pcls_save_gadget2_lightcone")
        return written.value

    def load_express(self, path):
        """loadExpress: replaces the particles with the file's; returns its Gadget-2 header"""
        hdr = gev_gadget2_header()
        self.ctx.bind_stream()
        check(lib.gev_pcls_load_express(self._h, os.fsencode(path), C.byref(hdr)), "gev_pcls_load_express")
        return hdr

    # --------------------------------------------------------- checkpoints
    def state_bytes(self):
        b = C.c_int64()
        check(lib.gev_pcls_state_bytes(self._h, C.byref(b)), "gev_pcls_state_bytes")
        return b.value

    def save(self, buf=None):
        """copy the sorted container state into a device byte tensor"""
        nb = self.state_bytes()
        buf = torch.empty(nb, dtype=torch.uint8, device=self.ctx.device) if buf is None else buf
        pb = _checked(buf, "save(buf)", torch.uint8, nb, self.ctx.device, at_least=True)
        self.ctx.bind_stream()
        check(lib.gev_pcls_save(self._h, pb, buf.numel()), "gev_pcls_save")
        return buf

    def restore(self, buf):
        """the state a save wrote into buf (the library checks its count and size)"""
        pb = _checked(buf, "restore(buf)", torch.uint8, 8, self.ctx.device, at_least=True)
        self.ctx.bind_stream()
        check(lib.gev_pcls_restore(self._h, pb, buf.numel()), "gev_pcls_restore")

    # ------------------------------------------------ adjoint (Newtonian)
    def adj_deposit(self, lq, u):
        self.ctx.bind_stream()
        check(lib.gev_adj_deposit(self._h, self.per_particle(lq, "adj_deposit(lq)"), self.f(u, 3, "adj_deposit(u)")),
              "gev_adj_deposit")

    # ------------------------------------------------ adjoint (GR cycle)
    def adj_gr_kick_drift(self, phi, chi, Bi, lx, lq, lphi, lchi, lB, nf_kick, nf_drift, a_kick, a_drift, dtau_kick,
                          dtau_drift):
        """reverse of kick_drift(gr=1) from the pre-kick state: lx, lq (after the
        drift) become the adjoints before the kick; lphi, lchi, lB += """
        self.ctx.bind_stream()
        nm = "adj_gr_kick_drift"
        check(lib.gev_adj_gr_kick_drift(self._h, self.f(phi, 1, nm + "(phi)"), self.f(chi, 1, nm + "(chi)"),
                                        self.f(Bi, 3, nm + "(Bi)"), self.per_particle(lx, nm + "(lx)"),
                                        self.per_particle(lq, nm + "(lq)"), self.f(lphi, 1, nm + "(lphi)"),
                                        self.f(lchi, 1, nm + "(lchi)"), self.f(lB, 3, nm + "(lB)"), int(nf_kick),
                                        int(nf_drift), a_kick, a_drift, dtau_kick, dtau_drift), "gev_adj_gr_kick_drift")

    def adj_gr_deposit(self, phi, lT00, lTij, lL, lx, lq, lphi, a, tij_scale=1.0):
        """reverse of project_T00_Tij_T0i_links (coeff 1), the Tij adjoint being
        tij_scale lTij: lx, lq += , lphi += """
        self.ctx.bind_stream()
        nm = "adj_gr_deposit"
        check(lib.gev_adj_gr_deposit(self._h, self.f(phi, 1, nm + "(phi)", True), self.f(lT00, 1, nm + "(lT00)"),
                                     self.f(lTij, 6, nm + "(lTij)"), self.f(lL, 3, nm + "(lL)", True),
                                     self.per_particle(lx, nm + "(lx)"), self.per_particle(lq, nm + "(lq)"),
                                     self.f(lphi, 1, nm + "(lphi)", True), a, tij_scale), "gev_adj_gr_deposit")

    def adj_gather(self, lx, lq, phi=None, lrho=None, K=0.0, mN=0.0, C_prev=0.0):
        self.ctx.bind_stream()
        check(lib.gev_adj_gather(self._h, self.f(phi, 1, "adj_gather(phi)", True), self.f(lrho, 1, "adj_gather(lrho)", True),
                                 self.per_particle(lq, "adj_gather(lq)"), self.per_particle(lx, "adj_gather(lx)"), K, mN,
                                 C_prev), "gev_adj_gather")


class Fields:
    """The metric / source / work fields of main.cu (phi, chi, Bi, source, Sij and their FT work)."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.phi, self.chi, self.Bi = ctx.real(1), ctx.real(1), ctx.real(3)
        self.source, self.Sij = ctx.real(1), ctx.real(6)
        self.scalarFT, self.SijFT, self.BiFT = ctx.fourier(1), ctx.fourier(6), ctx.fourier(3)

    def c_struct(self):
        f = gev_fields()
        for name in ("phi", "chi", "Bi", "source", "Sij", "scalarFT", "SijFT", "BiFT"):
            setattr(f, name, getattr(self, name).data_ptr())
        return f

    def copy_from(self, other):
        for name in ("phi", "chi", "Bi"):
            getattr(self, name).copy_(getattr(other, name).to(getattr(self, name).dtype))


class StepState:
    """Scalars carried by the main loop (a, tau, dtau, dtau_old, maxvel ...)."""

    def __init__(self, gr, fourpiG, Omega_m, Omega_Lambda, Omega_rad=0.0, z_relax=None, elliptic=True):
        s = gev_step_state()
        s.gr = int(gr); s.vector_elliptic = int(elliptic)
        s.fourpiG = fourpiG; s.Omega_m = Omega_m; s.Omega_Lambda = Omega_Lambda; s.Omega_rad = Omega_rad
        s.z_relax = -2.0 if z_relax is None else z_relax
        self.s = s

    def __getattr__(self, k):
        if k == "s":
            raise AttributeError(k)
        return getattr(self.s, k)

    def __setattr__(self, k, v):
        if k == "s":
            object.__setattr__(self, k, v)
        else:
            setattr(self.s, k, v)


def fourpiG_of(boxsize):
    """4 pi G in box units (main.cu:418, C_SPEED_OF_LIGHT of metadata.hpp:97)."""
    return 1.5 * boxsize * boxsize / 2997.92458 / 2997.92458


def hconf(a, fourpiG, om, ol, orad=0.0):
    return lib.gev_hconf(a, fourpiG, om, ol, orad)


def rk4bg(a, fourpiG, om, ol, dtau, orad=0.0):
    return lib.gev_rk4bg(a, fourpiG, om, ol, orad, dtau)


def particle_horizon(a, fourpiG, om, ol, orad=0.0):
    return lib.gev_particle_horizon(a, fourpiG, om, ol, orad)


def ic_generate(pcls, fields, state, seed, boxsize, z_in, k_hmpc, P_mpch3, ksphere=1):
    import numpy as np
    k = np.ascontiguousarray(k_hmpc, np.float64)
    P = np.ascontiguousarray(P_mpch3, np.float64)
    if state.z_relax < -1.0:
        state.z_relax = z_in  # parser.hpp:1468 default
    md = C.c_double()
    pcls.ctx.bind_stream()
    f = fields.c_struct()
    check(lib.gev_ic_generate(pcls.handle, C.byref(f), C.byref(state.s), seed, ksphere, boxsize, z_in, len(k),
                              k.ctypes.data_as(C.POINTER(C.c_double)), P.ctypes.data_as(C.POINTER(C.c_double)),
                              C.byref(md)), "gev_ic_generate")
    return md.value


def step(pcls, fields, state):
    pcls.ctx.bind_stream()
    f = fields.c_struct()
    check(lib.gev_step(pcls.handle, C.byref(f), C.byref(state.s)), "gev_step")
    return state


PHASES = ["deposit_T00_Tij", "source_T00", "solve_phi", "source_Tij", "deposit_T0i", "solve_chi", "solve_B",
          "kick_drift", "rebin"]


def step_timing(ctx, enable=True):
    check(lib.gev_step_timing(ctx.handle, int(enable)), "gev_step_timing")


def step_timings(ctx):
    """{phase name: ms} of the last gev_step (with timing enabled)"""
    ids = (C.c_int * 24)()
    ms = (C.c_float * 24)()
    n = lib.gev_step_timings(ctx.handle, ids, ms, 24)
    if n < 0:
        check(n, "gev_step_timings")
    out = {}
    for i in range(n):
        out[PHASES[ids[i]]] = out.get(PHASES[ids[i]], 0.0) + ms[i]
    return out
