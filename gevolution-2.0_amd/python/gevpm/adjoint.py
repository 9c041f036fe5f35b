"""Differentiable Newtonian particle-mesh run.

`PMRun` maps an initial displacement potential psi (an N^3 mesh field, box
units squared) to the final CIC mass density after `steps` Newtonian cycles
(main.cu:655-765 and 1227-1266 with the Newton functors):

    x0 = x_lat + cx N grad1(psi),  q0 = cq N grad1(psi)   (lattice Zel'dovich map)
    per cycle: rho = CIC(x); phi = Poisson(rho; 4 pi G / a); kick; drift; re-bin
    output: rho_final = CIC(x_final)

`pm_density(psi, run)` is the torch.autograd.Function entry: its backward runs
the hand-written reverse kernels of the library (gev_adj_*), recomputing each
cycle's fields from per-cycle container checkpoints (every `ckpt_every`
cycles, with the cycles in between replayed from the checkpoint), so memory
grows with steps / ckpt_every states plus ckpt_every replayed ones.

x_lat is the cell-centred lattice (one particle per cell, ID = cell); the
Zel'dovich velocity uses q = a dx/dtau with dx/dtau = f(a) H(a) (x - x_lat),
f = Omega_m(a)^0.55, so cq = a f H cx.
"""
import math

import torch

from . import (GEV_FP64, Context, Particles, fourpiG_of, hconf, particle_horizon, rk4bg)


class PMRun:
    def __init__(self, N, boxsize, z_in, z_end, steps, precision=GEV_FP64, ckpt_every=1, omega_m=0.31,
                 omega_lambda=0.69, device="cuda"):
        self.N, self.steps, self.ckpt_every = N, int(steps), max(1, int(ckpt_every))
        self.om, self.ol = omega_m, omega_lambda
        self.ctx = Context(N, precision, device)
        self.M = N ** 3
        self.pcls = Particles(self.ctx, self.M, omega_m / self.M)
        self.fpg = fourpiG_of(boxsize)
        c = self.ctx
        self.rho, self.phi, self.lphi, self.lrho = c.real(1), c.real(1), c.real(1), c.real(1)
        self.u = c.real(3)
        self.work = c.fourier(1)
        # cycle schedule: uniform steps in conformal time (the same as bench.py)
        a = 1.0 / (1.0 + z_in)
        tau0 = particle_horizon(a, self.fpg, omega_m, omega_lambda)
        dtau = (particle_horizon(1.0 / (1.0 + z_end), self.fpg, omega_m, omega_lambda) - tau0) / self.steps
        H = hconf(a, self.fpg, omega_m, omega_lambda)
        om_a = omega_m / (omega_m + omega_lambda * a ** 3)
        self.cx = 1.0
        self.cq = a * om_a ** 0.55 * H * self.cx
        self.sched = []
        dtau_old = 0.0
        for _ in range(self.steps):
            a_d = rk4bg(a, self.fpg, omega_m, omega_lambda, 0.5 * dtau)
            self.sched.append(dict(a=a, a_d=a_d, dtau_k=0.5 * (dtau + dtau_old), dtau=dtau))
            a = rk4bg(a_d, self.fpg, omega_m, omega_lambda, 0.5 * dtau)
            dtau_old = dtau
        self.a_final = a
        self.mN = omega_m * N  # particle mass / dx^3 * N
        self._ckpts = {}
        self._final = None
        self.mem_probe = None  # optional callable, run where device memory use peaks

    # ------------------------------------------------------------ forward
    def _ic(self, psi):
        p = self.pcls
        p.init_lattice()
        p.initialize_q(psi, -self.cq)  # q = cq N grad1(psi) at the lattice point
        p.displace(psi, self.cx)       # x = x_lat + cx N grad1(psi)

    def _cycle(self, n):
        s, p, c = self.sched[n], self.pcls, self.ctx
        p.project_mass(self.rho)
        c.solve_poisson(self.rho, self.fpg / s["a"], 0.0, out=self.phi, work=self.work)
        p.kick_drift(0, s["dtau_k"], s["dtau"], self.phi, None, None, 1, 0, s["a"], s["a_d"], 1.0, 1.0, want_max=False)

    def forward(self, psi):
        psi = psi.reshape(1, self.N, self.N, self.N).to(self.ctx.rdtype).contiguous()
        self._ic(psi)
        self._ckpts = {}
        for n in range(self.steps):
            if n % self.ckpt_every == 0:
                self._ckpts[n] = self.pcls.save()
            self._cycle(n)
        self._final = self.pcls.save()
        out = self.ctx.real(1)
        self.pcls.project_mass(out)
        if self.mem_probe:
            self.mem_probe()
        return out

    def checkpoint_bytes(self):
        return sum(b.numel() for b in self._ckpts.values()) + (self._final.numel() if self._final is not None else 0)

    # ----------------------------------------------------------- backward
    def backward(self, g_rho):
        p, c, n = self.pcls, self.ctx, self.M
        g_rho = g_rho.reshape(1, self.N, self.N, self.N).to(c.rdtype).contiguous()
        lx = torch.zeros(3, n, dtype=c.rdtype, device=c.device)
        lq = torch.zeros_like(lx)
        def c_of(k):  # drift factor of cycle k (0 before the first cycle)
            return self.sched[k]["dtau"] / self.sched[k]["a_d"] if k >= 0 else 0.0

        p.restore(self._final)
        # d rho_final / d x_final, then lq += C lx for the last cycle
        p.adj_gather(lx, lq, lrho=g_rho, mN=self.mN, C_prev=c_of(self.steps - 1))
        starts = sorted(self._ckpts)
        for si in range(len(starts) - 1, -1, -1):
            k0 = starts[si]
            k1 = starts[si + 1] if si + 1 < len(starts) else self.steps
            # replay the segment from its checkpoint to get every cycle's start
            # state and potential (the last cycle's potential is formed below)
            states = {k0: self._ckpts[k0]}
            phis = {}
            if k1 - k0 > 1:
                p.restore(self._ckpts[k0])
                for k in range(k0, k1 - 1):
                    self._cycle(k)
                    phis[k] = self.phi.clone()
                    if k + 1 < k1 - 1:
                        states[k + 1] = p.save()
                # the container now holds the start state of cycle k1 - 1
            if self.mem_probe:
                self.mem_probe()
            for k in range(k1 - 1, k0 - 1, -1):
                s = self.sched[k]
                K = s["dtau_k"] * s["a"] * self.N
                if k in states:  # the backward kernels leave the particles as they are
                    p.restore(states[k])
                phi = phis.pop(k, None)
                if phi is None:
                    p.project_mass(self.rho)
                    phi = c.solve_poisson(self.rho, self.fpg / s["a"], 0.0, out=self.phi, work=self.work)
                p.adj_deposit(lq, self.u)
                c.adj_divergence(self.u, self.lphi, K)
                c.solve_poisson(self.lphi, self.fpg / s["a"], 0.0, out=self.lrho, work=self.work)
                p.adj_gather(lx, lq, phi=phi, lrho=self.lrho, K=K, mN=self.mN, C_prev=c_of(k - 1))
                del phi
            del states
        lpsi = c.real(1)
        c.adj_ic_lattice(lx, lq, self.cx, self.cq, lpsi)
        return lpsi


class _PMFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, psi, run):
        ctx.run = run
        return run.forward(psi.detach())

    @staticmethod
    def backward(ctx, g):
        return ctx.run.backward(g.contiguous()), None


def pm_density(psi, run):
    """final CIC mass density of `run` from the displacement potential psi, differentiable in psi"""
    return _PMFunction.apply(psi, run)


class PMRunGR:
    """Differentiable weak-field GR run: the lattice Zel'dovich map of PMRun,
    then `steps` cycles of main.cu's GR loop with the elliptic B
    (main.cu:608-1317, in gev_step's order), then the final CIC mass density:

        T00, Tij, L = deposits(x, q; phi_old)   (L: T0i link sums)
        phi = Poisson_modified(prepareFTsource_T00(T00, phi_old, chi_old))
        chi = projectFTscalar(prepareFTsource_Tij(phi, Tij))
        B   = projectFTvector(L (1 + phi_a + phi_b))
        kick (update_q, nfield 3) and drift (update_pos, nfield 3)

    The run starts from phi = chi = 0; its first cycle takes dtau_old = dtau
    for the modified Poisson solve (modif 3 H / dtau_old) and half a kick.
    The backward replays every cycle's fields from a per-cycle checkpoint of
    the particles and of phi and chi (segments of ckpt_every cycles, each
    replayed once from its first checkpoint) and runs the reverse kernels
    (gev_adj_gr_*, gev_adj_source_*, gev_adj_link_factor, the adjoint chi
    solve; the Poisson and B solves are their own adjoints)."""

    def __init__(self, N, boxsize, z_in, z_end, steps, precision=GEV_FP64, ckpt_every=1, omega_m=0.31,
                 omega_lambda=0.69, device="cuda"):
        self.N, self.steps, self.ckpt_every = N, int(steps), max(1, int(ckpt_every))
        self.om, self.ol = omega_m, omega_lambda
        self.ctx = Context(N, precision, device)
        self.M = N ** 3
        self.pcls = Particles(self.ctx, self.M, omega_m / self.M)
        self.fpg = fourpiG_of(boxsize)
        c = self.ctx
        self.dx = 1.0 / N
        self.phi, self.chi = c.real(1), c.real(1)
        self.T00, self.Tij, self.L = c.real(1), c.real(6), c.real(3)
        self.src, self.Bi = c.real(1), c.real(3)
        self.work1, self.work3, self.work6 = c.fourier(1), c.fourier(3), c.fourier(6)
        a = 1.0 / (1.0 + z_in)
        tau0 = particle_horizon(a, self.fpg, omega_m, omega_lambda)
        dtau = (particle_horizon(1.0 / (1.0 + z_end), self.fpg, omega_m, omega_lambda) - tau0) / self.steps
        H = hconf(a, self.fpg, omega_m, omega_lambda)
        om_a = omega_m / (omega_m + omega_lambda * a ** 3)
        self.cx = 1.0
        self.cq = a * om_a ** 0.55 * H * self.cx
        self.sched = []
        dtau_prev = 0.0
        for _ in range(self.steps):
            a_d = rk4bg(a, self.fpg, omega_m, omega_lambda, 0.5 * dtau)
            self.sched.append(dict(a=a, a_d=a_d, dtau_k=0.5 * (dtau + dtau_prev), dtau=dtau,
                                   dtau_old=dtau_prev if dtau_prev > 0 else dtau,
                                   H=hconf(a, self.fpg, omega_m, omega_lambda)))
            a = rk4bg(a_d, self.fpg, omega_m, omega_lambda, 0.5 * dtau)
            dtau_prev = dtau
        self.a_final = a
        self.mN = omega_m * N
        self._ckpts = {}
        self._final = None
        self.mem_probe = None

    def _coef(self, k):
        s, dx = self.sched[k], self.dx
        return dict(c1=3.0 * s["H"] * dx * dx / s["dtau_old"], c2=self.fpg * dx * dx / s["a"],
                    c3=3.0 * s["H"] ** 2 * dx * dx, pcoeff=1.0 / (dx * dx), modif=3.0 * s["H"] / s["dtau_old"],
                    ccoeff=2.0 * self.fpg * dx * dx / s["a"], bcoeff=self.fpg * dx * dx)

    def _fields(self, k, phi_old, chi_old):
        """the cycle's deposits and metric solves from the container's state;
        returns (phi, chi, B) of the cycle (new tensors) and fills self.T00,
        self.Tij, self.L"""
        s, c, p, k_ = self.sched[k], self.ctx, self.pcls, self._coef(k)
        p.project_T00_Tij_T0i_links(self.T00, self.Tij, self.L, a=s["a"], phi=phi_old, coeff=1.0)
        c.prepare_source_T00(phi_old, chi_old, self.T00, self.om, k_["c1"], k_["c2"], k_["c3"], result=self.src)
        phi = c.solve_poisson(self.src, k_["pcoeff"], k_["modif"], work=self.work1)
        Sij = c.prepare_source_Tij(phi, self.Tij, k_["ccoeff"])
        chi = c.solve_chi(Sij, work6=self.work6, work1=self.work1)
        B = self.L.clone()
        c.solve_vector_T0i_links(B, phi, k_["bcoeff"], 0.0, work=self.work3)
        return phi, chi, B

    def _cycle(self, k, keep=False):
        """cycle k; with keep, returns the fields the reverse cycle reads
        (phi, chi, B, T00, links) so that it need not recompute them"""
        s, p = self.sched[k], self.pcls
        phi, chi, B = self._fields(k, self.phi, self.chi)
        kept = None
        if keep:  # hand the deposit buffers over and deposit into fresh ones next time
            kept = (phi, chi, B, self.T00, self.L)
            self.T00, self.L = torch.empty_like(self.T00), torch.empty_like(self.L)
        N = self.N
        p.kick_drift(1, s["dtau_k"], s["dtau"], phi, chi, B, 3, 3, s["a"], s["a_d"], s["a"] ** 2 * N,
                     s["a_d"] ** 2 * N, want_max=False)
        self.phi, self.chi = phi, chi
        return kept

    def _ic(self, psi):
        p = self.pcls
        p.init_lattice()
        p.initialize_q(psi, -self.cq)
        p.displace(psi, self.cx)
        self.phi, self.chi = self.ctx.real(1), self.ctx.real(1)  # fresh: saved states may share the old ones

    def _save(self):
        return (self.pcls.save(), self.phi.clone(), self.chi.clone())

    def _restore(self, st):
        # phi and chi are never written in place once saved: no copies needed
        self.pcls.restore(st[0])
        self.phi, self.chi = st[1], st[2]

    def forward(self, psi):
        psi = psi.reshape(1, self.N, self.N, self.N).to(self.ctx.rdtype).contiguous()
        self._ic(psi)
        self._ckpts = {}
        for k in range(self.steps):
            if k % self.ckpt_every == 0:
                self._ckpts[k] = self._save()
            self._cycle(k)
        self._final = self.pcls.save()
        out = self.ctx.real(1)
        self.pcls.project_mass(out)
        if self.mem_probe:
            self.mem_probe()
        return out

    def checkpoint_bytes(self):
        nb = sum(t[0].numel() + t[1].numel() * t[1].element_size() + t[2].numel() * t[2].element_size()
                 for t in self._ckpts.values())
        return nb + (self._final.numel() if self._final is not None else 0)

    def _reverse_cycle(self, k, phi_old, chi_old, lx, lq, lphi, lchi, kept=None):
        """the adjoint of cycle k from its start state (in the container, with
        phi_old, chi_old); lphi, lchi: adjoints of the cycle's phi and chi
        (consumed); kept: the cycle's fields from the replay (phi, chi, B,
        T00, links), else they are recomputed.  Returns the adjoints of
        phi_old and chi_old."""
        s, c, p, k_ = self.sched[k], self.ctx, self.pcls, self._coef(k)
        N = self.N
        if kept is None:
            phi, chi, B = self._fields(k, phi_old, chi_old)
            T00, L = self.T00, self.L
        else:
            phi, chi, B, T00, L = kept
        lB = c.real(3)
        p.adj_gr_kick_drift(phi, chi, B, lx, lq, lphi, lchi, lB, 3, 3, s["a"], s["a_d"], s["dtau_k"], s["dtau"])
        # B = V (L (1 + phi_a + phi_b)), V self-adjoint
        c.solve_vector(lB, k_["bcoeff"], 0.0, work=self.work3)
        c.adj_link_factor(phi, L, lB, lphi)                # lB now holds lL
        # chi = C Sij, Sij = stencil(phi) + ccoeff Tij
        lS = c.adj_solve_chi(lchi, work1=self.work1, work6=self.work6)
        c.adj_source_Tij(phi, lS, lphi)  # Tij's adjoint is ccoeff lS: the deposit's reverse scales it
        # phi = P src, P self-adjoint
        lsrc = c.solve_poisson(lphi, k_["pcoeff"], k_["modif"], work=self.work1)
        lphi_old, lchi_old = c.real(1), c.real(1)
        lT00 = torch.empty_like(lphi_old)  # overwritten by adj_source_T00
        c.adj_source_T00(phi_old, T00, lsrc, lT00, lphi_old, lchi_old, self.om, k_["c1"], k_["c2"], k_["c3"])
        p.adj_gr_deposit(phi_old, lT00, lS, lB, lx, lq, lphi_old, s["a"], k_["ccoeff"])
        return lphi_old, lchi_old

    def backward(self, g_rho):
        p, c, n = self.pcls, self.ctx, self.M
        g_rho = g_rho.reshape(1, self.N, self.N, self.N).to(c.rdtype).contiguous()
        lx = torch.zeros(3, n, dtype=c.rdtype, device=c.device)
        lq = torch.zeros_like(lx)
        p.restore(self._final)
        p.adj_gather(lx, lq, lrho=g_rho, mN=self.mN)
        lphi, lchi = c.real(1), c.real(1)
        starts = sorted(self._ckpts)
        for si in range(len(starts) - 1, -1, -1):
            k0 = starts[si]
            k1 = starts[si + 1] if si + 1 < len(starts) else self.steps
            states = {k0: self._ckpts[k0]}
            kept = {}  # the fields the replay formed, by cycle
            held = -1
            if k1 - k0 > 1:
                self._restore(self._ckpts[k0])
                for k in range(k0, k1 - 1):
                    kept[k] = self._cycle(k, keep=True)
                    if k + 1 < k1 - 1:
                        states[k + 1] = self._save()
                # the container and phi, chi now hold the start state of cycle k1 - 1
                held = k1 - 1
            if self.mem_probe:
                self.mem_probe()
            for k in range(k1 - 1, k0 - 1, -1):
                if k != held:  # the reverse cycle leaves the particles and fields as they are
                    self._restore(states[k])
                lphi, lchi = self._reverse_cycle(k, self.phi, self.chi, lx, lq, lphi, lchi, kept.pop(k, None))
            del states, kept
        lpsi = c.real(1)
        c.adj_ic_lattice(lx, lq, self.cx, self.cq, lpsi)
        return lpsi


def pm_density_gr(psi, run):
    """final CIC mass density of the GR run `run` (PMRunGR), differentiable in psi"""
    return _PMFunction.apply(psi, run)
