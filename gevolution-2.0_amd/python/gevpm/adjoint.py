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
        p.kick_drift(0, s["dtau_k"], s["dtau"], self.phi, None, None, 1, 0, s["a"], s["a_d"], 1.0, 1.0)

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
