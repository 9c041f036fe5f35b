"""Physics pins of the oracle's whole cycle and IC generator (CPU only).

* linear growth: a single small Zel'dovich mode in Einstein-de Sitter grows
  like D(a) = a (Newton, and GR on a sub-horizon box);
* momentum conservation of the Newtonian PM cycle (CIC deposit + CIC
  interpolation of an antisymmetric gradient);
* IC power spectrum: the realised linear density field has the input P(k).
"""
import numpy as np
import pytest

C_LIGHT = 2997.92458


def lattice(N):
    g = (np.arange(N) + 0.5) / N
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()])


def run_mode(O, gr, boxsize, N=32, steps=24):
    fpg = 1.5 * boxsize ** 2 / C_LIGHT ** 2
    om, ol = 1.0, 0.0
    a_in, a_end = 0.01, 0.02
    x0 = lattice(N)
    amp = 0.05 / N
    d = amp * np.sin(2 * np.pi * x0[0])
    pos = x0.copy()
    pos[0] = (pos[0] + d) % 1.0
    H = O.hconf(a_in, fpg, om, ol)
    q = np.zeros_like(pos)
    q[0] = a_in * H * d  # growing mode, D ~ a
    M = N ** 3
    st = dict(pos=pos, q=q, phi=np.zeros(M), chi=np.zeros(M), Bi=np.zeros((3, M)), source=np.zeros(M),
              Sij=np.zeros((6, M)), a=a_in, tau=O.particle_horizon(a_in, fpg, om, ol), dtau_old=0.0)
    if gr:
        # initial phi as the IC generator sets it (ic_basic.hpp:2406-2428)
        rho = O.project_T00(N, pos, q, om / M)
        F = O.fft_r2c(rho)
        F[0, 0, 0] = 0.0
        st["phi"] = O.fft_c2r(O.poisson(F, fpg / a_in, 3.0 * (H * H + fpg * om / a_in)))
    t1 = O.particle_horizon(a_end, fpg, om, ol)
    dtau = (t1 - st["tau"]) / steps
    pmass = om / M
    for _ in range(steps):
        st["dtau"] = dtau
        O.step(N, st, gr, pmass, fpg, om, ol, 1.0 / a_in - 1.0)
    disp = (st["pos"][0] - x0[0] + 0.5) % 1.0 - 0.5
    s = np.sin(2 * np.pi * x0[0])
    A = np.sum(disp * s) / np.sum(s * s)
    return A / amp, st["a"] / a_in, st


def test_newton_linear_growth(O):
    g, ratio, st = run_mode(O, gr=0, boxsize=320.0)
    # discrete Laplacian / CIC error at k = 2 pi / L on 32^3 is ~0.3%
    assert abs(g / ratio - 1) < 0.01, (g, ratio)


def test_gr_linear_growth_subhorizon(O):
    g, ratio, st = run_mode(O, gr=1, boxsize=8.0)
    assert abs(g / ratio - 1) < 0.01, (g, ratio)


def test_newton_momentum_conservation(O):
    N = 16
    rng = np.random.default_rng(80)
    pos = rng.uniform(0, 1, (3, 5000))
    q = np.zeros_like(pos)
    M = N ** 3
    fpg = 1.5 * 100.0 ** 2 / C_LIGHT ** 2
    st = dict(pos=pos, q=q, phi=np.zeros(M), chi=np.zeros(M), Bi=np.zeros((3, M)), source=np.zeros(M),
              Sij=np.zeros((6, M)), a=0.1, tau=1.0, dtau=0.01, dtau_old=0.01)
    O.step(N, st, 0, 1.0 / 5000, fpg, 1.0, 0.0, 0.0)
    P = st["q"].sum(1)
    assert np.abs(P).max() < 1e-12 * np.abs(st["q"]).sum()


def test_ic_power_spectrum_matches_input(O):
    import gevpm.cosmology as cosmo
    N, L = 64, 320.0
    k, P = cosmo.bbks_table(z=100.0)
    fpg = 1.5 * L ** 2 / C_LIGHT ** 2
    ic = O.ic_generate(N, 42, k, P, L, fpg, 0.31, 0.69, 1 / 101.0, gr=0)
    rho = O.project_T00(N, ic["pos"], ic["q"], 1.0 / N ** 3)  # mean 1
    dk = O.fft_r2c(rho - 1.0) / N ** 3
    kx = np.arange(N // 2 + 1)
    kk = np.fft.fftfreq(N, 1.0 / N)
    n = np.sqrt(kx[None, None, :] ** 2 + kk[None, :, None] ** 2 + kk[:, None, None] ** 2)
    kphys = 2 * np.pi * n / L
    pw = L ** 3 * np.abs(dk) ** 2
    w = np.where((kx == 0)[None, None, :], 1.0, 2.0) * np.ones_like(pw)
    sel = (n >= 1) & (n < N / 8)
    ratio = np.sum((w * pw)[sel]) / np.sum((w * np.interp(kphys, k, P))[sel])
    # ~ 2 * 4 pi 8^3 / 3 modes: statistical error ~ 3%
    assert abs(ratio - 1) < 0.12, ratio
    assert ic["max_disp"] * N < 1.0  # displacements below a cell at z = 100
