"""Background cosmology helpers for the synthetic initial conditions.

The reference reads the linear power spectrum from an "mPk file" (two
columns: k in h/Mpc, P in (Mpc/h)^3, ic_basic.hpp:370-495).  Here the table
is produced from the BBKS fitting formula (Bardeen, Bond, Kaiser & Szalay
1986, eq. G3) with shape parameter Gamma = Omega_m h, normalised to sigma_8
at z = 0 with a real-space top hat of radius 8 Mpc/h, and scaled to the
initial redshift with the LCDM linear growth factor
D(a) ~ H(a) int_0^a da' / (a' H(a'))^3 (Heath 1977).
"""
import numpy as np
from scipy import integrate


def bbks_transfer(k_hmpc, omega_m, h):
    q = np.asarray(k_hmpc, np.float64) / (omega_m * h)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.log1p(2.34 * q) / (2.34 * q) * (1.0 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3
                                               + (6.71 * q) ** 4) ** -0.25
    return np.where(q > 0, t, 1.0)


def tophat_window(x):
    x = np.asarray(x, np.float64)
    small = x < 1e-4
    xs = np.where(small, 1.0, x)
    w = 3.0 * (np.sin(xs) - xs * np.cos(xs)) / xs ** 3
    return np.where(small, 1.0 - x * x / 10.0, w)


def sigma_R(k, P, R):
    """sqrt( int dk k^2 P(k) W(kR)^2 / (2 pi^2) ) on a log-spaced k grid (trapezoid in ln k)."""
    lnk = np.log(k)
    integrand = k ** 3 * P * tophat_window(k * R) ** 2 / (2.0 * np.pi ** 2)
    return np.sqrt(np.trapezoid(integrand, lnk))


def growth_factor(a, omega_m, omega_l):
    """Linear growth D(a), D(1) = 1, flat or curved LCDM without radiation."""
    def E(x):
        return np.sqrt(omega_m / x ** 3 + omega_l + (1.0 - omega_m - omega_l) / x ** 2)

    def D_unnorm(x):
        val, _ = integrate.quad(lambda y: 1.0 / (y * E(y)) ** 3, 0.0, x, epsabs=0, epsrel=1e-12, limit=200)
        return 2.5 * omega_m * E(x) * val

    return D_unnorm(a) / D_unnorm(1.0)


def bbks_table(omega_m=0.31, h=0.67, n_s=0.965, sigma8=0.81, z=100.0, omega_l=None, kmin=1e-5, kmax=200.0, n=1024):
    """mPk-style table (k [h/Mpc], P [(Mpc/h)^3]) of the BBKS spectrum at redshift z."""
    if omega_l is None:
        omega_l = 1.0 - omega_m
    k = np.logspace(np.log10(kmin), np.log10(kmax), n)
    P = k ** n_s * bbks_transfer(k, omega_m, h) ** 2
    P *= (sigma8 / sigma_R(k, P, 8.0)) ** 2
    D = growth_factor(1.0 / (1.0 + z), omega_m, omega_l)
    return k, P * D * D


def write_mpk(path, k, P):
    with open(path, "w") as f:
        for a, b in zip(k, P):
            f.write(f"{a:.10e} {b:.10e}\n")


def tau_schedule(hconf, particle_horizon, rk4bg, a_in, a_end, steps):
    """Uniform conformal-time steps between a_in and a_end: the harness's
    fixed-cycle stand-in for the reference's Courant / step-limit rule
    (main.cu:597-603, 1307-1310).  Returns (tau_in, dtau)."""
    t0, t1 = particle_horizon(a_in), particle_horizon(a_end)
    return t0, (t1 - t0) / steps
