"""Pins the oracle's CIC projections, kick/drift functors, IC helpers,
background, RNG and spline (CPU only).  Expected values come from sum rules,
first moments, closed forms for one-dimensional fields and textbook routines."""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline


def rand_particles(N, n, seed, vmax=0.05, margin=0.0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(margin, 1.0 - margin, (3, n))
    q = rng.uniform(-vmax, vmax, (3, n))
    return pos, q


def moment(field, N, axis):
    """sum_s field(s) * coordinate_axis(s) for a flat [z,y,x] field"""
    f = field.reshape(N, N, N)
    c = np.arange(N, dtype=np.float64)
    shape = [1, 1, 1]
    shape[2 - axis] = N
    return np.sum(f * c.reshape(shape))


# ------------------------------------------------------------ deposits ----
def test_mass_deposit_sum_and_first_moment(O):
    N = 16
    pos, q = rand_particles(N, 4000, 1, margin=0.1)  # keep CIC clouds away from the wrap
    m = 0.37
    rho = O.project_T00(N, pos, q, m)  # phi = None: mass density
    mass = m * N ** 3
    assert np.isclose(rho.sum(), mass * pos.shape[1], rtol=1e-13)
    for ax in range(3):
        assert np.isclose(moment(rho, N, ax), mass * np.sum(pos[ax] * N), rtol=1e-12)


def test_mass_deposit_single_particle_and_lattice(O):
    N = 8
    pos = np.array([[0.3], [0.55], [0.95]])  # z-cloud wraps
    rho = O.project_T00(N, pos, np.zeros((3, 1)), 1.0).reshape(N, N, N)
    # product of 1-D linear interpolation weights
    def w1(p):
        x = p * N
        c = int(np.floor(x)); r = x - c
        v = np.zeros(N); v[c % N] += 1 - r; v[(c + 1) % N] += r
        return v
    ref = np.einsum("k,j,i->kji", w1(0.95), w1(0.55), w1(0.3)) * N ** 3
    assert np.allclose(rho, ref, rtol=1e-13, atol=1e-12)
    # cell-centred lattice at rest -> uniform
    g = (np.arange(N) + 0.5) / N
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([x.ravel(), y.ravel(), z.ravel()])
    rho = O.project_T00(N, P, np.zeros_like(P), 1.0 / N ** 3)
    assert np.allclose(rho, 1.0, rtol=1e-13)


def test_T00_energy_sum_rule_and_phi_terms(O):
    N = 16
    pos, q = rand_particles(N, 3000, 2, vmax=0.3, margin=0.1)
    m, a = 0.5, 0.02
    zero = np.zeros(N ** 3)
    T = O.project_T00(N, pos, q, m, a=a, phi=zero)
    e = np.sqrt(np.sum(q * q, 0) + a * a)
    mass = m * N ** 3 / a
    assert np.isclose(T.sum(), mass * e.sum(), rtol=1e-13)       # relativistic energy
    for ax in range(3):
        assert np.isclose(moment(T, N, ax), mass * np.sum(e * (pos[ax] * N)), rtol=1e-11,
                          atol=1e-9 * abs(T.sum()))
    # constant phi: the first-order metric correction (3e + q^2/e) per particle
    p = 1e-6
    Tp = O.project_T00(N, pos, q, m, a=a, phi=np.full(N ** 3, p))
    slope = (Tp.sum() - T.sum()) / p
    q2 = np.sum(q * q, 0)
    assert np.isclose(slope, mass * np.sum(3 * e + q2 / e), rtol=1e-5)


def test_T0i_momentum_sum_rule_and_staggering(O):
    N = 16
    pos, q = rand_particles(N, 3000, 3, margin=0.1)
    m = 0.8
    T = O.project_T0i(N, pos, q, m)
    mass = m * N ** 3
    cell = np.floor(pos * N)
    frac = pos * N - cell
    for i in range(3):
        assert np.isclose(T[i].sum(), mass * q[i].sum(), rtol=1e-12)
        for ax in range(3):
            # component i sits on the link x + e_i/2: no CIC spread along axis i
            expect = np.sum(q[i] * (cell[ax] if ax == i else cell[ax] + frac[ax]))
            assert np.isclose(moment(T[i], N, ax), mass * expect, rtol=1e-11)
    # constant phi multiplies every deposit by (1 + 2 phi)
    p = 1e-3
    Tp = O.project_T0i(N, pos, q, m, phi=np.full(N ** 3, p))
    assert np.allclose(Tp, T * (1 + 2 * p), rtol=1e-13, atol=1e-13 * np.abs(T).max())


def test_Tij_sum_rules_and_staggering(O):
    N = 16
    pos, q = rand_particles(N, 3000, 4, vmax=0.3, margin=0.1)
    m, a = 0.6, 0.05
    T = O.project_Tij(N, pos, q, m, a=a, phi=np.zeros(N ** 3))
    e = np.sqrt(np.sum(q * q, 0) + a * a)
    mass = m * N ** 3 / a
    cell = np.floor(pos * N)
    frac = pos * N - cell
    comps = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 4, (2, 2): 5}
    for (i, j), c in comps.items():
        w = q[i] * q[j] / e
        assert np.isclose(T[c].sum(), mass * w.sum(), rtol=1e-12)
        for ax in range(3):
            if i == j:
                coord = cell[ax] + frac[ax]  # full CIC
            else:
                k = 3 - i - j               # off-diagonal: edge, spread only along k
                coord = cell[ax] + frac[ax] if ax == k else cell[ax]
            assert np.isclose(moment(T[c], N, ax), mass * np.sum(w * coord), rtol=1e-10,
                              atol=1e-10 * abs(mass * np.abs(w).sum()) * N)
    # constant phi: f = 1 + (4 + a^2/(q^2+a^2)) phi per particle
    p = 1e-4
    Tp = O.project_Tij(N, pos, q, m, a=a, phi=np.full(N ** 3, p))
    fac = 1 + (4 + a * a / e ** 2) * p
    for (i, j), c in comps.items():
        assert np.isclose(Tp[c].sum(), mass * np.sum(fac * q[i] * q[j] / e), rtol=1e-12)


# ------------------------------------------------------------- kicks ------
def field_1d(N, axis, vals):
    f = np.zeros((N, N, N))
    shape = [1, 1, 1]
    shape[2 - axis] = N
    return (f + vals.reshape(shape)).ravel()


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_newton_kick_1d_closed_form(O, axis):
    """psi(x_axis) only: dq = -dtau a N [(1-r) D(c) + r D(c+1)], D(i) = (psi(i+1)-psi(i-1))/2."""
    N = 16
    rng = np.random.default_rng(10 + axis)
    vals = rng.standard_normal(N)
    psi = field_1d(N, axis, vals)
    pos, q = rand_particles(N, 500, 20 + axis)
    q0 = q.copy()
    dtau, a = 0.013, 0.4
    O.update_q_newton(N, pos, q, dtau, psi, a)
    c = np.floor(pos[axis] * N).astype(int)
    r = pos[axis] * N - c
    D = 0.5 * (np.roll(vals, -1) - np.roll(vals, 1))
    g = (1 - r) * D[c % N] + r * D[(c + 1) % N]
    expect = q0.copy()
    expect[axis] -= dtau * a * N * g
    assert np.allclose(q, expect, rtol=1e-12, atol=1e-14)


def test_gr_kick_newton_limit_chi_and_B(O):
    N = 16
    rng = np.random.default_rng(30)
    vals = 1e-3 * rng.standard_normal(N)
    phi = field_1d(N, 0, vals)
    pos, _ = rand_particles(N, 400, 31)
    q = np.zeros((3, 400))
    a, dtau, bc = 0.3, 0.01, 0.3 * 0.3 * N
    ret = O.update_q(N, pos, q, dtau, phi, None, None, 1, a, bc)
    c = np.floor(pos[0] * N).astype(int); r = pos[0] * N - c
    D = 0.5 * (np.roll(vals, -1) - np.roll(vals, 1))
    g = (1 - r) * D[c % N] + r * D[(c + 1) % N]
    interp = (1 - r) * vals[c % N] + r * vals[(c + 1) % N]
    expect = -dtau * a * N * (1 + interp) * g  # q = 0: phiint = 1 + phi
    assert np.allclose(q[0], expect, rtol=1e-12, atol=1e-18)
    assert np.allclose(q[1:], 0)
    assert np.isclose(ret, np.max(np.sum(q * q, 0)) / a / a, rtol=1e-14)
    # chi(x) only: dq_x = + dtau e N (chi(c+1) - chi(c)), independent of r
    chv = rng.standard_normal(N)
    chi = field_1d(N, 0, chv)
    q = np.zeros((3, 400))
    O.update_q(N, pos, q, dtau, np.zeros(N ** 3), chi, None, 2, a, bc)
    assert np.allclose(q[0], dtau * a * N * (chv[(c + 1) % N] - chv[c % N]), rtol=1e-12)
    # B_1(x) only: dq_0 = -dtau N v1 (B1(c+1) - B1(c)) / bcoeff, others unchanged
    bv = rng.standard_normal(N)
    B = np.zeros((3, N ** 3)); B[1] = field_1d(N, 0, bv)
    q = rng.uniform(-0.1, 0.1, (3, 400)); q0 = q.copy()
    O.update_q(N, pos, q, dtau, np.zeros(N ** 3), np.zeros(N ** 3), B, 3, a, bc)
    assert np.allclose(q[0], q0[0] - dtau * N * q0[1] * (bv[(c + 1) % N] - bv[c % N]) / bc, rtol=1e-11)
    assert np.allclose(q[1:], q0[1:], rtol=1e-14)
    # uniform B: no force
    q = q0.copy()
    O.update_q(N, pos, q, dtau, np.zeros(N ** 3), np.zeros(N ** 3), np.ones((3, N ** 3)), 3, a, bc)
    assert np.allclose(q, q0, rtol=1e-13, atol=1e-15)


def test_drift_closed_forms(O):
    N = 16
    pos, q = rand_particles(N, 300, 40, vmax=0.4, margin=0.2)
    a, dtau, bc = 0.5, 0.01, 0.25 * N
    e = np.sqrt(np.sum(q * q, 0) + a * a)
    p1 = pos.copy()
    O.update_pos(N, p1, q, dtau, None, None, None, 0, a, bc)
    assert np.allclose(p1, pos + dtau * q / e, rtol=1e-14)
    ph = 1e-3
    p2 = pos.copy()
    O.update_pos(N, p2, q, dtau, np.full(N ** 3, ph), np.zeros(N ** 3), None, 2, a, bc)
    fac = (1 + (3 - np.sum(q * q, 0) / e ** 2) * ph) / e
    assert np.allclose(p2, pos + dtau * q * fac, rtol=1e-14)
    p3 = pos.copy()
    Bc = np.array([0.1, -0.2, 0.3])
    O.update_pos(N, p3, q, dtau, np.zeros(N ** 3), np.zeros(N ** 3), np.repeat(Bc[:, None], N ** 3, 1), 3, a, bc)
    assert np.allclose(p3, pos + dtau * (q / e + Bc[:, None] / bc), rtol=1e-13)
    p4 = pos.copy()
    O.update_pos_newton(N, p4, q, dtau, a)
    assert np.allclose(p4, pos + dtau * q / a, rtol=1e-14)
    # periodic wrap
    p5 = np.array([[0.999], [0.0001], [0.5]])
    O.update_pos_newton(N, p5, np.array([[0.02], [-0.05], [0.0]]), 1.0, 1.0)
    assert np.allclose(p5[:, 0], [0.019, 0.9501, 0.5], atol=1e-14)


def test_ic_displace_and_velocity_closed_forms(O):
    N = 16
    vals = np.random.default_rng(50).standard_normal(N)
    xi = field_1d(N, 1, vals)
    pos, q = rand_particles(N, 300, 51, margin=0.2)
    p0 = pos.copy()
    dmax = O.displace_particles(N, pos, xi, 0.001)
    c = np.floor(p0[1] * N).astype(int)
    d = 0.001 * N * (vals[(c + 1) % N] - vals[c % N])
    assert np.allclose(pos[1], p0[1] + d, rtol=1e-13)
    assert np.allclose(pos[[0, 2]], p0[[0, 2]])
    assert np.isclose(dmax, np.abs(d).max(), rtol=1e-13)
    qq = np.zeros((3, 300))
    ret = O.initialize_q(N, p0, qq, xi, 2.0)
    assert np.allclose(qq[1], -2.0 * N * (vals[(c + 1) % N] - vals[c % N]), rtol=1e-13)
    assert np.isclose(ret, np.max(qq[1] ** 2), rtol=1e-13)


# ---------------------------------------------------- sources (closed) ----
def test_prepare_source_T00_closed_forms(O):
    N = 16
    rng = np.random.default_rng(60)
    s = rng.standard_normal(N ** 3)
    c1, c2, c3, bg = 0.3, 1.7, 0.05, 0.31
    zero = np.zeros(N ** 3)
    r, tot = O.prepare_source_T00(zero, zero, s, bg, c1, c2, c3)
    assert np.allclose(r, c2 * (s - bg), rtol=1e-14)
    assert np.isclose(tot, s.sum(), rtol=1e-13)
    p, ch = 2e-3, -1e-3
    r, _ = O.prepare_source_T00(np.full(N ** 3, p), np.full(N ** 3, ch), s, bg, c1, c2, c3)
    ref = c2 * (s - bg) * (1 - 2 * p * (1 - p)) + (c3 * (1 - 3 * p) - c1) * p - c3 * ch
    assert np.allclose(r, ref, rtol=1e-13)
    # phi = eps cos(2 pi x/N): gradient term 0.125 (phi(x-1)-phi(x+1))^2 = 0.5 eps^2 sin^2(th) sin^2(2 pi/N)
    eps = 1e-2
    th = 2 * np.pi * np.arange(N) / N
    phi = field_1d(N, 0, eps * np.cos(th))
    r, _ = O.prepare_source_T00(phi, zero, np.full(N ** 3, bg), bg, 0.0, c2, 0.0)
    grad = field_1d(N, 0, 0.5 * eps ** 2 * np.sin(th) ** 2 * np.sin(2 * np.pi / N) ** 2)
    assert np.allclose(r, grad, rtol=1e-12, atol=1e-18)


def test_prepare_source_Tij_closed_forms(O):
    N = 8
    rng = np.random.default_rng(61)
    T = rng.standard_normal((6, N ** 3))
    S = O.prepare_source_Tij(np.zeros(N ** 3), T, 0.7)
    assert np.allclose(S, 0.7 * T, rtol=1e-14)
    u = 1e-2 * rng.standard_normal(N)
    v = 1e-2 * rng.standard_normal(N)
    phi = field_1d(N, 0, u) + field_1d(N, 1, v)
    S = O.prepare_source_Tij(phi, np.zeros((6, N ** 3)), 0.7)
    du = np.roll(u, -1) - u
    dv = np.roll(v, -1) - v
    # S01 = 2 (u(x+1)-u(x)) (v(y+1)-v(y)) for phi = u(x) + v(y)
    assert np.allclose(S[1], 2 * np.outer(dv, du)[None].repeat(N, 0).ravel(), rtol=1e-12, atol=1e-18)
    assert np.allclose(S[2], 0, atol=1e-18) and np.allclose(S[4], 0, atol=1e-18)
    cu = np.roll(u, -1) - np.roll(u, 1)
    assert np.allclose(S[0], field_1d(N, 0, 0.5 * cu ** 2), rtol=1e-12)


# --------------------------------------------------------- background -----
def test_background_eds_closed_forms(O):
    fpg = 0.7
    C = 2 * fpg / 3
    for a in (0.01, 0.1, 0.5, 1.0):
        assert np.isclose(O.hconf(a, fpg, 1.0, 0.0), np.sqrt(C / a), rtol=1e-15)
        # the integral starts at sqrt(a) 1e-7 (background.hpp:286): tau = 2 (1 - 1e-7) sqrt(a/C)
        assert np.isclose(O.particle_horizon(a, fpg, 1.0, 0.0), 2 * (1 - 1e-7) * np.sqrt(a / C), rtol=1e-13)
    # RK4 converges at 4th order to a(tau) = C tau^2/4
    t0, t1 = 2 * np.sqrt(0.01 / C), 2 * np.sqrt(0.5 / C)
    errs = []
    for n in (20, 40, 80):
        a = 0.01
        for _ in range(n):
            a = O.rk4bg(a, fpg, 1.0, 0.0, (t1 - t0) / n)
        errs.append(abs(a - 0.5))
    assert 12 < errs[0] / errs[1] < 20 and 12 < errs[1] / errs[2] < 20


# ------------------------------------------------------- RNG / spline -----
def test_philox_known_answers(O):
    """Random123 philox4x32-10 known-answer vectors (kat_vectors)."""
    kat = [([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
           ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
           ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
            [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])]
    for ctr, key, out in kat:
        assert list(O.philox4x32_10(ctr, key)) == out


def test_natural_spline_matches_scipy(O):
    rng = np.random.default_rng(70)
    x = np.sort(rng.uniform(0, 10, 25))
    y = np.sin(x) + 0.1 * x
    cs = CubicSpline(x, y, bc_type="natural")
    for xv in np.linspace(x[0], x[-1], 57):
        assert np.isclose(O.spline_eval(x, y, xv), cs(xv), rtol=1e-12, atol=1e-13)
