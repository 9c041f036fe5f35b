"""Pins the oracle terms that sum rules and one-dimensional fields leave open
(CPU only).

Every expectation here is built from one-dimensional lattice operators whose
meaning is fixed by where a quantity lives on the staggered lattice, not from
the oracle's expanded corner-by-corner expressions:

  lin(f, c, r)     linear interpolation between sites c and c+1
  ngp(f, c)        the value at site c (a staggered value that covers the cell)
  fwd(f, c)        f(c+1) - f(c), the difference across the link (c, c+1)
  bwdlin(f, c, r)  linear interpolation of the site differences f(i) - f(i-1)
                   between i = c and i = c+1 (the derivative of a link field
                   at the particle along its own axis)
  cen(f, c, r)     linear interpolation of the central differences
                   (f(i+1) - f(i-1))/2 between i = c and c+1 (GRADIENT_ORDER 2)

A three-dimensional stencil is the tensor product of one such operator per
axis, evaluated with numpy fancy indexing on random fields.  A swapped corner,
a weight paired with the wrong axis or a dropped term changes the result by
O(1) of the term, far above the tolerances.
"""
import numpy as np
import pytest

COMPS = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 4, (2, 2): 5}


def particles(N, n, seed, vmax=0.2):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, 1, (3, n)), rng.uniform(-vmax, vmax, (3, n))


def cell_frac(N, pos):
    c = np.floor(pos * N).astype(np.int64)
    return c, pos * N - c


# one-dimensional stencils: list of (offset, weight array) pairs
def st_lin(r):
    return [(0, 1 - r), (1, r)]


def st_ngp(r):
    return [(0, np.ones_like(r))]


def st_fwd(r):
    return [(1, np.ones_like(r)), (0, -np.ones_like(r))]


def st_bwdlin(r):
    # (1-r)(f(c) - f(c-1)) + r(f(c+1) - f(c))
    return [(-1, -(1 - r)), (0, (1 - r) - r), (1, r)]


def st_cen(r):
    # (1-r)(f(c+1) - f(c-1))/2 + r(f(c+2) - f(c))/2
    return [(-1, -0.5 * (1 - r)), (0, -0.5 * r), (1, 0.5 * (1 - r)), (2, 0.5 * r)]


def apply3(f, N, c, ops):
    """sum over the tensor product of the per-axis stencils ops[axis] of the
    [z,y,x] field f (flat or 3-D) at cells c[3, n]"""
    f = f.reshape(N, N, N)
    out = 0.0
    for ox, wx in ops[0]:
        for oy, wy in ops[1]:
            for oz, wz in ops[2]:
                out = out + wx * wy * wz * f[(c[2] + oz) % N, (c[1] + oy) % N, (c[0] + ox) % N]
    return out


def cic(f, N, c, r):
    return apply3(f, N, c, [st_lin(r[0]), st_lin(r[1]), st_lin(r[2])])


def w1(N, p):
    """one-dimensional linear CIC weights of a particle at p (box units) on N sites"""
    x = p * N
    c = int(np.floor(x)); r = x - c
    v = np.zeros(N); v[c % N] += 1 - r; v[(c + 1) % N] += r
    return v


def ngp1(N, p):
    v = np.zeros(N); v[int(np.floor(p * N)) % N] = 1.0
    return v


def outer3(vx, vy, vz):
    return np.einsum("k,j,i->kji", vz, vy, vx).ravel()


# ------------------------------------------------------------- kick: B ----
def pgradB_expected(N, B, c, r, v):
    """sum_c' v_c' D_d B_c' at the particle, B_c' living on the links x + e_c'/2:
    d != c': the link difference along d, NGP along c', linear along the third axis;
    d == c': bwdlin along d, linear along the other two axes."""
    out = np.zeros((3, c.shape[1]))
    for d in range(3):
        for cp in range(3):
            ops = [None] * 3
            if d == cp:
                for ax in range(3):
                    ops[ax] = st_bwdlin(r[ax]) if ax == d else st_lin(r[ax])
            else:
                e = 3 - d - cp
                ops[d] = st_fwd(r[d]); ops[cp] = st_ngp(r[cp]); ops[e] = st_lin(r[e])
            out[d] += v[cp] * apply3(B[cp], N, c, ops)
    return out


def test_kick_B_term_random_fields(O):
    """update_q with phi = chi = 0 and a random B: dq = -dtau N pgradB / bcoeff
    (gevolution.hpp:968-994 and 1003-1006: e cancels against the 1/e of the B term)"""
    N = 8
    rng = np.random.default_rng(101)
    pos, q = particles(N, 600, 102)
    B = rng.standard_normal((3, N ** 3))
    a, dtau, bc = 0.3, 0.011, 0.09 * N
    q1 = q.copy()
    zero = np.zeros(N ** 3)
    O.update_q(N, pos, q1, dtau, zero, zero, B, 3, a, bc)
    c, r = cell_frac(N, pos)
    expect = q - dtau * N * pgradB_expected(N, B, c, r, q) / bc
    assert np.abs(q1 - expect).max() <= 1e-12 * np.abs(q).max()


@pytest.mark.parametrize("cp,d", [(cp, d) for cp in range(3) for d in range(3)])
def test_kick_B_separable_closed_forms(O, cp, d):
    """B_c'(x) = f(x_d) g(x_e) (e the third axis when d != c', else e = c'+1):
    closed forms of the force from 1-D sums, one (component, axis) pair at a time"""
    N = 8
    rng = np.random.default_rng(110 + 3 * cp + d)
    fv, gv = rng.standard_normal(N), rng.standard_normal(N)
    e = 3 - d - cp if d != cp else (cp + 1) % 3
    idx = np.indices((N, N, N))  # [z, y, x] coordinates
    coord = [idx[2], idx[1], idx[0]]
    B = np.zeros((3, N ** 3))
    B[cp] = (fv[coord[d]] * gv[coord[e]]).ravel()
    pos, q = particles(N, 400, 120 + cp)
    a, dtau, bc = 0.5, 0.02, 0.25 * N
    q1 = q.copy()
    zero = np.zeros(N ** 3)
    O.update_q(N, pos, q1, dtau, zero, zero, B, 3, a, bc)
    c, r = cell_frac(N, pos)
    cd, rd, ce, re_ = c[d], r[d], c[e], r[e]
    F = np.zeros((3, pos.shape[1]))
    if d != cp:
        # force along d: link difference of f, linear in g along e
        F[d] += q[cp] * (fv[(cd + 1) % N] - fv[cd % N]) * ((1 - re_) * gv[ce % N] + re_ * gv[(ce + 1) % N])
        # force along e: link difference of g, f linear along d
        F[e] += q[cp] * (gv[(ce + 1) % N] - gv[ce % N]) * ((1 - rd) * fv[cd % N] + rd * fv[(cd + 1) % N])
    else:
        # own axis: interpolated site differences of f; g linear along e
        bw = (1 - rd) * (fv[cd % N] - fv[(cd - 1) % N]) + rd * (fv[(cd + 1) % N] - fv[cd % N])
        F[d] += q[cp] * bw * ((1 - re_) * gv[ce % N] + re_ * gv[(ce + 1) % N])
        # along e: link difference of g, f taken at the link that covers the cell (NGP along c')
        F[e] += q[cp] * (gv[(ce + 1) % N] - gv[ce % N]) * fv[cd % N]
    expect = q - dtau * N * F / bc
    assert np.abs(q1 - expect).max() <= 1e-12 * np.abs(q).max()


# ------------------------------------------------------ kick: gradients ----
def test_newton_kick_random_potential(O):
    """update_q_Newton: dq = -dtau a N grad2(psi) with the CIC-weighted central
    differences, plus +dtau a N grad1(chi) (links, linear across) when nfield = 2"""
    N = 8
    rng = np.random.default_rng(130)
    psi, chi = rng.standard_normal(N ** 3), rng.standard_normal(N ** 3)
    pos, q = particles(N, 500, 131)
    a, dtau = 0.7, 0.013
    c, r = cell_frac(N, pos)
    g2 = np.stack([apply3(psi, N, c, [st_cen(r[ax]) if ax == d else st_lin(r[ax]) for ax in range(3)])
                   for d in range(3)])
    g1 = np.stack([apply3(chi, N, c, [st_fwd(r[ax]) if ax == d else st_lin(r[ax]) for ax in range(3)])
                   for d in range(3)])
    q1 = q.copy()
    ret = O.update_q_newton(N, pos, q1, dtau, psi, a)
    assert np.abs(q1 - (q - dtau * a * N * g2)).max() <= 1e-12 * np.abs(q).max()
    assert np.isclose(ret, np.max(np.sum(q1 * q1, 0)) / a / a, rtol=1e-14)
    q2 = q.copy()
    O.update_q_newton(N, pos, q2, dtau, psi, a, chi=chi, nfield=2)
    assert np.abs(q2 - (q - dtau * a * N * (g2 - g1))).max() <= 1e-12 * np.abs(q).max()


def test_gr_kick_first_order_geodesic(O):
    """GR kick with small random phi, chi: to first order in the potentials,
    dq/dtau = -(q^2 + eps^2)/eps grad phi + eps grad chi (eps = sqrt(q^2 + a^2);
    the weak-field geodesic equation of Adamek et al. 2016, eq. 2.11 with
    Psi = phi - chi), with the lattice gradients above.  The O(phi^2) part is
    compared separately through its scaling: it changes by 4x when phi doubles."""
    N = 8
    rng = np.random.default_rng(140)
    phi0, chi0 = rng.standard_normal(N ** 3), rng.standard_normal(N ** 3)
    pos, q = particles(N, 500, 141, vmax=0.5)
    a, dtau = 0.3, 0.01
    c, r = cell_frac(N, pos)
    q2 = np.sum(q * q, 0)
    eps = np.sqrt(q2 + a * a)
    g2 = np.stack([apply3(phi0, N, c, [st_cen(r[ax]) if ax == d else st_lin(r[ax]) for ax in range(3)])
                   for d in range(3)])
    g1 = np.stack([apply3(chi0, N, c, [st_fwd(r[ax]) if ax == d else st_lin(r[ax]) for ax in range(3)])
                   for d in range(3)])

    def run(s):
        qq = q.copy()
        O.update_q(N, pos, qq, dtau, s * phi0, s * chi0, None, 2, a, a * a * N)
        return (qq - q) / (dtau * N)

    s = 1e-4
    lin = -(q2 + eps * eps) / eps * g2 + eps * g1
    d1, d2 = run(s), run(2 * s)
    # second-order remainders: d(s) = s lin + s^2 quad  ->  (4 d(s) - d(2s)) / 2 = s lin
    first = (4 * d1 - d2) / 2
    assert np.abs(first - s * lin).max() <= 1e-9 * s * np.abs(lin).max()
    quad = (d2 - 2 * d1) / 2
    quad2 = (run(4 * s) - 2 * d2) / 2
    assert np.abs(quad2 - 4 * quad).max() <= 1e-3 * np.abs(quad2).max()
    # the exact q = 0 form: dq = -dtau a N (1 + CIC(phi)) grad2(phi) + dtau a N grad1(chi)
    z = np.zeros_like(q)
    O.update_q(N, pos, z, dtau, 1e-2 * phi0, 1e-2 * chi0, None, 2, a, a * a * N)
    phint = cic(1e-2 * phi0, N, c, r)
    expect = -dtau * a * N * ((1 + phint) * 1e-2 * g2 - 1e-2 * g1)
    assert np.abs(z - expect).max() <= 1e-12 * np.abs(expect).max()


# ------------------------------------------------------------- drift -------
def test_drift_B_random_fields(O):
    """update_pos: b_c = B_c at the particle, NGP along c (the link covering the
    cell) and linear along the other two axes; pos += dtau (v + b / bcoeff)"""
    N = 8
    rng = np.random.default_rng(150)
    B = rng.standard_normal((3, N ** 3))
    pos, q = particles(N, 500, 151)
    a, dtau, bc = 0.4, 1e-3, 0.16 * N
    c, r = cell_frac(N, pos)
    b = np.stack([apply3(B[cc], N, c, [st_ngp(r[ax]) if ax == cc else st_lin(r[ax]) for ax in range(3)])
                  for cc in range(3)])
    zero = np.zeros(N ** 3)
    p1 = pos.copy()
    O.update_pos(N, p1, q, dtau, zero, zero, B, 3, a, bc)
    e = np.sqrt(np.sum(q * q, 0) + a * a)
    expect = (pos + dtau * (q / e + b / bc)) % 1.0
    d = (p1 - expect + 0.5) % 1.0 - 0.5
    assert np.abs(d).max() <= 1e-14


@pytest.mark.parametrize("cc", [0, 1, 2])
def test_drift_B_separable_closed_form(O, cc):
    """B_c = f(x_d) g(x_e), d, e the other two axes: b_c = lin_d(f) lin_e(g);
    B_c = h(x_c): b_c = h(cell_c)"""
    N = 8
    rng = np.random.default_rng(160 + cc)
    fv, gv, hv = rng.standard_normal(N), rng.standard_normal(N), rng.standard_normal(N)
    d, e = (cc + 1) % 3, (cc + 2) % 3
    idx = np.indices((N, N, N))
    coord = [idx[2], idx[1], idx[0]]
    pos, q = particles(N, 300, 170 + cc)
    a, dtau, bc = 0.4, 1e-3, 0.16 * N
    ev = np.sqrt(np.sum(q * q, 0) + a * a)
    c, r = cell_frac(N, pos)
    zero = np.zeros(N ** 3)
    for field, expect_b in (
            ((fv[coord[d]] * gv[coord[e]]).ravel(),
             ((1 - r[d]) * fv[c[d] % N] + r[d] * fv[(c[d] + 1) % N]) *
             ((1 - r[e]) * gv[c[e] % N] + r[e] * gv[(c[e] + 1) % N])),
            (hv[coord[cc]].ravel(), hv[c[cc] % N])):
        B = np.zeros((3, N ** 3)); B[cc] = field
        p1 = pos.copy()
        O.update_pos(N, p1, q, dtau, zero, zero, B, 3, a, bc)
        expect = pos + dtau * q / ev
        expect[cc] += dtau * expect_b / bc
        dd = (p1 - expect % 1.0 + 0.5) % 1.0 - 0.5
        assert np.abs(dd).max() <= 1e-14


def test_drift_metric_factor_random_fields(O):
    """update_pos with random phi, chi: v = q (1 + (3 - q^2/e^2) CIC(phi) - CIC(chi)) / e"""
    N = 8
    rng = np.random.default_rng(180)
    phi, chi = 1e-2 * rng.standard_normal(N ** 3), 1e-2 * rng.standard_normal(N ** 3)
    pos, q = particles(N, 400, 181, vmax=0.5)
    a, dtau = 0.3, 1e-3
    c, r = cell_frac(N, pos)
    q2 = np.sum(q * q, 0)
    e2 = q2 + a * a
    v = q * (1 + (3 - q2 / e2) * cic(phi, N, c, r) - cic(chi, N, c, r)) / np.sqrt(e2)
    p1 = pos.copy()
    O.update_pos(N, p1, q, dtau, phi, chi, None, 2, a, a * a * N)
    dd = (p1 - (pos + dtau * v) % 1.0 + 0.5) % 1.0 - 0.5
    assert np.abs(dd).max() <= 1e-14


# ----------------------------------------------------------- deposits ------
def test_T0i_single_particle_links(O):
    """one particle: component i sits on the links (s, s + e_i) with s_i = cell_i
    (no spread along i) and the linear CIC weights across; with phi, each link
    value is multiplied by 1 + phi(s) + phi(s + e_i)"""
    N = 8
    rng = np.random.default_rng(190)
    phi = 0.05 * rng.standard_normal(N ** 3)
    m = 0.3
    for k in range(6):
        p = rng.uniform(0, 1, 3)
        if k == 0:
            p = np.array([0.99, 0.001, 0.93])  # links across the periodic wrap
        v = rng.uniform(-0.3, 0.3, 3)
        P, Q = p[:, None], v[:, None]
        T = O.project_T0i(N, P, Q, m)
        Tp = O.project_T0i(N, P, Q, m, phi=phi)
        ph3 = phi.reshape(N, N, N)
        for i in range(3):
            vecs = [ngp1(N, p[ax]) if ax == i else w1(N, p[ax]) for ax in range(3)]
            ref = m * N ** 3 * v[i] * outer3(*vecs)
            assert np.allclose(T[i], ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
            fac = 1 + ph3 + np.roll(ph3, -1, axis=2 - i)
            assert np.allclose(Tp[i], ref * fac.ravel(), rtol=1e-13, atol=1e-13 * np.abs(ref).max())


def test_T0i_linear_phi_sum_rule(O):
    """phi = alpha + sum_d beta_d x_d (away from the wrap): the links of component
    i sum to m v_i (1 + 2 phi at the link midpoints), i.e. phi evaluated at
    x_i = cell_i + 1/2 and at the particle across"""
    N = 16
    rng = np.random.default_rng(200)
    pos = rng.uniform(2.0 / N, 1 - 3.0 / N, (3, 500))
    q = rng.uniform(-0.3, 0.3, (3, 500))
    al, be = 0.01, np.array([3e-3, -5e-3, 7e-3])
    idx = np.indices((N, N, N))
    phi = (al + be[0] * idx[2] + be[1] * idx[1] + be[2] * idx[0]).ravel()
    m = 0.4
    c, r = cell_frac(N, pos)
    for j in range(0, 500, 100):
        T = O.project_T0i(N, pos[:, j:j + 1], q[:, j:j + 1], m, phi=phi)
        for i in range(3):
            xm = [c[ax, j] + 0.5 if ax == i else c[ax, j] + r[ax, j] for ax in range(3)]
            expect = m * N ** 3 * q[i, j] * (1 + 2 * (al + np.dot(be, xm)))
            assert np.isclose(T[i].sum(), expect, rtol=1e-13)


def test_Tij_single_particle(O):
    """one particle: diagonal S_ii by full CIC; off-diagonal ij on the edge
    (cell_i + 1/2, cell_j + 1/2), NGP along i and j, linear along the third axis;
    with phi all six are multiplied by f = 1 + (4 + a^2/e^2) CIC(phi)"""
    N = 8
    rng = np.random.default_rng(210)
    phi = 0.05 * rng.standard_normal(N ** 3)
    m, a = 0.7, 0.2
    for k in range(6):
        p = rng.uniform(0, 1, 3)
        if k == 0:
            p = np.array([0.995, 0.5, 0.002])
        v = rng.uniform(-0.4, 0.4, 3)
        P, Q = p[:, None], v[:, None]
        T = O.project_Tij(N, P, Q, m, a=a)
        Tp = O.project_Tij(N, P, Q, m, a=a, phi=phi)
        e = np.sqrt(v @ v + a * a)
        mass = m * N ** 3 / a
        cicw = outer3(w1(N, p[0]), w1(N, p[1]), w1(N, p[2]))
        f = 1 + (4 + a * a / e ** 2) * np.dot(cicw, phi)
        for (i, j), cmp in COMPS.items():
            if i == j:
                shape = cicw
            else:
                vecs = [ngp1(N, p[ax]) if ax in (i, j) else w1(N, p[ax]) for ax in range(3)]
                shape = outer3(*vecs)
            ref = mass * v[i] * v[j] / e * shape
            tol = 1e-13 * np.abs(ref).max() + 1e-300
            assert np.allclose(T[cmp], ref, rtol=1e-13, atol=tol)
            assert np.allclose(Tp[cmp], f * ref, rtol=1e-13, atol=tol)


def test_Tij_linear_phi_sum_rule(O):
    """phi linear in x, y, z (away from the wrap): CIC(phi) = phi(x_p) exactly, so
    sum T_ij = sum_p m' (1 + (4 + a^2/e^2) phi(x_p)) q_i q_j / e"""
    N = 16
    rng = np.random.default_rng(220)
    pos = rng.uniform(2.0 / N, 1 - 3.0 / N, (3, 2000))
    q = rng.uniform(-0.3, 0.3, (3, 2000))
    al, be = -0.02, np.array([4e-3, 2e-3, -6e-3])
    idx = np.indices((N, N, N))
    phi = (al + be[0] * idx[2] + be[1] * idx[1] + be[2] * idx[0]).ravel()
    m, a = 0.5, 0.1
    T = O.project_Tij(N, pos, q, m, a=a, phi=phi)
    e = np.sqrt(np.sum(q * q, 0) + a * a)
    php = al + be @ (pos * N)
    f = 1 + (4 + a * a / e ** 2) * php
    for (i, j), cmp in COMPS.items():
        assert np.isclose(T[cmp].sum(), m * N ** 3 / a * np.sum(f * q[i] * q[j] / e), rtol=1e-12)


def test_T00_single_particle_corners(O):
    """one particle: corner s gets w_s (e + f phi_s + f2 phi_s^2) m' with its own
    phi_s; f = 3e + q^2/e, f2 = (f^2/2 + q^2/(1 + q^2/a^2))/e"""
    N = 8
    rng = np.random.default_rng(230)
    phi = 0.1 * rng.standard_normal(N ** 3)
    m, a = 0.9, 0.3
    for k in range(5):
        p = rng.uniform(0, 1, 3)
        if k == 0:
            p = np.array([0.001, 0.996, 0.5])
        v = rng.uniform(-0.5, 0.5, 3)
        T = O.project_T00(N, p[:, None], v[:, None], m, a=a, phi=phi)
        q2 = v @ v
        e = np.sqrt(q2 + a * a)
        f = 3 * e + q2 / e
        f2 = (0.5 * f * f + q2 / (1 + q2 / a / a)) / e
        w = outer3(w1(N, p[0]), w1(N, p[1]), w1(N, p[2]))
        ref = w * (e + f * phi + f2 * phi * phi) * m * N ** 3 / a
        assert np.allclose(T, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


def test_T00_linear_phi_quadratic_term(O):
    """phi linear along the axes: sum_s w_s phi_s = phi(x_p) and
    sum_s w_s phi_s^2 = phi(x_p)^2 + sum_d beta_d^2 r_d (1 - r_d) (the variance of
    the independent two-point weights); the total of T00 follows in closed form.
    The f2 phi^2 term is also isolated by the second difference in a constant phi."""
    N = 16
    rng = np.random.default_rng(240)
    pos = rng.uniform(2.0 / N, 1 - 3.0 / N, (3, 2000))
    q = rng.uniform(-0.5, 0.5, (3, 2000))
    al, be = 0.03, np.array([0.02, -0.015, 0.01])
    idx = np.indices((N, N, N))
    phi = (al + be[0] * idx[2] + be[1] * idx[1] + be[2] * idx[0]).ravel()
    m, a = 0.5, 0.2
    T = O.project_T00(N, pos, q, m, a=a, phi=phi)
    q2 = np.sum(q * q, 0)
    e = np.sqrt(q2 + a * a)
    f = 3 * e + q2 / e
    f2 = (0.5 * f * f + q2 / (1 + q2 / a / a)) / e
    c, r = cell_frac(N, pos)
    php = al + be @ (pos * N)
    var = np.sum(be[:, None] ** 2 * r * (1 - r), 0)
    mass = m * N ** 3 / a
    assert np.isclose(T.sum(), mass * np.sum(e + f * php + f2 * (php ** 2 + var)), rtol=1e-12)
    # constant phi: (T(p) - 2 T(0) + T(-p)) / p^2 = 2 sum m' f2
    p = 0.05
    Tp = O.project_T00(N, pos, q, m, a=a, phi=np.full(N ** 3, p)).sum()
    T0 = O.project_T00(N, pos, q, m, a=a, phi=np.zeros(N ** 3)).sum()
    Tm = O.project_T00(N, pos, q, m, a=a, phi=np.full(N ** 3, -p)).sum()
    assert np.isclose((Tp - 2 * T0 + Tm) / p ** 2, 2 * mass * f2.sum(), rtol=1e-9)


# ------------------------------------------------------------ sources ------
def axis_field(N, axis, vals):
    idx = np.indices((N, N, N))
    return vals[idx[2 - axis]].ravel()


@pytest.mark.parametrize("i,j", [(0, 1), (0, 2), (1, 2)])
def test_prepare_source_Tij_all_pairs(O, i, j):
    """phi = u(x_i) + v(x_j): S_ij = 2 (u(c+1) - u(c)) (v(c+1) - v(c)) on the edge,
    the other off-diagonal components vanish, S_ii = (u(x+1) - u(x-1))^2 / 2,
    and S_kk (k the third axis) = 0"""
    N = 8
    rng = np.random.default_rng(250 + 3 * i + j)
    u, v = 1e-2 * rng.standard_normal(N), 1e-2 * rng.standard_normal(N)
    phi = axis_field(N, i, u) + axis_field(N, j, v)
    S = O.prepare_source_Tij(phi, np.zeros((6, N ** 3)), 0.7)
    k = 3 - i - j
    du = axis_field(N, i, np.roll(u, -1) - u)
    dv = axis_field(N, j, np.roll(v, -1) - v)
    assert np.allclose(S[COMPS[(i, j)]], 2 * du * dv, rtol=1e-12, atol=1e-18)
    for other in ((min(i, k), max(i, k)), (min(j, k), max(j, k))):
        assert np.abs(S[COMPS[other]]).max() <= 1e-18
    assert np.allclose(S[COMPS[(i, i)]], axis_field(N, i, 0.5 * (np.roll(u, -1) - np.roll(u, 1)) ** 2), rtol=1e-12)
    assert np.allclose(S[COMPS[(j, j)]], axis_field(N, j, 0.5 * (np.roll(v, -1) - np.roll(v, 1)) ** 2), rtol=1e-12)
    assert np.abs(S[COMPS[(k, k)]]).max() <= 1e-18


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_prepare_source_T00_gradient_term_each_axis(O, axis):
    """phi = eps cos(2 pi x_axis / N): the source gains (phi(x-1) - phi(x+1))^2 / 8"""
    N = 16
    eps = 1e-2
    th = 2 * np.pi * np.arange(N) / N
    phi = axis_field(N, axis, eps * np.cos(th))
    bg = 0.31
    r, _ = O.prepare_source_T00(phi, np.zeros(N ** 3), np.full(N ** 3, bg), bg, 0.0, 1.3, 0.0)
    grad = axis_field(N, axis, 0.5 * eps ** 2 * np.sin(th) ** 2 * np.sin(2 * np.pi / N) ** 2)
    assert np.allclose(r, grad, rtol=1e-12, atol=1e-18)


# ------------------------------------------------------------ k-space ------
def _roll(f, axis, s):
    return np.roll(f, s, 2 - axis)


def lap_idx(f):
    """7-point Laplacian in lattice units of a [z,y,x] array"""
    return sum(np.roll(f, 1, ax) + np.roll(f, -1, ax) for ax in range(3)) - 6 * f


def staggered_divergence(S):
    """V_i on the link x + e_i/2: forward difference of S_ii along i, backward
    differences of the edge components S_ij along j (lattice units)"""
    N = S.shape[1]
    V = np.zeros((3, N, N, N))
    for i in range(3):
        V[i] = _roll(S[COMPS[(i, i)]], i, -1) - S[COMPS[(i, i)]]
        for j in range(3):
            if j != i:
                Sij = S[COMPS[(min(i, j), max(i, j))]]
                V[i] += Sij - _roll(Sij, j, 1)
    return V


def transverse_part(V):
    """discrete Helmholtz projection of a link field: V - grad+ Lap^-1 div- V,
    in Fourier space with the exact lattice multipliers (e^{i th} - 1) and
    (1 - e^{-i th}); the k = 0 mode is dropped"""
    N = V.shape[1]
    th = 2 * np.pi * np.fft.fftfreq(N)
    thx = 2 * np.pi * np.arange(N // 2 + 1) / N
    TH = [thx[None, None, :], th[None, :, None], th[:, None, None]]
    fwd = [np.exp(1j * t) - 1 for t in TH]
    bwd = [1 - np.exp(-1j * t) for t in TH]
    VF = np.stack([np.fft.rfftn(V[i]) for i in range(3)])
    k2 = sum(fwd[i] * bwd[i] for i in range(3))  # = -sum 4 sin^2(th/2), real
    k2 = np.where(k2 == 0, 1.0, k2)
    div = sum(bwd[i] * VF[i] for i in range(3)) / k2
    out = np.stack([VF[i] - fwd[i] * div for i in range(3)])
    out[:, 0, 0, 0] = 0.0
    return np.stack([np.fft.irfftn(out[i], s=(N, N, N), axes=(0, 1, 2)) for i in range(3)])


def test_evolve_ft_vector_amplitude(O):
    """evolveFTvector (gevolution.hpp:359-429) on a random tensor: the increment dB
    satisfies -Lap(dB) = -(2 a^2 dtau / N) T(div S) in lattice units, where div S
    is the staggered divergence onto B's links and T the transverse projection,
    with the unnormalised transforms of the reference (c2r = N^3 irfftn)"""
    N = 8
    rng = np.random.default_rng(260)
    S = rng.standard_normal((6, N, N, N))
    SF = np.stack([np.fft.rfftn(S[c]) for c in range(6)])
    a2dtau = 0.37
    inc = O.evolve_ft_vector(SF, np.zeros((3, N, N, N // 2 + 1), np.complex128), a2dtau)
    dB = np.stack([np.fft.irfftn(inc[i], s=(N, N, N), axes=(0, 1, 2)) for i in range(3)])
    lhs = -np.stack([lap_idx(dB[i]) for i in range(3)])
    rhs = -(2 * a2dtau / N) * transverse_part(staggered_divergence(S))
    assert np.abs(lhs - rhs).max() <= 1e-12 * np.abs(rhs).max()
    # and the increment adds to B
    B0 = rng.standard_normal((3, N, N, N // 2 + 1)) + 1j * rng.standard_normal((3, N, N, N // 2 + 1))
    B1 = O.evolve_ft_vector(SF, B0, a2dtau)
    mask = np.ones(B0.shape, bool); mask[:, 0, 0, 0] = False
    assert np.allclose((B1 - B0)[mask], inc[mask], rtol=1e-13, atol=1e-13 * np.abs(inc).max())
    assert np.all(B1[:, 0, 0, 0] == 0)


def test_project_ft_scalar_random_tensor(O):
    """projectFTscalar (gevolution.hpp:251-340) on a random tensor:
    chi_hat = (k^2 tr S - 3 DD S) / (2 k^4 N) with DD S the staggered double
    divergence; in real space (lattice units, chi = irfftn of chi_hat)
    Lap^2 chi = (3 DD S - Lap tr S) / (2 N^3), which also fixes S's vector and
    tensor parts (a longitudinal test alone leaves four of the six components
    per mode open)"""
    N = 8
    rng = np.random.default_rng(270)
    S = rng.standard_normal((6, N, N, N))
    SF = np.stack([np.fft.rfftn(S[c]) for c in range(6)])
    chiF = O.project_ft_scalar(SF)
    chi = np.fft.irfftn(chiF, s=(N, N, N), axes=(0, 1, 2))
    tr = S[0] + S[3] + S[5]
    dd = 0.0
    for i in range(3):
        dd = dd + _roll(S[COMPS[(i, i)]], i, -1) - 2 * S[COMPS[(i, i)]] + _roll(S[COMPS[(i, i)]], i, 1)
    for (i, j), cmp in COMPS.items():
        if i != j:
            Sij = S[cmp]
            # backward differences along i and j bring the edge value back to the site; twice (ij and ji)
            d = Sij - _roll(Sij, i, 1)
            d = d - _roll(d, j, 1)
            dd = dd + 2 * d
    lhs = lap_idx(lap_idx(chi))
    rhs = (3 * dd - lap_idx(tr)) / (2 * N ** 3)
    rhs = rhs - rhs.mean()
    assert np.abs(lhs - rhs).max() <= 1e-11 * np.abs(rhs).max()


# --------------------------------------- metric factors from the geodesic ----
# The weak-field metric ds^2 = a^2 [-e^{2 psi} dtau^2 + e^{-2 phi} dx^2] with
# psi = phi - chi (gevolution's phi, chi) fixes, for a particle of unit mass
# with covariant spatial momentum q (u_i = q_i):
#   dq_i/dtau = -sqrt(A/Bm) sqrt(Bm + q^2) [d_i psi + q^2/(Bm + q^2) d_i phi]
#   dx^i/dtau = q_i sqrt(A / (Bm (Bm + q^2)))
#   a^3 (-T^0_0) = e^{3 phi} sqrt(a^2 + q^2 e^{2 phi}) / a        per unit density
#   a^3 T^i_j    = q_i q_j e^{5 phi} / (a sqrt(a^2 + q^2 e^{2 phi}))
# with A = a^2 e^{2 psi}, Bm = a^2 e^{-2 phi} (the Euler-Lagrange equations of
# L = -sqrt(A - Bm xdot^2) and the point-particle stress tensor
# T^mu_nu = u^mu u_nu delta^3 / (u^0 sqrt(-g))).  The reference keeps the kick to
# O(phi grad phi) and T00 to O(phi^2) (PHINONLINEAR), the drift and Tij to O(phi):
# the remainders must scale with the next power of the potential.
def _order(err1, err2):
    """log2 of err(s) / err(2 s) ~ -(order of the remainder)"""
    return np.log2(np.abs(err2).max() / np.abs(err1).max())


def test_gr_kick_exact_geodesic_to_second_order(O):
    N = 8
    rng = np.random.default_rng(280)
    phi0 = rng.standard_normal(N ** 3)
    pos, q = particles(N, 400, 281, vmax=0.6)
    a, dtau = 0.4, 0.01
    c, r = cell_frac(N, pos)
    q2 = np.sum(q * q, 0)
    g2 = np.stack([apply3(phi0, N, c, [st_cen(r[ax]) if ax == d else st_lin(r[ax]) for ax in range(3)])
                   for d in range(3)])
    ph0 = cic(phi0, N, c, r)
    zero = np.zeros(N ** 3)

    def err(s):
        qq = q.copy()
        O.update_q(N, pos, qq, dtau, s * phi0, zero, None, 2, a, a * a * N)
        ph = s * ph0
        A, Bm = a * a * np.exp(2 * ph), a * a * np.exp(-2 * ph)
        F = -np.sqrt(A / Bm) * np.sqrt(Bm + q2) * (1 + q2 / (Bm + q2)) * s * g2
        return (qq - q) / (dtau * N) - F

    e1, e2 = err(1e-3), err(2e-3)
    assert 2.7 < _order(e1, e2) < 3.3  # remainder O(phi^2 grad phi)
    assert np.abs(e1).max() < 1e-7 * np.abs(g2).max()


def test_gr_drift_exact_geodesic_to_first_order(O):
    N = 8
    rng = np.random.default_rng(290)
    phi0, chi0 = rng.standard_normal(N ** 3), rng.standard_normal(N ** 3)
    pos, q = particles(N, 400, 291, vmax=0.6)
    a, dtau = 0.4, 1e-4
    c, r = cell_frac(N, pos)
    q2 = np.sum(q * q, 0)
    ph0, ch0 = cic(phi0, N, c, r), cic(chi0, N, c, r)

    def err(s):
        p1 = pos.copy()
        O.update_pos(N, p1, q, dtau, s * phi0, s * chi0, None, 2, a, a * a * N)
        ph, ps = s * ph0, s * (ph0 - ch0)
        A, Bm = a * a * np.exp(2 * ps), a * a * np.exp(-2 * ph)
        v = q * np.sqrt(A / (Bm * (Bm + q2)))
        d = (p1 - pos + 0.5) % 1.0 - 0.5
        return d / dtau - v

    e1, e2 = err(1e-3), err(2e-3)
    assert 1.8 < _order(e1, e2) < 2.2  # remainder O(phi^2)


def test_T00_exact_point_particle_to_second_order(O):
    N = 8
    rng = np.random.default_rng(300)
    m, a = 0.5, 0.3
    for k in range(4):
        p = rng.uniform(0, 1, (3, 1))
        v = rng.uniform(-0.6, 0.6, (3, 1))
        q2 = float(np.sum(v * v))

        def err(s):
            T = O.project_T00(N, p, v, m, a=a, phi=np.full(N ** 3, s)).sum()
            return T / (m * N ** 3 / a) - np.exp(3 * s) * np.sqrt(a * a + q2 * np.exp(2 * s))

        e1, e2 = err(1e-2), err(2e-2)
        assert 2.8 < _order(np.array([e1]), np.array([e2])) < 3.2  # remainder O(phi^3)


def test_Tij_exact_point_particle_to_first_order(O):
    N = 8
    rng = np.random.default_rng(310)
    m, a = 0.5, 0.3
    for k in range(4):
        p = rng.uniform(0, 1, (3, 1))
        v = rng.uniform(-0.6, 0.6, (3, 1))
        q2 = float(np.sum(v * v))

        def err(s):
            T = O.project_Tij(N, p, v, m, a=a, phi=np.full(N ** 3, s)).sum(1)
            exact = np.array([v[i, 0] * v[j, 0] for (i, j) in COMPS]) * np.exp(5 * s) / np.sqrt(
                a * a + q2 * np.exp(2 * s))
            return T / (m * N ** 3 / a) - exact

        e1, e2 = err(1e-3), err(2e-3)
        assert 1.8 < _order(e1, e2) < 2.2  # remainder O(phi^2)
