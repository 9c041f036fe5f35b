"""Pins of the oracle's lightcone particle output (CPU only).

* vertex search (findIntersectingLightcones, tools.hpp:619-862): every
  periodic copy of the observer whose shell [inner, outer] and cone meet a
  dense sample of the box is found, the full-sky set equals the copies whose
  shell meets the box (nearest and farthest box points), the visiting order
  is u-major, and a copy that looks away from the box is not found;
* particle selection against the geometric definition, over all periodic
  copies (any hit, not only the first), for particles at rest;
* the crossing: a particle at rest stays put, a radially moving one ends on
  the observer's past light cone exactly (dadtau = dtau_old = 0);
* the kick to the crossing for a potential linear in x (exact CIC);
* ID backlog and prelog zones from the output geometry.
"""
import itertools

import numpy as np
import pytest

GVC = 3.335640952e-6
ZONE = 0.05


def box_domain(N):
    # writeLightcones' domain for one rank (output.hpp:1175-1190), box units
    return np.array([-0.5 / N] * 3 + [(N + 0.5) / N] * 3)


def copies(lcv, rng_):
    r = range(-rng_, rng_ + 1)
    return np.array([[lcv[0] + u, lcv[1] + v, lcv[2] + w] for u, v, w in itertools.product(r, r, r)])


def box_samples(dom, m):
    g = [np.linspace(dom[d], dom[d + 3], m) for d in range(3)]
    z, y, x = np.meshgrid(g[2], g[1], g[0], indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1)


@pytest.mark.parametrize("inner,outer", [(0.2, 0.3), (0.55, 0.7), (1.05, 1.2)])
def test_vertices_full_sky_exact(O, inner, outer):
    N = 16
    dom = box_domain(N)
    lcv = np.array([0.3, 0.6, 0.45])
    n, v = O.lightcone_vertices(lcv, [0, 0, 1], -1.0, outer, inner, dom, cap=200)
    lo, hi = dom[:3], dom[3:]
    want = []
    rg = int(np.ceil(outer)) + 1
    for c in copies(lcv, rg):
        near = np.sqrt((np.maximum(0, np.maximum(lo - c, c - hi)) ** 2).sum())
        far = np.sqrt((np.maximum(np.abs(c - lo), np.abs(c - hi)) ** 2).sum())
        if near <= outer and far >= inner:
            want.append(c)
    want = np.array(want)
    assert n == len(want) == len(v)
    assert np.allclose(v, want)  # same u-major order
    # the samples' view: each found copy sees a box point inside the shell
    S = box_samples(dom, 25)
    for c in v:
        d = np.linalg.norm(S - c, axis=1)
        h = np.sqrt(3) * (dom[3] - dom[0]) / 24
        assert ((d >= inner - h) & (d <= outer + h)).any()


def test_vertices_cone(O):
    N = 8
    dom = box_domain(N)
    lcv = np.array([0.5, 0.5, 0.5])
    dirn = np.array([1.0, 0.0, 0.0])
    opening = 0.9
    inner, outer = 1.9, 2.1
    n, v = O.lightcone_vertices(lcv, dirn, opening, outer, inner, dom, cap=200)
    assert n == len(v)
    found = {tuple(np.round(c - lcv).astype(int)) for c in v}
    # every copy with a sampled box point inside both the shell and the cone
    S = box_samples(dom, 21)
    for c in copies(lcv, int(np.ceil(outer)) + 1):
        r = S - c
        d = np.linalg.norm(r, axis=1)
        ok = (d >= inner) & (d <= outer) & ((r @ dirn) / np.maximum(d, 1e-300) > opening)
        if ok.any():
            assert tuple(np.round(c - lcv).astype(int)) in found
    assert (-2, 0, 0) in found     # looks at the box centre along +x
    assert (2, 0, 0) not in found  # the box lies behind it


def _setup(N, n, seed, q_scale=0.0, tf=1):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, 1, (3, n))
    q = rng.normal(0, q_scale, (3, n)) if q_scale else np.zeros((3, n))
    pid = rng.permutation(n).astype(np.int64) * 3 + 1
    return pos, q, pid


@pytest.mark.parametrize("opening", [-1.0, 0.6])
def test_selection_at_rest(O, opening):
    N, n = 16, 40000
    pos, q, pid = _setup(N, n, 5)
    lcv = np.array([0.25, 0.5, 0.7])
    dirn = np.array([0.0, 0.6, 0.8])
    dist, dtau, dtau_old, a, box = 0.62, 0.02, 0.018, 0.5, 1000.0
    inner, outer = dist - 0.5 * dtau, dist + (0.5 + ZONE) * dtau_old
    nv, verts = O.lightcone_vertices(lcv, dirn, opening, outer, inner, box_domain(N))
    assert nv <= O.MAX_INTERSECTS
    po, vo, io, pl, margin = O.lightcone_select(N, pos, q, pid, None, 1, lcv, dirn, opening, dist, dtau, dtau_old,
                                                0.0, a, box, verts)
    # geometric definition over all periodic copies
    P = pos.T
    hit = np.zeros(n, bool)
    for c in copies(lcv, 2):
        r = P - c
        d = np.linalg.norm(r, axis=1)
        hit |= (d >= inner) & (d < outer) & ((opening == -1.0) | ((r @ dirn) / d > opening))
    assert np.array_equal(np.sort(io), np.sort(pid[hit]))
    assert len(io) > 500
    # at rest: the output is the particle seen from its copy, unmoved
    rel = po.astype(np.float64) / box - lcv
    d = np.linalg.norm(rel, axis=1)
    tol = 4e-7 * (1 + np.abs(lcv).max() + 1)
    assert (d >= inner - tol).all() and (d < outer + tol).all()
    if opening != -1.0:
        assert ((rel @ dirn) / d > opening - 1e-6).all()
    o = {k: i for i, k in enumerate(pid)}
    src = P[[o[k] for k in io]]
    vv = src - rel  # the copy of the observer each particle was seen from
    assert np.allclose(vv - lcv, np.round(vv - lcv), atol=1e-5)
    assert (vo == 0).all()
    # it is the first copy in the vertex list that sees the particle in the shell
    for x, w in zip(src[:400], vv[:400]):
        r = x - verts
        dv = np.linalg.norm(r, axis=1)
        ok = (dv >= inner) & (dv < outer) & ((opening == -1.0) | ((r @ dirn) / dv > opening))
        assert np.allclose(verts[np.argmax(ok)], w, atol=1e-5)


def test_tracer_factor_and_backlog(O):
    N, n = 16, 30000
    pos, q, pid = _setup(N, n, 7)
    lcv = np.array([0.5, 0.5, 0.5])
    dist, dtau, dtau_old, a, box = 0.35, 0.04, 0.05, 0.4, 100.0
    inner, outer = dist - 0.5 * dtau, dist + (0.5 + ZONE) * dtau_old
    _, verts = O.lightcone_vertices(lcv, [1, 0, 0], -1.0, outer, inner, box_domain(N))
    args = (lcv, [1, 0, 0], -1.0, dist, dtau, dtau_old, 0.0, a, box, verts)
    _, _, io_all, _, _ = O.lightcone_select(N, pos, q, pid, None, 1, *args)
    _, _, io_tf, _, _ = O.lightcone_select(N, pos, q, pid, None, 4, *args)
    assert np.array_equal(np.sort(io_tf), np.sort(io_all[io_all % 4 == 0]))
    po, _, io, pl, _ = O.lightcone_select(N, pos, q, pid, None, 1, *args)
    d = np.linalg.norm(po.astype(np.float64) / box - lcv, axis=1)
    far = outer - d > 2 * ZONE * dtau_old + 1e-5
    near = outer - d < 2 * ZONE * dtau_old - 1e-5
    assert near.sum() > 20 and far.sum() > 20
    # prelog: within 2 ZONE dtau of the inner radius
    pz = d - inner
    sure = np.abs(pz - 2 * ZONE * dtau) > 1e-5
    assert np.array_equal(pl[sure], (pz < 2 * ZONE * dtau)[sure])
    # backlog of every ID: only particles outside the check zone remain
    bl = np.sort(pid)
    _, _, io_b, _, _ = O.lightcone_select(N, pos, q, pid, None, 1, *args, backlog=bl)
    assert set(io[far]) <= set(io_b) and not (set(io[near]) & set(io_b))
    # backlog of the even IDs: odd check-zone particles stay
    bl2 = np.sort(pid[pid % 2 == 0])
    _, _, io_b2, _, _ = O.lightcone_select(N, pos, q, pid, None, 1, *args, backlog=bl2)
    keep = far | (io % 2 == 1)
    sure2 = far | near
    assert set(io[sure2 & keep]) <= set(io_b2) and not (set(io[sure2 & ~keep]) & set(io_b2))


def test_radial_motion_ends_on_the_light_cone(O):
    """with dadtau = dtau_old = 0 a particle moving along the line of sight is
    moved to |x - observer| = dist - dtau_pos, the observer's past light cone
    at the crossing time; its velocity is q / (a^1.5 GVC)"""
    N, n = 16, 20000
    rng = np.random.default_rng(11)
    lcv = np.array([0.5, 0.5, 0.5])
    dist, dtau, a, box = 0.3, 0.05, 0.3, 1.0
    u = rng.normal(size=(3, n))
    u /= np.linalg.norm(u, axis=0)
    r0 = rng.uniform(dist - 0.02, dist + 0.0, n)
    pos = lcv[:, None] + u * r0
    speed = rng.uniform(-0.4, 0.4, n)  # q, radial
    q = u * speed
    pid = np.arange(n, dtype=np.int64)
    inner, outer = dist - 0.5 * dtau, dist
    _, verts = O.lightcone_vertices(lcv, [1, 0, 0], -1.0, outer, inner, box_domain(N))
    po, vo, io, _, _ = O.lightcone_select(N, pos, q, pid, None, 1, lcv, [1, 0, 0], -1.0, dist, dtau, 0.0, 0.0, a,
                                          box, verts)
    assert len(io) > 1000
    e = np.sqrt(speed[io] ** 2 + a * a)
    v = speed[io] / e                       # radial coordinate velocity
    d0 = r0[io]
    dtp = (dist - d0) / (1 + v)             # light-cone crossing time
    rel = po.astype(np.float64) - lcv
    dn = np.linalg.norm(rel, axis=1)
    assert np.allclose(dn, dist - dtp, atol=3e-7)
    assert np.allclose(vo, (q[:, io] / (a ** 1.5 * GVC)).T, rtol=2e-7, atol=1e-3)


def test_kick_linear_potential(O):
    """at rest in phi = s x_cell: vel_x = -(dtau_pos + dtau_old/2) s N / (sqrt(a') GVC)
    with dtau_pos = dist - d and a' = a + dtau_pos dadtau"""
    N, n = 16, 20000
    rng = np.random.default_rng(13)
    pos = np.stack([rng.uniform(0.2, 0.7, n), rng.uniform(0, 1, n), rng.uniform(0, 1, n)])
    q = np.zeros((3, n))
    pid = np.arange(n, dtype=np.int64)
    s = 1e-4
    phi = np.broadcast_to(s * np.arange(N, dtype=np.float64)[None, None, :], (N, N, N)).copy()
    lcv = np.array([0.45, 0.5, 0.5])
    dist, dtau, dtau_old, dadtau, a, box = 0.2, 0.03, 0.02, 0.7, 0.6, 1.0
    inner, outer = dist - 0.5 * dtau, dist + (0.5 + ZONE) * dtau_old
    _, verts = O.lightcone_vertices(lcv, [1, 0, 0], -1.0, outer, inner, box_domain(N))
    po, vo, io, _, _ = O.lightcone_select(N, pos, q, pid, phi, 1, lcv, [1, 0, 0], -1.0, dist, dtau, dtau_old,
                                          dadtau, a, box, verts)
    keep = (pos[0, io] > 0.26) & (pos[0, io] < 0.64)  # same cells' gradient, away from the x break
    assert keep.sum() > 500
    d = np.linalg.norm(pos[:, io].T - lcv, axis=1)
    dtp = dist - d
    ap = a + dtp * dadtau
    want = -(dtp + 0.5 * dtau_old) * s * N / (np.sqrt(ap) * GVC)
    assert np.allclose(vo[keep, 0], want[keep], rtol=3e-6)
    assert (vo[:, 1:] == 0).all()
    assert np.allclose(po.astype(np.float64), pos[:, io].T - verts[0] + lcv, atol=3e-7)
