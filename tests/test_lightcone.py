"""Lightcone particle output on the GPU (the lightcone form of saveGadget2,
Particles_gevolution.hpp:1270-2317; findIntersectingLightcones,
tools.hpp:619-862; the ID backlog of lightcone_id_backlog.hpp and
writeLightcones, output.hpp:966-2880) against the oracle on the same
particles: vertex lists equal, written IDs equal in container order (the
selection is the same IEEE arithmetic on both sides), positions and
velocities equal to float rounding, prelog and backlog hand-over equal."""
import numpy as np
import pytest

GVC = 3.335640952e-6
ZONE = 0.05


def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def read_snapshot(path):
    raw = open(path, "rb").read()
    o, recs = 0, []
    while o < len(raw):
        b = int(np.frombuffer(raw, np.uint32, 1, o)[0])
        recs.append(raw[o + 4:o + 4 + b])
        assert int(np.frombuffer(raw, np.uint32, 1, o + 4 + b)[0]) == b
        o += 8 + b
    assert len(recs) == 4 and len(recs[0]) == 256
    n = len(recs[3]) // 8
    return (recs[0], np.frombuffer(recs[1], np.float32).reshape(n, 3), np.frombuffer(recs[2], np.float32).reshape(n, 3),
            np.frombuffer(recs[3], np.int64))


def make(torch, gp, N, n, prec, seed, clustered=False):
    rng = np.random.default_rng(seed)
    ctx = gp.Context(N, gp.GEV_FP64 if prec == "fp64" else gp.GEV_MIXED)
    if clustered:
        c = rng.uniform(0, 1, (3, 8))
        pos = (c[:, rng.integers(0, 8, n)] + rng.normal(0, 0.04, (3, n))) % 1.0
    else:
        pos = rng.uniform(0, 1, (3, n))
    q = rng.normal(0, 0.05, (3, n))
    pid = rng.permutation(n).astype(np.int64) * 2 + 3
    p = gp.Particles(ctx, n, 1.0)
    p.load(pos, q, pid)
    phi_np = 1e-5 * rng.standard_normal((N, N, N))
    phi = torch.from_numpy(phi_np).to(ctx.rdtype).reshape(1, N, N, N).cuda()
    return ctx, p, phi


def oracle_view(p, phi):
    spos, sq, sid = (t.cpu().numpy() for t in p.store())
    return spos, sq, sid, phi.double().cpu().numpy().reshape(-1)


CASES = [
    # (lightcone vertex, direction, opening, dist, dtau, dtau_old, dadtau, a)
    ((0.5, 0.5, 0.5), (0.0, 0.0, 1.0), -1.0, 0.31, 0.02, 0.025, 0.6, 0.45),
    ((0.1, 0.8, 0.3), (0.6, 0.0, 0.8), 0.5, 0.74, 0.03, 0.02, 0.3, 0.8),
]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("clustered", [False, True])
def test_lightcone_matches_oracle(tmp_path, O, prec, case, clustered):
    torch = _gpu()
    import gevpm as gp
    from gevpm import lightcone as lcm
    N, n = 32, 60000
    lcv, dirn, opening, dist, dtau, dtau_old, dadtau, a = CASES[case]
    print("lightcone parity", prec, case, clustered, flush=True)
    ctx, p, phi = make(torch, gp, N, n, prec, 21 + case, clustered)
    spos, sq, sid, sphi = oracle_view(p, phi)
    geo = lcm.Geometry(vertex=lcv, direction=dirn, opening=opening)
    inner, outer = dist - 0.5 * dtau, dist + (0.5 + ZONE) * dtau_old
    nv, verts = lcm.vertices(geo, outer, inner, lcm.box_domain(N))
    onv, overts = O.lightcone_vertices(lcv, dirn, opening, outer, inner, lcm.box_domain(N), cap=32)
    assert nv == onv and np.array_equal(verts, overts)
    assert 0 < nv <= lcm.MAX_VERTICES
    # a backlog that holds every third ID
    rng = np.random.default_rng(5)
    bl = np.sort(rng.choice(sid, len(sid) // 3, replace=False))
    log = lcm.IDLog(ctx)
    log.set_backlog(bl[::-1].copy())  # the log sorts
    assert np.array_equal(log.backlog, bl)
    hdr = gp.gev_gadget2_header()
    hdr.num_files = 1
    hdr.time, hdr.BoxSize = a, 320000.0
    path = str(tmp_path / "lc_cdm")
    w = p.save_gadget2_lightcone(path, hdr, geo, dist, dtau, dtau_old, dadtau, verts, log, phi, tracer_factor=1)
    po, vo, io, pl, margin = O.lightcone_select(N, spos, sq, sid, sphi, 1, lcv, dirn, opening, dist, dtau, dtau_old,
                                                dadtau, a, 320000.0, verts, backlog=bl)
    assert w == len(io) and w > 200
    assert hdr.npart[1] == w and hdr.npartTotal[1] == w and hdr.npartTotalHW[1] == 0
    hb, fpos, fvel, fid = read_snapshot(path)
    assert np.frombuffer(hb, np.uint32, 6, 0)[1] == w
    assert np.array_equal(fid, io)  # same particles, same (container) order
    assert np.allclose(fpos, po, rtol=3e-7, atol=320000.0 * 2e-7)
    assert np.allclose(fvel, vo, rtol=3e-7, atol=1e-2)
    assert np.array_equal(log.prelog, io[pl])
    assert pl.sum() > 5
    # the backlog took check-zone particles out, and only backlog IDs
    _, _, io_nb, _, _ = O.lightcone_select(N, spos, sq, sid, sphi, 1, lcv, dirn, opening, dist, dtau, dtau_old,
                                           dadtau, a, 320000.0, verts)
    dropped = set(io_nb) - set(io)
    assert len(dropped) > 0 and dropped <= set(bl) and set(io) <= set(io_nb)
    log.commit()
    assert np.array_equal(log.backlog, np.unique(io[pl]))
    assert len(log.prelog) == 0


@pytest.mark.gpu
def test_lightcone_tracer_empty_and_errors(tmp_path, O):
    torch = _gpu()
    import gevpm as gp
    from gevpm import lightcone as lcm
    N, n = 16, 20000
    ctx, p, phi = make(torch, gp, N, n, "fp64", 4)
    spos, sq, sid, sphi = oracle_view(p, phi)
    geo = lcm.Geometry(vertex=(0.4, 0.4, 0.6))
    dist, dtau, dtau_old, dadtau, a = 0.25, 0.03, 0.03, 0.5, 0.5
    inner, outer = dist - 0.5 * dtau, dist + (0.5 + ZONE) * dtau_old
    _, verts = lcm.vertices(geo, outer, inner, lcm.box_domain(N))
    hdr = gp.gev_gadget2_header()
    hdr.num_files = 1
    hdr.time, hdr.BoxSize = a, 1.0
    w = p.save_gadget2_lightcone(str(tmp_path / "t"), hdr, geo, dist, dtau, dtau_old, dadtau, verts, None, None,
                                 tracer_factor=5)
    _, _, io, _, _ = O.lightcone_select(N, spos, sq, sid, None, 5, geo.vertex, geo.direction, -1.0, dist, dtau,
                                        dtau_old, dadtau, a, 1.0, verts)
    assert w == len(io) > 0 and (io % 5 == 0).all()
    assert np.array_equal(read_snapshot(str(tmp_path / "t"))[3], io)
    # a shell outside the box: nothing written, no file, counts zero
    w = p.save_gadget2_lightcone(str(tmp_path / "none"), hdr, geo, 0.0, 0.0, 0.0, dadtau, verts, None, None)
    assert w == 0 and hdr.npart[1] == 0 and not (tmp_path / "none").exists()
    with pytest.raises(gp.GevError):
        p.save_gadget2_lightcone(str(tmp_path / "x"), hdr, geo, dist, dtau, dtau_old, dadtau,
                                 np.zeros((33, 3)), None, None)
    with pytest.raises(gp.GevError):
        p.save_gadget2_lightcone(str(tmp_path / "x"), hdr, geo, dist, dtau, dtau_old, dadtau, verts, None, None,
                                 tracer_factor=0)
    other = lcm.IDLog(gp.Context(N, gp.GEV_FP64))
    with pytest.raises(gp.GevError):
        p.save_gadget2_lightcone(str(tmp_path / "x"), hdr, geo, dist, dtau, dtau_old, dadtau, verts, other, None)


@pytest.mark.gpu
def test_write_particles_over_cycles(tmp_path, O):
    """writeLightcones' Gadget branch over three cycles of a Newtonian run: a
    file per cycle while the shell is inside the lightcone's distance range,
    the backlog handed over each cycle, and few IDs written by both of two
    consecutive cycles (the backlog drops those in the overlap of the shells)"""
    torch = _gpu()
    import gevpm as gp
    from gevpm import lightcone as lcm
    import gevpm.cosmology as cosmo
    N, L, z_in = 32, 320.0, 10.0
    fpg = gp.fourpiG_of(L)
    ctx = gp.Context(N, gp.GEV_FP64)
    p = gp.Particles(ctx, N ** 3, 0.31 / N ** 3)
    f = gp.Fields(ctx)
    st = gp.StepState(0, fpg, 0.31, 0.69, z_relax=z_in)
    k, P = cosmo.bbks_table(z=z_in)
    gp.ic_generate(p, f, st, 7, L, z_in, k, P)
    # observer whose light cone at z=0 sweeps through the box during these cycles
    tau0 = gp.particle_horizon(1.0, fpg, 0.31, 0.69)
    dtau = 0.02
    lcs = [lcm.Geometry(vertex=(0.5, 0.5, 0.5), z=0.0, distance=(10.0, 0.0))]
    log = lcm.IDLog(ctx)
    seen = []
    for cycle in range(3):
        st.dtau = dtau
        gp.step(p, f, st)
        # pretend the cycle ends at a conformal time where the shell radius is ~0.3 box
        tau = tau0 - 0.3 + 0.02 * cycle
        outs = lcm.write_particles(p, lcs, [log], f.phi, st.a, tau, dtau, st.dtau_old, cycle,
                                   str(tmp_path / "lcdm"), L, fpg, 0.31, 0.69)
        assert len(outs) == 1 and outs[0].written > 0
        _, fpos, _, fid = read_snapshot(outs[0].path)
        assert outs[0].path.endswith(f"_{cycle:04d}_cdm")
        assert len(np.unique(fid)) == len(fid)
        seen.append(fid)
        # the backlog is the sorted prelog of this cycle's output
        assert np.all(np.diff(log.backlog) > 0) and set(log.backlog) <= set(fid)
    # consecutive shells overlap in the check zone only; IDs in both outputs
    # can only be ones outside the previous backlog's reach
    assert len(set(seen[0]) & set(seen[1])) < 0.05 * len(seen[1])
