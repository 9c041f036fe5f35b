"""Particle files: Gadget-2 snapshots (saveGadget2 / loadGadget2,
Particles_gevolution.hpp:597-800, 2455-2640) and the express restart format
(saveExpress / loadExpress, :2642-2900).  The byte layout of each file read
back with numpy, the unit conversions, the tracer selection, the
exact-redshift drift and kick against closed forms, and save -> load round
trips in both precision forms."""
import ctypes as C
import os

import numpy as np
import pytest

GVC = 3.335640952e-6  # GADGET_VELOCITY_CONVERSION, metadata.hpp:117


def test_header_is_256_bytes():
    from gevpm._lib import gev_gadget2_header
    assert C.sizeof(gev_gadget2_header) == 256
    # field offsets of gadget2_header (metadata.hpp:159-178)
    off = {f[0]: getattr(gev_gadget2_header, f[0]).offset for f in gev_gadget2_header._fields_}
    assert (off["mass"], off["time"], off["redshift"], off["npartTotal"], off["num_files"], off["BoxSize"],
            off["HubbleParam"], off["npartTotalHW"], off["fill"]) == (24, 72, 80, 96, 124, 128, 152, 168, 192)


def read_gadget2(path):
    """independent numpy reader of a single-file snapshot: (header bytes, pos, vel, ids)"""
    raw = open(path, "rb").read()
    o = 0
    recs = []
    while o < len(raw):
        b = np.frombuffer(raw, np.uint32, 1, o)[0]
        recs.append(raw[o + 4:o + 4 + b])
        assert np.frombuffer(raw, np.uint32, 1, o + 4 + b)[0] == b
        o += 8 + b
    assert len(recs) == 4 and len(recs[0]) == 256
    n = len(recs[1]) // 12
    assert len(recs[2]) == 12 * n and len(recs[3]) == 8 * n
    return (recs[0], np.frombuffer(recs[1], np.float32).reshape(n, 3), np.frombuffer(recs[2], np.float32).reshape(n, 3),
            np.frombuffer(recs[3], np.int64))


def header(gp, n, a, boxsize):
    from gevpm._lib import gev_gadget2_header
    h = gev_gadget2_header()
    h.npart[1] = n
    h.npartTotal[1] = n
    h.time = a
    h.redshift = 1 / a - 1
    h.num_files = 1
    h.BoxSize = boxsize
    h.Omega0, h.OmegaLambda, h.HubbleParam = 0.31, 0.69, 0.67
    return h


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_gadget2_layout_units_and_round_trip(tmp_path, prec):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gevpm as gp
    N, n, a, box = 16, 5000, 0.25, 320000.0
    ctx = gp.Context(N, gp.GEV_FP64 if prec == "fp64" else gp.GEV_MIXED)
    rng = np.random.default_rng(3)
    pos, q = rng.uniform(0, 1, (3, n)), rng.uniform(-0.2, 0.2, (3, n))
    pid = rng.permutation(n).astype(np.int64) * 2 + 1
    p = gp.Particles(ctx, n, 1.0)
    p.load(pos, q, pid)
    gpos, gq, gid = (t.cpu().numpy() for t in p.store())  # the container's (possibly FP32-rounded) state
    path = str(tmp_path / "snap_cdm")
    p.save_gadget2(path, header(gp, n, a, box))
    hb, fpos, fvel, fid = read_gadget2(path)
    assert np.frombuffer(hb, np.uint32, 6, 0)[1] == n
    assert np.frombuffer(hb, np.float64, 1, 72)[0] == a and np.frombuffer(hb, np.float64, 1, 128)[0] == box
    assert np.array_equal(fid, gid)  # container (cell) order
    # the IDs are odd, so a tracer factor of 2 selects none: empty records
    p.save_gadget2(str(tmp_path / "none"), header(gp, 0, a, box), tracer_factor=2)
    assert all(len(r) == 0 for r in read_gadget2(str(tmp_path / "none"))[1:])
    o = np.argsort(gid)
    f = np.argsort(fid)
    # positions p * BoxSize and velocities q / (a^1.5 GVC), rounded once to float
    assert np.allclose(fpos[f], (gpos[:, o] * box).T.astype(np.float32), rtol=2e-7, atol=box * 2e-7)
    assert np.allclose(fvel[f], (gq[:, o] / (a ** 1.5 * GVC)).T.astype(np.float32), rtol=2e-7)
    # loadGadget2 gives back the container state to float precision
    p2 = gp.Particles(ctx, n, 1.0)
    h2 = p2.load_gadget2(path)
    assert h2.npart[1] == n and h2.BoxSize == box
    lpos, lq, lid = (t.cpu().numpy() for t in p2.store())
    o2 = np.argsort(lid)
    assert np.array_equal(lid[o2], gid[o])
    d = np.abs(((lpos[:, o2] - gpos[:, o]) + 0.5) % 1.0 - 0.5)
    assert d.max() < 2e-7
    assert np.allclose(lq[:, o2], gq[:, o], rtol=3e-7, atol=1e-9)
    cells = p2.cells().cpu().numpy()
    assert np.all(np.diff(cells) >= 0)


@pytest.mark.gpu
def test_gadget2_tracer_and_exact_redshift(tmp_path):
    """ID % tracer_factor selection; the EXACT_OUTPUT_REDSHIFTS drift and kick
    against closed forms for a potential linear in x (CIC interpolation and
    the edge gradient of a linear field are exact)"""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gevpm as gp
    N, n, a, box = 16, 3000, 0.5, 1000.0
    ctx = gp.Context(N, gp.GEV_FP64)
    rng = np.random.default_rng(9)
    # keep x away from the periodic break of the linear potential
    pos = np.stack([rng.uniform(0.2, 0.7, n), rng.uniform(0, 1, n), rng.uniform(0, 1, n)])
    q = rng.uniform(-0.1, 0.1, (3, n))
    pid = np.arange(n, dtype=np.int64)
    p = gp.Particles(ctx, n, 1.0)
    p.load(pos, q, pid)
    s = 1e-4  # phi(x) = s * x_cell
    xs = np.arange(N, dtype=np.float64)
    phi = torch.from_numpy(np.broadcast_to(s * xs[None, None, :], (N, N, N)).copy()).reshape(1, N, N, N).cuda()
    sel = pid[pid % 3 == 0]
    dtp, dtv = 0.01, 0.02
    path = str(tmp_path / "snap_tr")
    p.save_gadget2(path, header(gp, len(sel), a, box), tracer_factor=3, dtau_pos=dtp, dtau_vel=dtv, phi=phi)
    _, fpos, fvel, fid = read_gadget2(path)
    cid = p.store()[2].cpu().numpy()
    assert np.array_equal(fid, cid[cid % 3 == 0])  # the tracers, in container order
    assert np.array_equal(np.sort(fid), sel)
    o = np.argsort(fid)
    P, Q = pos[:, sel], q[:, sel]
    phip = s * (P[0] * N)                       # CIC value of the linear field at the particle
    grad = np.stack([np.full(len(sel), s), np.zeros(len(sel)), np.zeros(len(sel))])  # edge difference per cell
    q2 = (Q ** 2).sum(0)
    e2 = q2 + a * a
    w = 2 * q2 + a * a
    want_pos = (P + dtp * Q * (1 + (4 - w / e2) * phip) / np.sqrt(e2)) % 1.0 * box
    want_vel = (Q - dtv * w * grad * N / np.sqrt(e2)) / np.sqrt(a) / GVC / a
    assert np.allclose(fpos[o], want_pos.T.astype(np.float32), rtol=3e-7, atol=box * 3e-7)
    assert np.allclose(fvel[o], want_vel.T.astype(np.float32), rtol=3e-7, atol=1e-3)


@pytest.mark.gpu
def test_gadget2_rejects_bad_requests(tmp_path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gevpm as gp
    ctx = gp.Context(8, gp.GEV_FP64)
    p = gp.Particles(ctx, 10, 1.0)
    p.load(np.full((3, 10), 0.5), np.zeros((3, 10)), np.arange(10, dtype=np.int64))
    with pytest.raises(gp.GevError):
        p.save_gadget2(str(tmp_path / "x"), header(gp, 7, 1.0, 1.0))  # npart does not match
    bad = str(tmp_path / "not_gadget")
    with open(bad, "wb") as fh:
        fh.write(os.urandom(64))
    with pytest.raises(gp.GevError):
        p.load_gadget2(bad)
    small = gp.Particles(ctx, 5, 1.0)
    p.save_gadget2(str(tmp_path / "ok"), header(gp, 10, 1.0, 1.0))
    with pytest.raises(gp.GevError):
        small.load_gadget2(str(tmp_path / "ok"))  # capacity 5 < 10 particles


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_express_layout_and_round_trip(tmp_path, prec):
    """saveExpress / loadExpress: 120-byte express header, Real triplets of
    positions and momenta, int64 IDs, the Gadget-2 header last; the file name
    gets the rank suffix .0 in place of a numeric one"""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gevpm as gp
    N, n = 16, 4000
    ctx = gp.Context(N, gp.GEV_FP64 if prec == "fp64" else gp.GEV_MIXED)
    rng = np.random.default_rng(12)
    p = gp.Particles(ctx, n, 1.0)
    p.load(rng.uniform(0, 1, (3, n)), rng.uniform(-0.3, 0.3, (3, n)), rng.permutation(n).astype(np.int64) + 100)
    gpos, gq, gid = (t.cpu().numpy() for t in p.store())
    h = p.save_express(str(tmp_path / "restart.7"), header(gp, 0, 0.1, 640000.0))
    assert h.npart[1] == n and h.npartTotal[1] == n
    path = tmp_path / "restart.0"
    raw = path.read_bytes()
    rs = 8 if prec == "fp64" else 4
    assert raw[:8] == b"GEXPCDM\x00"
    u32 = np.frombuffer(raw, np.uint32, 6, 8)
    assert tuple(u32) == (1, 120, 0x01020304, rs, 8, 256)
    # rank, size, grid_size[2], lat_size[3], lat_size_local[3], coord_skip[2]
    assert tuple(np.frombuffer(raw, np.int32, 12, 32)) == (0, 1, 1, 1, N, N, N, N, N, N, 0, 0)
    u64 = np.frombuffer(raw, np.uint64, 5, 80)
    assert tuple(u64) == (n, n, 3 * n * rs, 3 * n * rs, 8 * n)
    assert len(raw) == 120 + 6 * n * rs + 8 * n + 256
    dt = np.float64 if rs == 8 else np.float32
    fpos = np.frombuffer(raw, dt, 3 * n, 120).reshape(n, 3)
    fid = np.frombuffer(raw, np.int64, n, 120 + 6 * n * rs)
    assert np.array_equal(fid, gid)  # container order
    assert np.array_equal(fpos, gpos.T.astype(dt))
    p2 = gp.Particles(ctx, n, 1.0)
    h2 = p2.load_express(str(tmp_path / "restart"))
    assert h2.npart[1] == n and h2.BoxSize == 640000.0
    lpos, lq, lid = (t.cpu().numpy() for t in p2.store())
    o, o2 = np.argsort(gid), np.argsort(lid)
    assert np.array_equal(lid[o2], gid[o])
    tol = 0.0 if prec == "fp64" else 1.2e-7
    assert np.abs(((lpos[:, o2] - gpos[:, o]) + 0.5) % 1.0 - 0.5).max() <= tol
    assert np.array_equal(lq[:, o2], gq[:, o])  # momenta are stored in the container's own precision
    # another lattice is refused
    other = gp.Particles(gp.Context(32, ctx.precision), n, 1.0)
    with pytest.raises(gp.GevError):
        other.load_express(str(tmp_path / "restart"))
