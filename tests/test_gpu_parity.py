"""GPU parity: every HIP entry point against the FP64 oracle on identical
seeded inputs, called through the C ABI.  Tolerances: FP64 paths are held to
relative max-norm errors of 1e-13..1e-11 (reordered sums); FP32 (mixed form)
paths to the single-precision tolerances written beside each check.  Particle
results are compared per particle ID (the GPU keeps particles cell-sorted)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def gp():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gevpm
    return gevpm


def relmax(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def to_np(t):
    return t.detach().cpu().numpy()


def real_np(t):
    """(c,N,N,N) tensor -> (c, N^3) float64 array"""
    a = to_np(t).astype(np.float64)
    return a.reshape(a.shape[0], -1)


def ft_np(ctx, T):
    return to_np(ctx.unpad(T)).astype(np.complex128)


def ft_gpu(ctx, F):
    """numpy (c,N,N,N/2+1) complex -> padded device tensor"""
    F = np.asarray(F)
    if F.ndim == 3:
        F = F[None]
    out = ctx.fourier(F.shape[0])
    out[..., : ctx.N // 2 + 1] = torch.from_numpy(F).to(ctx.cdtype).to(ctx.device)
    return out


def field_gpu(ctx, f, ncomp=1):
    return torch.from_numpy(np.ascontiguousarray(f).reshape(ncomp, ctx.N, ctx.N, ctx.N)).to(ctx.rdtype).to(ctx.device)


# --------------------------------------------------------------------- FFT
@pytest.mark.parametrize("N", [4, 8, 16, 32, 64, 128])
def test_fft_fp64_vs_oracle_and_numpy(gp, O, N):
    ctx = gp.Context(N, gp.GEV_FP64)
    rng = np.random.default_rng(N)
    f = rng.standard_normal((2, N ** 3))
    F = ctx.fft_r2c(field_gpu(ctx, f, 2))
    Fn = ft_np(ctx, F)
    for c in range(2):
        ref = np.fft.rfftn(f[c].reshape(N, N, N))
        assert relmax(Fn[c], ref) < 1e-13
        if N <= 64:
            assert relmax(Fn[c], O.fft_r2c(f[c])) < 1e-13
    # padding columns are zero
    assert torch.all(F[..., N // 2 + 1:] == 0)
    g = ctx.fft_c2r(F)
    assert relmax(real_np(g) / N ** 3, f) < 1e-13
    # non-Hermitian input: imaginary DC / Nyquist ignored like the oracle
    G = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    g2 = ctx.fft_c2r(ft_gpu(ctx, G))
    assert relmax(real_np(g2)[0], O.fft_c2r(G)) < 1e-12


@pytest.mark.parametrize("N", [16, 64, 256])
def test_fft_fp32(gp, N):
    ctx = gp.Context(N, gp.GEV_MIXED)
    rng = np.random.default_rng(N + 1)
    f = rng.standard_normal(N ** 3).astype(np.float32)
    F = ft_np(ctx, ctx.fft_r2c(field_gpu(ctx, f)))[0]
    ref = np.fft.rfftn(f.astype(np.float64).reshape(N, N, N))
    assert relmax(F, ref) < 3e-6  # FP32: ~eps * log2(N^3)
    g = real_np(ctx.fft_c2r(ctx.fft_r2c(field_gpu(ctx, f))))[0] / N ** 3
    assert relmax(g, f) < 3e-6


def test_fft_512_roundtrip(gp):
    N = 512
    for prec, tol in ((gp.GEV_FP64, 1e-13), (gp.GEV_MIXED, 5e-6)):
        ctx = gp.Context(N, prec)
        g = torch.Generator(device="cuda").manual_seed(5)
        f = torch.randn((1, N, N, N), generator=g, device="cuda", dtype=ctx.rdtype)
        F = ctx.fft_r2c(f)
        # Parseval
        w = torch.full((N // 2 + 1,), 2.0, device="cuda", dtype=torch.float64)
        w[0] = w[N // 2] = 1.0
        e_k = (w * (ctx.unpad(F).abs().double() ** 2)).sum().item() / N ** 3
        e_x = (f.double() ** 2).sum().item()
        assert abs(e_k / e_x - 1) < tol * 10
        h = ctx.fft_c2r(F, preserve=False)
        err = ((h / N ** 3 - f).abs().max() / f.abs().max()).item()
        assert err < tol
        del F, h, f, ctx
        torch.cuda.empty_cache()


# ----------------------------------------------------------------- k-space
def hermitian_ft(O, rng, N, ncomp):
    return np.stack([O.fft_r2c(rng.standard_normal(N ** 3)) for _ in range(ncomp)])


@pytest.mark.parametrize("N", [8, 32])
def test_kspace_ops_fp64(gp, O, N):
    ctx = gp.Context(N, gp.GEV_FP64)
    rng = np.random.default_rng(100 + N)
    S1 = hermitian_ft(O, rng, N, 1)[0]
    for modif in (0.0, 12.5):
        got = ft_np(ctx, ctx.poisson(ft_gpu(ctx, S1), 0.37, modif))[0]
        assert relmax(got, O.poisson(S1, 0.37, modif)) < 1e-14
    S6 = hermitian_ft(O, rng, N, 6)
    got = ft_np(ctx, ctx.project_ft_scalar(ft_gpu(ctx, S6)))[0]
    assert relmax(got, O.project_ft_scalar(S6)) < 1e-13
    base = ft_gpu(ctx, S1)
    got = ft_np(ctx, ctx.project_ft_scalar(ft_gpu(ctx, S6), base))[0]
    assert relmax(got, O.project_ft_scalar(S6, S1)) < 1e-13
    S3 = hermitian_ft(O, rng, N, 3)
    for modif in (0.0, 3.0):
        got = ft_np(ctx, ctx.project_ft_vector(ft_gpu(ctx, S3), 0.7, modif))
        assert relmax(got, O.project_ft_vector(S3, 0.7, modif)) < 1e-13
    B = ft_gpu(ctx, S3)
    got = ft_np(ctx, ctx.evolve_ft_vector(ft_gpu(ctx, S6), B, 0.21))
    assert relmax(got, O.evolve_ft_vector(S6, S3, 0.21)) < 1e-13


def test_kspace_ops_mixed(gp, O):
    N = 32
    ctx = gp.Context(N, gp.GEV_MIXED)
    rng = np.random.default_rng(7)
    S6 = hermitian_ft(O, rng, N, 6).astype(np.complex64).astype(np.complex128)
    got = ft_np(ctx, ctx.project_ft_scalar(ft_gpu(ctx, S6)))[0]
    assert relmax(got, O.project_ft_scalar(S6)) < 1e-6  # FP64 combination, FP32 storage
    S1 = S6[0]
    got = ft_np(ctx, ctx.poisson(ft_gpu(ctx, S1), 0.37, 2.0))[0]
    assert relmax(got, O.poisson(S1, 0.37, 2.0)) < 1e-6


# ----------------------------------------------------------------- sources
@pytest.mark.parametrize("N", [8, 32])
def test_sources_fp64(gp, O, N):
    ctx = gp.Context(N, gp.GEV_FP64)
    rng = np.random.default_rng(200 + N)
    phi = 1e-3 * rng.standard_normal(N ** 3)
    chi = 1e-4 * rng.standard_normal(N ** 3)
    src = 0.31 + rng.standard_normal(N ** 3)
    args = (0.31, 0.2, 1.3, 0.07)
    res, s1, s2 = ctx.prepare_source_T00(field_gpu(ctx, phi), field_gpu(ctx, chi), field_gpu(ctx, src), *args)
    ref, sref = O.prepare_source_T00(phi, chi, src, *args)
    assert relmax(real_np(res)[0], ref) < 1e-14
    assert np.isclose(s1, sref, rtol=1e-12)
    assert np.isclose(s2, ref.sum(), rtol=1e-12, atol=1e-12 * np.abs(ref).sum())
    T = rng.standard_normal((6, N ** 3))
    got = real_np(ctx.prepare_source_Tij(field_gpu(ctx, phi), field_gpu(ctx, T, 6), 0.9))
    assert relmax(got, O.prepare_source_Tij(phi, T, 0.9)) < 1e-14
    # in place
    Tg = field_gpu(ctx, T, 6)
    ctx.prepare_source_Tij(field_gpu(ctx, phi), Tg, 0.9, Sij=Tg)
    assert relmax(real_np(Tg), O.prepare_source_Tij(phi, T, 0.9)) < 1e-14


# --------------------------------------------------------------- particles
def make_particles(N, n, seed, vmax=0.05, cluster=False):
    rng = np.random.default_rng(seed)
    if cluster:
        # half the particles crammed into a handful of cells, the rest uniform
        c = rng.integers(0, N, (3, 1)) / N
        pos = np.concatenate([c + rng.uniform(0, 2.0 / N, (3, n // 2)), rng.uniform(0, 1, (3, n - n // 2))], 1) % 1.0
    else:
        pos = rng.uniform(0, 1, (3, n))
    q = rng.uniform(-vmax, vmax, (3, n))
    pid = rng.permutation(n).astype(np.int64) * 3 + 7
    return pos, q, pid


def gpu_pcls(gp, ctx, pos, q, pid, mass, cap=None):
    p = gp.Particles(ctx, cap or max(pos.shape[1], 1), mass)
    p.load(pos, q, pid)
    return p


def by_id(p):
    pos, q, pid = p.store()
    pos, q, pid = to_np(pos), to_np(q), to_np(pid)
    o = np.argsort(pid)
    return np.ascontiguousarray(pos[:, o]), np.ascontiguousarray(q[:, o]), pid[o]


@pytest.mark.parametrize("N,n", [(32, 50000), (256, 400000)])
def test_sort_invariants(gp, N, n):
    """cell keys, order and the cell-start table; at 256^3 the block sums of
    the scan are long enough for its 16-byte path"""
    ctx = gp.Context(N, gp.GEV_FP64)
    pos, q, pid = make_particles(N, n, 1, cluster=True)
    p = gpu_pcls(gp, ctx, pos, q, pid, 1.0)
    gpos, gq, gid = (to_np(t) for t in p.store())
    assert np.array_equal(np.sort(gid), np.sort(pid))
    cells = to_np(p.cells())
    c = np.floor(gpos * N).astype(np.int64)
    assert np.array_equal(cells, c[0] + N * (c[1] + N * c[2]))          # bit-exact keys
    assert np.all(np.diff(cells) >= 0)                                    # sorted by cell
    start = to_np(p.cell_start())
    assert start[0] == 0 and start[-1] == n
    assert np.array_equal(np.diff(start), np.bincount(cells, minlength=N ** 3))
    o = np.argsort(gid); oi = np.argsort(pid)
    assert np.array_equal(gpos[:, o], pos[:, oi]) and np.array_equal(gq[:, o], q[:, oi])


@pytest.mark.parametrize("N,n,cluster", [(4, 300, False), (16, 20000, False), (32, 60000, True), (64, 100000, False)])
def test_deposits_fp64(gp, O, N, n, cluster):
    ctx = gp.Context(N, gp.GEV_FP64)
    pos, q, pid = make_particles(N, n, 300 + N, vmax=0.2, cluster=cluster)
    m, a = 0.31 / n, 0.05
    p = gpu_pcls(gp, ctx, pos, q, pid, m)
    phi = 1e-3 * np.random.default_rng(N).standard_normal(N ** 3)
    gphi = field_gpu(ctx, phi)
    out = ctx.real(1)
    p.project_mass(out)
    assert relmax(real_np(out)[0], O.project_T00(N, pos, q, m)) < 1e-13
    p.project_T00(out, a, gphi, 1.3)
    assert relmax(real_np(out)[0], O.project_T00(N, pos, q, m, a=a, phi=phi, coeff=1.3)) < 1e-13
    p.project_T00(out, a, gphi, 1.3, accumulate=True)
    assert relmax(real_np(out)[0], 2 * O.project_T00(N, pos, q, m, a=a, phi=phi, coeff=1.3)) < 1e-13
    o3 = ctx.real(3)
    for ph in (None, phi):
        p.project_T0i(o3, None if ph is None else gphi, 0.8)
        assert relmax(real_np(o3), O.project_T0i(N, pos, q, m, phi=ph, coeff=0.8)) < 1e-13
    o6 = ctx.real(6)
    for ph in (None, phi):
        p.project_Tij(o6, a, None if ph is None else gphi, 0.9)
        assert relmax(real_np(o6), O.project_Tij(N, pos, q, m, a=a, phi=ph, coeff=0.9)) < 1e-13
    # T00 and Tij in one pass
    o1 = ctx.real(1)
    for ph in (None, phi):
        p.project_T00_Tij(o1, o6, a, None if ph is None else gphi, 0.9)
        assert relmax(real_np(o1)[0], O.project_T00(N, pos, q, m, a=a, phi=ph, coeff=0.9)) < 1e-13
        assert relmax(real_np(o6), O.project_Tij(N, pos, q, m, a=a, phi=ph, coeff=0.9)) < 1e-13


def test_deposits_empty_and_single(gp, O):
    N = 8
    ctx = gp.Context(N, gp.GEV_FP64)
    p = gp.Particles(ctx, 4, 1.0)
    p.load(np.zeros((3, 0)), np.zeros((3, 0)), np.zeros(0, np.int64))
    out = ctx.real(1)
    out.fill_(5.0)
    p.project_mass(out)
    assert torch.all(out == 0)
    pos = np.array([[0.999], [0.0], [0.5]])
    p.load(pos, np.array([[0.1], [0.2], [0.3]]), np.array([1], np.int64))
    p.project_Tij(ctx.real(6), 0.3)
    p.project_mass(out)
    assert relmax(real_np(out)[0], O.project_T00(N, pos, np.zeros((3, 1)), 1.0)) < 1e-14


@pytest.mark.parametrize("N,cluster", [(16, False), (64, False), (64, True)])
def test_deposits_mixed(gp, O, N, cluster):
    """clustered: half the particles in 8 cells (LDS atomics on shared sites).
    The FP32 form sums each tile in 32-bit fixed point scaled to its largest
    per-site particle count: a site fed by k of the tile's particles carries a
    rounding error of about sqrt(k) half-units of (k_max e_max / 2^30), the
    order of an FP32 running sum; with 65536 particles on one site that is
    ~1e-5 of the largest site value"""
    tol = 3e-5 if cluster else 2e-6
    ctx = gp.Context(N, gp.GEV_MIXED)
    n = 4 * N ** 3 // 8
    pos, q, pid = make_particles(N, n, 400 + N, vmax=0.2, cluster=cluster)
    m, a = 0.31 / n, 0.05
    p = gpu_pcls(gp, ctx, pos, q, pid, m)
    # the mixed form rounds offsets and momenta to FP32: compare with the oracle on those inputs
    gpos, gq, gid = by_id(p)
    o = np.argsort(pid)
    assert np.abs(((gpos - pos[:, o]) + 0.5) % 1.0 - 0.5).max() < 1.0 / N * 1.2e-7
    phi = 1e-3 * np.random.default_rng(N).standard_normal(N ** 3).astype(np.float32).astype(np.float64)
    out = ctx.real(1)
    p.project_T00(out, a, field_gpu(ctx, phi))
    errs = [relmax(real_np(out)[0], O.project_T00(N, gpos, gq, m, a=a, phi=phi))]
    o6 = ctx.real(6)
    p.project_Tij(o6, a, field_gpu(ctx, phi))
    errs.append(relmax(real_np(o6), O.project_Tij(N, gpos, gq, m, a=a, phi=phi)))
    o1 = ctx.real(1)
    p.project_T00_Tij(o1, o6, a, field_gpu(ctx, phi))
    errs.append(relmax(real_np(o1)[0], O.project_T00(N, gpos, gq, m, a=a, phi=phi)))
    errs.append(relmax(real_np(o6), O.project_Tij(N, gpos, gq, m, a=a, phi=phi)))
    print(f"mixed deposits N={N} cluster={cluster}: rel max errors T00, Tij, T00 (pair), Tij (pair) {errs}", flush=True)
    assert max(errs) < tol


@pytest.mark.parametrize("N,cluster", [(32, False), (64, True), (128, False)])
def test_deposits_mixed_after_update(gp, O, N, cluster):
    """After a particle update, the fixed-point deposits bound each tile by
    the update's largest |q|^2 over the 8^3-cell blocks the tile's halo box
    touches, instead of a pass over the tile's particles.  The drift moves
    particles up to ~2.5 cells at N = 128, so some land outside their tile's
    window of blocks.  Both bounds must give T00, Tij and the T0i link sums
    within the fixed-point tolerance of the FP64 oracle on the updated
    particles; a particle above its block bound would overflow its tile's
    32-bit accumulators and fail by orders of magnitude."""
    ctx = gp.Context(N, gp.GEV_MIXED)
    n = N ** 3 // 2
    pos, q, pid = make_particles(N, n, 2100 + N, vmax=0.2, cluster=cluster)
    rng = np.random.default_rng(N + 11)
    f32 = lambda x: x.astype(np.float32).astype(np.float64)
    phi, chi = f32(1e-3 * rng.standard_normal(N ** 3)), f32(1e-4 * rng.standard_normal(N ** 3))
    B = f32(1e-2 * rng.standard_normal((3, N ** 3)))
    m = 0.31 / n
    p = gpu_pcls(gp, ctx, pos, q, pid, m)
    a_k, a_d = 0.2, 0.21
    p.kick_drift(1, 0.013, 0.02, field_gpu(ctx, phi), field_gpu(ctx, chi), field_gpu(ctx, B, 3), 3, 3, a_k, a_d,
                 a_k ** 2 * N, a_d ** 2 * N)
    gpos, gq, gid = by_id(p)
    moved = np.abs(((gpos - pos[:, np.argsort(pid)]) + 0.5) % 1.0 - 0.5).max() * N
    a = 0.05
    ref00 = O.project_T00(N, gpos, gq, m, a=a, phi=phi)
    refij = O.project_Tij(N, gpos, gq, m, a=a, phi=phi)
    ref0i = O.project_T0i(N, gpos, gq, m)
    tol = 3e-5 if cluster else 2e-6
    tol_l = 2e-3 if cluster else 2e-6
    got = {}
    for form in ("block bounds", "particle pass"):
        if form == "particle pass":
            p.sort()  # particles may have been written: the next deposits bound their tiles themselves
        T00, Tij, Lk = p.project_T00_Tij_T0i_links(ctx.real(1), ctx.real(6), ctx.real(3), a, field_gpu(ctx, phi))
        got[form] = (to_np(T00), to_np(Tij), to_np(Lk))
        errs = (relmax(real_np(T00)[0], ref00), relmax(real_np(Tij), refij), relmax(real_np(Lk), ref0i))
        print(f"mixed deposits after an update N={N} cluster={cluster} ({form}, largest move {moved:.2f} cells): "
              f"rel max errors T00, Tij, T0i links {errs}", flush=True)
        assert errs[0] < tol and errs[1] < tol and errs[2] < tol_l
    # the two bounds differ, so the fixed-point roundings do too
    assert any(not np.array_equal(x, y) for x, y in zip(got["block bounds"], got["particle pass"]))


# ---------------------------------------------------------- kick / drift --
@pytest.mark.parametrize("N", [8, 32])
def test_kick_drift_fp64(gp, O, N):
    ctx = gp.Context(N, gp.GEV_FP64)
    n = 20000
    pos, q, pid = make_particles(N, n, 500 + N, vmax=0.3)
    o = np.argsort(pid)
    rng = np.random.default_rng(N)
    phi, chi = 1e-3 * rng.standard_normal(N ** 3), 1e-4 * rng.standard_normal(N ** 3)
    B = 1e-2 * rng.standard_normal((3, N ** 3))
    gphi, gchi, gB = field_gpu(ctx, phi), field_gpu(ctx, chi), field_gpu(ctx, B, 3)
    a, bc = 0.2, 0.04 * N
    for nf in (1, 2, 3):
        p = gpu_pcls(gp, ctx, pos, q, pid, 1.0)
        qo = q.copy()
        ret = p.update_q(0.013, gphi, gchi if nf >= 2 else None, gB if nf >= 3 else None, nf, a, bc)
        rref = O.update_q(N, pos, qo, 0.013, phi, chi, B, nf, a, bc)
        _, gq, _ = by_id(p)
        assert relmax(gq, qo[:, o]) < 1e-13
        assert np.isclose(ret, rref, rtol=1e-13)
    p = gpu_pcls(gp, ctx, pos, q, pid, 1.0)
    qo = q.copy()
    ret = p.update_q_newton(0.013, gphi, a)
    rref = O.update_q_newton(N, pos, qo, 0.013, phi, a)
    assert relmax(by_id(p)[1], qo[:, o]) < 1e-13 and np.isclose(ret, rref, rtol=1e-13)
    for nf in (0, 1, 2, 3):
        p = gpu_pcls(gp, ctx, pos, q, pid, 1.0)
        po = pos.copy()
        p.update_pos(0.02, gphi, gchi, gB, nf, a, bc)
        O.update_pos(N, po, q, 0.02, phi, chi, B, nf, a, bc)
        gpos = by_id(p)[0]
        assert np.abs(((gpos - po[:, o]) + 0.5) % 1.0 - 0.5).max() < 1e-15
        cells = to_np(p.cells())
        assert np.all(np.diff(cells) >= 0)
    p = gpu_pcls(gp, ctx, pos, q, pid, 1.0)
    po = pos.copy()
    p.update_pos_newton(0.02, a)
    O.update_pos_newton(N, po, q, 0.02, a)
    assert np.abs(((by_id(p)[0] - po[:, o]) + 0.5) % 1.0 - 0.5).max() < 1e-15
    p = gpu_pcls(gp, ctx, pos, q, pid, 1.0)
    po = pos.copy()
    d = p.displace(gphi, 0.5)
    dref = O.displace_particles(N, po, phi, 0.5)
    assert np.abs(((by_id(p)[0] - po[:, o]) + 0.5) % 1.0 - 0.5).max() < 1e-15 and np.isclose(d, dref, rtol=1e-13)
    p = gpu_pcls(gp, ctx, pos, q, pid, 1.0)
    qo = np.zeros_like(q)
    r = p.initialize_q(gphi, 3.0)
    rref = O.initialize_q(N, pos, qo, phi, 3.0)
    assert relmax(by_id(p)[1], qo[:, o]) < 1e-13 and np.isclose(r, rref, rtol=1e-13)


# -------------------------------------------------------------- IC / step --
C_LIGHT = 2997.92458


def gpu_ic(gp, N, L, gr, seed=42, z_in=100.0, om=0.31):
    import gevpm.cosmology as cosmo
    k, P = cosmo.bbks_table(omega_m=om, z=z_in)
    ctx = gp.Context(N, gp.GEV_FP64)
    p = gp.Particles(ctx, N ** 3, om / N ** 3)
    f = gp.Fields(ctx)
    st = gp.StepState(gr, gp.fourpiG_of(L), om, 1 - om, z_relax=z_in)
    md = gp.ic_generate(p, f, st, seed, L, z_in, k, P)
    return ctx, p, f, st, md, (k, P)


@pytest.mark.parametrize("gr", [0, 1])
def test_ic_generate_vs_oracle(gp, O, gr):
    N, L = 32, 320.0
    ctx, p, f, st, md, (k, P) = gpu_ic(gp, N, L, gr)
    ic = O.ic_generate(N, 42, k, P, L, gp.fourpiG_of(L), 0.31, 0.69, 1 / 101.0, gr=gr)
    gpos, gq, gid = by_id(p)
    assert np.array_equal(gid, np.arange(N ** 3))
    assert np.abs(((gpos - ic["pos"]) + 0.5) % 1.0 - 0.5).max() < 1e-12 / N
    assert relmax(gq, ic["q"]) < 1e-10
    assert relmax(real_np(f.phi)[0], ic["phi"]) < 1e-10
    assert relmax(real_np(f.chi)[0], ic["chi"]) < 1e-8
    assert relmax(real_np(f.Bi), ic["Bi"]) < 1e-8
    assert np.isclose(st.maxvel, ic["maxvel"], rtol=1e-10)
    assert np.isclose(md, ic["max_disp"], rtol=1e-10)


@pytest.mark.parametrize("gr,N,steps", [(0, 64, 10), (1, 32, 6)])
def test_steps_vs_oracle(gp, O, gr, N, steps):
    """C1-type Newton cycles and GR weak-field cycles from the same IC."""
    L = 320.0 if gr == 0 else 640.0
    ctx, p, f, st, md, _ = gpu_ic(gp, N, L, gr)
    pos0, q0, pid = by_id(p)
    M = N ** 3
    ost = dict(pos=pos0.copy(), q=q0.copy(), phi=real_np(f.phi)[0].copy(), chi=real_np(f.chi)[0].copy(),
               Bi=real_np(f.Bi).copy(), source=np.zeros(M), Sij=np.zeros((6, M)), a=st.a, tau=st.tau,
               dtau_old=0.0, maxvel=st.maxvel)
    fpg = gp.fourpiG_of(L)
    a_end = 1 / 51.0 if gr == 0 else 1 / 80.0
    t1 = gp.particle_horizon(a_end, fpg, 0.31, 0.69)
    dtau = (t1 - st.tau) / steps
    for _ in range(steps):
        st.dtau = dtau
        gp.step(p, f, st)
        ost["dtau"] = dtau
        O.step(N, ost, gr, 0.31 / M, fpg, 0.31, 0.69, 100.0)
    gpos, gq, _ = by_id(p)
    assert np.isclose(st.a, ost["a"], rtol=1e-14)
    dpos = np.abs(((gpos - ost["pos"]) + 0.5) % 1.0 - 0.5).max()
    assert dpos < 1e-11 / N, dpos
    assert relmax(gq, ost["q"]) < 1e-9
    assert relmax(real_np(f.phi)[0], ost["phi"]) < 1e-9
    if gr:
        assert relmax(real_np(f.chi)[0], ost["chi"]) < 1e-7
        assert relmax(real_np(f.Bi), ost["Bi"]) < 1e-7
    assert np.isclose(st.maxvel, ost["maxvel"], rtol=1e-9)
    # cell keys per ID across the two implementations: equal except within eps of a face
    eps = 1e-9
    xa, xb = gpos * N, ost["pos"] * N
    mism = np.any(np.floor(xa) != np.floor(xb), axis=0)
    near = np.any(np.abs(xa - np.rint(xa)) < eps, axis=0) | np.any(np.abs(xb - np.rint(xb)) < eps, axis=0)
    print(f"  cell keys: {int(mism.sum())} differ, {int(near.sum())} particles within {eps} cells of a face")
    assert np.all(near[mism])


@pytest.mark.parametrize("gr", [0, 1])
def test_fused_kick_drift_vs_oracle(gp, O, gr):
    """gev_kick_drift = update_q then update_pos with the kicked momentum (main.cu:1232-1266)."""
    N = 16
    ctx = gp.Context(N, gp.GEV_FP64)
    n = 30000
    pos, q, pid = make_particles(N, n, 900 + gr, vmax=0.3)
    o = np.argsort(pid)
    rng = np.random.default_rng(77)
    phi, chi = 1e-3 * rng.standard_normal(N ** 3), 1e-4 * rng.standard_normal(N ** 3)
    B = 1e-2 * rng.standard_normal((3, N ** 3))
    p = gpu_pcls(gp, ctx, pos, q, pid, 1.0)
    a_k, a_d = 0.2, 0.21
    args = (field_gpu(ctx, phi), field_gpu(ctx, chi), field_gpu(ctx, B, 3))
    if gr:
        ret = p.kick_drift(1, 0.013, 0.02, *args, 3, 3, a_k, a_d, a_k ** 2 * N, a_d ** 2 * N)
        qo, po = q.copy(), pos.copy()
        rref = O.update_q(N, po, qo, 0.013, phi, chi, B, 3, a_k, a_k ** 2 * N)
        O.update_pos(N, po, qo, 0.02, phi, chi, B, 3, a_d, a_d ** 2 * N)
    else:
        ret = p.kick_drift(0, 0.013, 0.02, args[0], None, None, 1, 0, a_k, a_d, 1.0, 1.0)
        qo, po = q.copy(), pos.copy()
        rref = O.update_q_newton(N, po, qo, 0.013, phi, a_k)
        O.update_pos_newton(N, po, qo, 0.02, a_d)
    gpos, gq, gid = by_id(p)
    assert np.array_equal(gid, np.sort(pid))
    assert relmax(gq, qo[:, o]) < 1e-13
    assert np.abs(((gpos - po[:, o]) + 0.5) % 1.0 - 0.5).max() < 1e-15
    assert np.isclose(ret, rref, rtol=1e-13)
    cells = to_np(p.cells())
    c = np.floor(gpos * N).astype(np.int64)
    assert np.array_equal(np.sort(cells), np.sort(c[0] + N * (c[1] + N * c[2])))


@pytest.mark.parametrize("N,cluster", [(32, False), (128, False), (128, True)])
def test_kick_drift_mixed(gp, O, N, cluster):
    """mixed-form fused GR kick+drift (FP32 offsets, momenta and fields) against
    the FP64 oracle run on the same FP32-rounded inputs; the clustered case
    puts half the particles in 8 cells (many passes over one tile)"""
    ctx = gp.Context(N, gp.GEV_MIXED)
    n = N ** 3 // 2
    pos, q, pid = make_particles(N, n, 1300 + N, vmax=0.3, cluster=cluster)
    rng = np.random.default_rng(N + 5)
    f32 = lambda x: x.astype(np.float32).astype(np.float64)
    phi, chi = f32(1e-3 * rng.standard_normal(N ** 3)), f32(1e-4 * rng.standard_normal(N ** 3))
    B = f32(1e-2 * rng.standard_normal((3, N ** 3)))
    p = gpu_pcls(gp, ctx, pos, q, pid, 1.0)
    gpos, gq, gid = by_id(p)
    a_k, a_d = 0.2, 0.21
    args = (field_gpu(ctx, phi), field_gpu(ctx, chi), field_gpu(ctx, B, 3))
    ret = p.kick_drift(1, 0.013, 0.02, *args, 3, 3, a_k, a_d, a_k ** 2 * N, a_d ** 2 * N)
    qo, po = gq.copy(), gpos.copy()
    rref = O.update_q(N, po, qo, 0.013, phi, chi, B, 3, a_k, a_k ** 2 * N)
    O.update_pos(N, po, qo, 0.02, phi, chi, B, 3, a_d, a_d ** 2 * N)
    npos, nq, nid = by_id(p)
    assert np.array_equal(nid, gid)
    # FP32 stencil sums: a few ulp of the kick; offsets round to 2^-24 of a
    # cell, and the drift carries the kick's FP32 error
    assert relmax(nq, qo) < 3e-6
    assert np.abs(((npos - po) + 0.5) % 1.0 - 0.5).max() < 2e-6 / N
    assert np.isclose(ret, rref, rtol=1e-5)
    cells = to_np(p.cells())
    assert np.all(np.diff(cells) >= 0)
    c = np.floor(npos * N).astype(np.int64) % N
    assert np.array_equal(np.sort(cells), np.sort(c[0] + N * (c[1] + N * c[2])))


# ------------------------------------------------------------------- P(k)
@pytest.mark.parametrize("N,prec,numbins,deconv", [(16, "fp64", 24, True), (64, "fp64", 24, True),
                                                    (64, "fp64", 7, False), (128, "fp64", 200, True),
                                                    (64, "mixed", 24, True), (128, "mixed", 64, False)])
def test_power_spectrum_vs_oracle(gp, O, N, prec, numbins, deconv):
    """gev_power_spectrum (extractPowerSpectrum semantics) against the oracle's
    on the same modes (for the mixed form: the complex64 modes, widened)"""
    from gevpm import spectra
    ctx = gp.Context(N, gp.GEV_FP64 if prec == "fp64" else gp.GEV_MIXED)
    rng = np.random.default_rng(5 + N)
    f = rng.standard_normal(N ** 3)
    F = O.fft_r2c(f)
    Fd = ft_gpu(ctx, F)
    print("power spectrum", N, prec, numbins, deconv, flush=True)
    res = spectra.power_spectrum(ctx, Fd, numbins, deconvolve=deconv)
    Fo = to_np(ctx.unpad(Fd)[0]).astype(np.complex128)
    ref = O.power_spectrum(Fo, numbins, deconvolve=deconv)
    # kbin and power: sums of positive terms, held to 1e-12.  The scatters are
    # square roots of differences of such sums: a bin whose modes share one |k|
    # (or one power) leaves only rounding, of size sqrt(eps) times the bin's
    # mean, which the order of the sums decides
    for got, want in zip(res[:2], ref[:2]):
        assert relmax(got.numpy(), want) < 1e-12
    for got, want, mean in zip(res[2:4], ref[2:4], ref[:2]):
        assert (np.abs(got.numpy() - want) <= 1e-8 * np.abs(want).max() + 4e-8 * np.abs(mean)).all()
    assert np.array_equal(res[4].numpy(), ref[4])
    # every mode below k_max except k = 0 is counted once, its conjugate too
    tk = (2 * np.pi * np.abs(np.fft.fftfreq(N, 1.0 / N))) ** 2
    sk = np.sqrt((tk[:, None, None] + tk[None, :, None]) + tk[None, None, :])
    inside = np.floor(numbins * sk / np.sqrt(3 * tk[N // 2])) < numbins
    assert res[4].sum() == int(inside.sum()) - 1


@pytest.mark.parametrize("N,prec,case", [(32, "fp64", "cross"), (32, "fp64", "grid"), (32, "fp64", "vector"),
                                         (32, "fp64", "symmetric"), (32, "fp64", "pair"), (64, "mixed", "cross"),
                                         (64, "mixed", "symmetric")])
def test_cross_spectrum_vs_oracle(gp, O, N, prec, case):
    """gev_cross_spectrum (extractCrossSpectrum, tools.hpp:65-330) against the
    oracle on the same modes: two scalar fields; KTYPE_GRID bins; a 3-vector
    with itself; a symmetric tensor (xx xy xz yy yz zz) against another; one
    component pair.  Products of two different fields can cancel in a bin, so
    power is held to 1e-12 of the bin's largest |power|."""
    from gevpm import spectra
    ctx = gp.Context(N, gp.GEV_FP64 if prec == "fp64" else gp.GEV_MIXED)
    rng = np.random.default_rng(41 + N)
    ncomp = {"cross": 1, "grid": 1, "vector": 3, "symmetric": 6, "pair": 3}[case]
    F1 = np.stack([O.fft_r2c(rng.standard_normal(N ** 3)) for _ in range(ncomp)])
    F2 = F1 if case in ("grid", "vector") else np.stack([O.fft_r2c(rng.standard_normal(N ** 3)) for _ in range(ncomp)])
    D1, D2 = ft_gpu(ctx, F1), ft_gpu(ctx, F2)
    kw = dict(deconvolve=True, ktype=0 if case == "grid" else 1, symmetric=case == "symmetric")
    if case == "pair":
        kw.update(comp1=2, comp2=0)
    print("cross spectrum", N, prec, case, flush=True)
    res = spectra.cross_spectrum(ctx, D1, D2, 20, **kw)
    ref = O.cross_spectrum(ft_np(ctx, D1), ft_np(ctx, D2), 20, **kw)
    assert np.array_equal(res[4].numpy(), ref[4])
    assert relmax(res[0].numpy(), ref[0]) < 1e-12
    assert np.abs(res[1].numpy() - ref[1]).max() <= 1e-12 * np.abs(ref[1]).max()
    assert np.abs(res[3].numpy() - ref[3]).max() <= 1e-7 * np.abs(ref[3]).max()
    with pytest.raises(gp.GevError):
        spectra.cross_spectrum(ctx, D1, D2, 20, ktype=2)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_power_spectrum_single_mode_and_errors(gp, prec):
    """a single real cosine mode: one bin holds power |F|^2 k^3 / W^2 at its k,
    the others none; bad arguments are refused"""
    from gevpm import spectra
    N = 32
    ctx = gp.Context(N, gp.GEV_FP64 if prec == "fp64" else gp.GEV_MIXED)
    F = ctx.fourier(1)
    m = (3, 1, 2)  # (kx, ky, kz)
    F[0, m[2], m[1], m[0]] = 2.0 + 1.0j
    nb = 40
    kb, pw, ks, ps, occ = spectra.power_spectrum(ctx, F, nb, deconvolve=True)
    k = 2 * np.pi * np.sqrt(sum(x * x for x in m))
    kmax = np.sqrt(3) * 2 * np.pi * N / 2
    b = int(np.floor(nb * k / kmax))
    W = np.prod([np.sin(np.pi * x / N) * N / (np.pi * x) for x in m])
    assert np.isclose(pw[b].item() * occ[b].item(), 2 * 5.0 * k ** 3 / W ** 2, rtol=1e-6)
    assert (pw.numpy()[np.arange(nb) != b] == 0).all()
    with pytest.raises(gp.GevError):
        spectra.power_spectrum(ctx, F, 0)
    with pytest.raises(gp.GevError):
        spectra.power_spectrum(ctx, F, 2000)
    with pytest.raises(gp.GevError):
        spectra.power_spectrum(ctx, F.to(torch.complex128 if prec == "mixed" else torch.complex64), 8)


# --------------------------------------------- fused z-pass solves
@pytest.mark.parametrize("N,prec", [(64, "fp64"), (128, "fp64"), (64, "mixed"), (256, "mixed")])
def test_fused_solves_vs_unfused(gp, N, prec):
    """gev_solve_poisson / _chi / _vector (operator inside the z pass) against
    the separate r2c -> k-space kernel -> c2r sequence on the same inputs"""
    ctx = gp.Context(N, gp.GEV_FP64 if prec == "fp64" else gp.GEV_MIXED)
    tol = 1e-12 if prec == "fp64" else 2e-5
    g = torch.Generator(device="cpu").manual_seed(N)
    src = torch.randn(1, N, N, N, generator=g, dtype=torch.float64).to(ctx.rdtype).to(ctx.device)
    # Poisson with modif (GR form) and with the k=0 mode replaced
    coeff, modif, k0 = N * N * 1.0, 0.37, 1234.5
    F = ctx.fft_r2c(src)
    ctx.set_k0(F, k0)
    ref = ctx.fft_c2r(ctx.poisson(F, coeff, modif), preserve=False)
    got = ctx.solve_poisson(src, coeff, modif, k0=k0)
    assert relmax(to_np(got), to_np(ref)) < tol
    # Poisson, Newton form (modif = 0, k=0 dropped)
    ref = ctx.fft_c2r(ctx.poisson(ctx.fft_r2c(src), 0.8), preserve=False)
    got = ctx.solve_poisson(src, 0.8)
    assert relmax(to_np(got), to_np(ref)) < tol
    # chi from a random symmetric tensor
    S = torch.randn(6, N, N, N, generator=g, dtype=torch.float64).to(ctx.rdtype).to(ctx.device)
    ref = ctx.fft_c2r(ctx.project_ft_scalar(ctx.fft_r2c(S)), preserve=False)
    got = ctx.solve_chi(S)
    assert relmax(to_np(got), to_np(ref)) < tol
    # chi with Sij formed inside the x passes (N >= 128; the sequence below that)
    phi = 1e-3 * torch.randn(1, N, N, N, generator=g, dtype=torch.float64).to(ctx.rdtype).to(ctx.device)
    T = torch.randn(6, N, N, N, generator=g, dtype=torch.float64).to(ctx.rdtype).to(ctx.device)
    Sref = ctx.prepare_source_Tij(phi, T, 0.7)
    ref = ctx.solve_chi(Sref)
    T2 = T.clone()
    got = ctx.solve_chi_Tij(phi, T2, 0.7)
    assert relmax(to_np(got), to_np(ref)) < tol
    if N >= 128:
        assert torch.equal(T2, T)  # Tij is only read
    # B from a random vector source, in place
    Bsrc = torch.randn(3, N, N, N, generator=g, dtype=torch.float64).to(ctx.rdtype).to(ctx.device)
    ref = ctx.fft_c2r(ctx.project_ft_vector(ctx.fft_r2c(Bsrc), 0.25, 0.0), preserve=False)
    got = ctx.solve_vector(Bsrc.clone(), 0.25, 0.0)
    assert relmax(to_np(got), to_np(ref)) < tol


# ------------------------------- T0i deposited as link sums (one particle pass)
@pytest.mark.parametrize("N,prec,cluster", [(32, "fp64", False), (128, "fp64", True), (128, "mixed", False),
                                            (128, "mixed", True)])
def test_t0i_links_match_separate_projection(gp, N, prec, cluster):
    """gev_project_T00_Tij_T0i_links gives gev_project_T00_Tij's T00 and Tij,
    and link sums that, times 1 + phi_a + phi_b of a second phi, are
    gev_project_T0i with that phi; gev_solve_vector_T0i_links equals
    gev_solve_vector of that T0i (x passes with the factor for N >= 128,
    the separate factor kernel below)"""
    ctx = gp.Context(N, gp.GEV_FP64 if prec == "fp64" else gp.GEV_MIXED)
    n = 3 * N ** 3 // 2
    pos, q, pid = make_particles(N, n, 1700 + N, vmax=0.2, cluster=cluster)
    p = gpu_pcls(gp, ctx, pos, q, pid, 0.31 / n)
    g = torch.Generator(device="cpu").manual_seed(N + 3)
    phi_old = (1e-4 * torch.randn(1, N, N, N, generator=g, dtype=torch.float64)).to(ctx.rdtype).to(ctx.device)
    phi_new = (1e-4 * torch.randn(1, N, N, N, generator=g, dtype=torch.float64)).to(ctx.rdtype).to(ctx.device)
    a = 0.3
    T00a, Tija = p.project_T00_Tij(ctx.real(1), ctx.real(6), a, phi_old)
    T00b, Tijb, Lk = p.project_T00_Tij_T0i_links(ctx.real(1), ctx.real(6), ctx.real(3), a, phi_old)
    tol = 1e-12 if prec == "fp64" else 1e-6
    assert relmax(to_np(T00b), to_np(T00a)) < tol
    assert relmax(to_np(Tijb), to_np(Tija)) < tol
    # the link factor applied on the host, against the projection with phi_new
    T0i = to_np(p.project_T0i(ctx.real(3), phi_new)).astype(np.float64)
    ph = to_np(phi_new)[0].astype(np.float64)  # [z, y, x]
    L = to_np(Lk).astype(np.float64)
    fac = [1 + ph + np.roll(ph, -1, axis=2), 1 + ph + np.roll(ph, -1, axis=1), 1 + ph + np.roll(ph, -1, axis=0)]
    # FP32 form: both sides sum in 32-bit fixed point scaled to the tile's
    # largest per-site count times the largest |q_i|; where half the particles
    # share 8 cells, their momenta (random signs) nearly cancel, and the
    # rounding of ~65000 adds per site is a few 1e-4 of the largest net sum,
    # the size an FP32 running sum would also have
    tol_l = 1e-12 if prec == "fp64" else (2e-3 if cluster else 2e-6)
    assert relmax(np.stack([L[c] * fac[c] for c in range(3)]), T0i) < tol_l
    # the B solve from the link sums
    Bref = ctx.solve_vector(p.project_T0i(ctx.real(3), phi_new), 0.25, 0.0)
    Bgot = ctx.solve_vector_T0i_links(Lk.clone(), phi_new, 0.25, 0.0)
    assert relmax(to_np(Bgot), to_np(Bref)) < (1e-10 if prec == "fp64" else (2e-3 if cluster else 2e-5))


def test_gr_steps_128_vs_oracle(gp, O):
    """GR cycles at 128^3, where gev_step deposits T0i on the links together
    with T00 and Tij and applies the link factor inside the B solve, against
    the oracle's main.cu-ordered cycle"""
    N, L, steps = 128, 640.0, 3
    ctx, p, f, st, md, _ = gpu_ic(gp, N, L, 1)
    pos0, q0, pid = by_id(p)
    M = N ** 3
    ost = dict(pos=pos0.copy(), q=q0.copy(), phi=real_np(f.phi)[0].copy(), chi=real_np(f.chi)[0].copy(),
               Bi=real_np(f.Bi).copy(), source=np.zeros(M), Sij=np.zeros((6, M)), a=st.a, tau=st.tau,
               dtau_old=0.0, maxvel=st.maxvel)
    fpg = gp.fourpiG_of(L)
    t1 = gp.particle_horizon(1 / 90.0, fpg, 0.31, 0.69)
    dtau = (t1 - st.tau) / steps
    for _ in range(steps):
        st.dtau = dtau
        gp.step(p, f, st)
        ost["dtau"] = dtau
        O.step(N, ost, 1, 0.31 / M, fpg, 0.31, 0.69, 100.0)
    gpos, gq, _ = by_id(p)
    dpos = np.abs(((gpos - ost["pos"]) + 0.5) % 1.0 - 0.5).max()
    assert dpos < 1e-11 / N, dpos
    assert relmax(gq, ost["q"]) < 1e-9
    assert relmax(real_np(f.phi)[0], ost["phi"]) < 1e-9
    assert relmax(real_np(f.chi)[0], ost["chi"]) < 1e-7
    assert relmax(real_np(f.Bi), ost["Bi"]) < 1e-7


# ------------------------- prepareFTsource T00 formed inside the Poisson solve
@pytest.mark.parametrize("N,prec", [(64, "fp64"), (128, "fp64"), (128, "mixed"), (256, "mixed")])
def test_solve_poisson_T00_vs_separate(gp, N, prec):
    """gev_solve_poisson_T00 (source formed in the x pass for N >= 128) against
    gev_prepare_source_T00 followed by gev_solve_poisson with the k=0 mode set
    from the FP64 sum; both sums returned as the source pass returns them"""
    ctx = gp.Context(N, gp.GEV_FP64 if prec == "fp64" else gp.GEV_MIXED)
    g = torch.Generator(device="cpu").manual_seed(5 * N)
    rnd = lambda s: (s * torch.randn(1, N, N, N, generator=g, dtype=torch.float64)).to(ctx.rdtype).to(ctx.device)
    phi, chi, src = rnd(1e-4), rnd(1e-5), 1.0 + rnd(0.3)
    bg, c1, c2, c3 = 0.31, 2.3e-3, 4.1e-4, 7.7e-7
    coeff, modif = float(N * N), 0.21
    k0set = prec == "mixed"
    res, s_src, s_res = ctx.prepare_source_T00(phi, chi, src, bg, c1, c2, c3)
    ref = ctx.solve_poisson(res, coeff, modif, k0=s_res if k0set else None)
    src2 = src.clone()
    got, (g_src, g_res) = ctx.solve_poisson_T00(phi, chi, src2, bg, c1, c2, c3, coeff, modif, k0set=k0set)
    tol = 1e-12 if prec == "fp64" else 2e-5
    assert relmax(to_np(got), to_np(ref)) < tol
    assert abs(g_src - s_src) <= 1e-12 * abs(s_src) and abs(g_res - s_res) <= 1e-9 * max(abs(s_res), 1.0)
    if N >= 128:
        assert torch.equal(src2, src)  # the source is only read


def test_mixed_gr_cycle_bit_reproducible(gp):
    """the mixed GR cycle is bit-reproducible run to run: the fixed-point deposit
    sums do not depend on the order of the LDS adds (the order of particles
    inside a cell, which follows the rank atomics, may differ), the transforms
    and the updates have fixed arithmetic, and the particle set is compared by
    ID.  Two runs of 6 cycles at 64^3 from the same IC give identical offsets,
    momenta, cell keys, fields and background"""
    N, L, steps = 64, 320.0, 6
    ctx, p, f, st, md, _ = gpu_ic(gp, N, L, 1)
    a_, b_, c_ = p.store()
    t1 = gp.particle_horizon(1 / 60.0, st.fourpiG, 0.31, 0.69)
    dtau = (t1 - st.tau) / steps
    runs = []
    for _ in range(2):
        cm = gp.Context(N, gp.GEV_MIXED)
        pm = gp.Particles(cm, N ** 3, 0.31 / N ** 3)
        pm.load(a_, b_, c_)
        fm = gp.Fields(cm)
        fm.copy_from(f)
        sm = gp.StepState(1, st.fourpiG, 0.31, 0.69, z_relax=100.0)
        for key in ("a", "tau", "dtau_old", "maxvel"):
            setattr(sm, key, getattr(st, key))
        for _ in range(steps):
            sm.dtau = dtau
            gp.step(pm, fm, sm)
        pos, q, pid = (t.cpu().numpy() for t in pm.store())
        o = np.argsort(pid)
        runs.append((pos[:, o], q[:, o], real_np(fm.phi).copy(), real_np(fm.chi).copy(), real_np(fm.Bi).copy(),
                     sm.a, sm.T00hom, sm.maxvel))
    r0, r1 = runs
    for x, y in zip(r0[:5], r1[:5]):
        assert np.array_equal(x, y)
    assert r0[5:] == r1[5:]


def test_gr_steps_128_mixed_vs_fp64(gp):
    """the mixed GR cycle at 128^3 (T00 source inside the phi solve with the FP64
    k=0 mode, T0i on the links, FP32 particles) against the FP64 cycle from the
    same IC"""
    N, L, steps = 128, 640.0, 3
    ctx, p, f, st, md, _ = gpu_ic(gp, N, L, 1)
    cm = gp.Context(N, gp.GEV_MIXED)
    pm = gp.Particles(cm, N ** 3, 0.31 / N ** 3)
    a_, b_, c_ = p.store()
    pm.load(a_, b_, c_)
    fm = gp.Fields(cm)
    fm.copy_from(f)
    sm = gp.StepState(1, st.fourpiG, 0.31, 0.69, z_relax=100.0)
    for key in ("a", "tau", "dtau_old", "maxvel"):
        setattr(sm, key, getattr(st, key))
    t1 = gp.particle_horizon(1 / 90.0, st.fourpiG, 0.31, 0.69)
    dtau = (t1 - st.tau) / steps
    for _ in range(steps):
        st.dtau = dtau
        sm.dtau = dtau
        gp.step(p, f, st)
        gp.step(pm, fm, sm)
    gpos, gq, _ = by_id(p)
    mpos, mq, _ = by_id(pm)
    assert np.abs(((mpos - gpos) + 0.5) % 1.0 - 0.5).max() * N < 2e-5
    assert relmax(mq, gq) < 2e-5
    assert relmax(real_np(fm.phi)[0], real_np(f.phi)[0]) < 1e-3
    assert np.isclose(sm.T00hom, st.T00hom, rtol=1e-6)
    assert np.isclose(sm.a, st.a, rtol=1e-12)
