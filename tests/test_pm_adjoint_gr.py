"""Adjoint of the weak-field GR cycle (csrc/gev_adjoint_gr.hip, gevpm.adjoint.PMRunGR).

Every reverse kernel is checked against central differences of the forward
library call it reverses (FP64, particles kept away from cell faces so no
CIC kink lies inside the difference interval), the FFT-diagonal solves by
dot-product tests, and whole GR runs against central differences of the
forward run and across checkpoint intervals."""
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


def dev_t(x, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype).cuda()


def dot(a, b):
    return float((a.double() * b.double()).sum())


def particles(gp, ctx, rng, n, margin=0.1):
    N = ctx.N
    c = rng.integers(2, N - 2, size=(3, n))
    r = rng.uniform(margin, 1 - margin, size=(3, n))
    pos = (c + r) / N
    pid = rng.permutation(n).astype(np.int64)
    p = gp.Particles(ctx, n, 0.31 / N ** 3)
    return p, pos, pid


def by_id(p):
    pos, q, pid = p.store()
    o = torch.argsort(pid)
    return pos[:, o].clone(), q[:, o].clone()


def check_fd(jfun, adj_dir, eps, tol, label):
    jp, jm = jfun(eps), jfun(-eps)
    fd = (jp - jm) / (2 * eps)
    print(f"{label}: adjoint {adj_dir:.12e} fd {fd:.12e} rel {abs(adj_dir - fd) / abs(fd):.2e}", flush=True)
    assert abs(adj_dir - fd) <= tol * abs(fd)


def test_adjoint_kick_drift_matches_differences(gp):
    N, n = 16, 300
    ctx = gp.Context(N, gp.GEV_FP64)
    rng = np.random.default_rng(1)
    p, pos, pid = particles(gp, ctx, rng, n)
    q = rng.normal(0, 0.02, (3, n))
    a_k, a_d, dt_k, dt_d = 0.1, 0.102, 0.015, 0.02
    phi = dev_t(rng.normal(0, 1e-3, (1, N, N, N)))
    chi = dev_t(rng.normal(0, 1e-3, (1, N, N, N)))
    B = dev_t(rng.normal(0, 1e-5, (3, N, N, N)))
    lxv, lqv = dev_t(rng.normal(size=(3, n))), dev_t(rng.normal(size=(3, n)))
    d = dict(x=rng.normal(size=(3, n)) / N, q=rng.normal(0, 0.02, (3, n)), phi=rng.normal(0, 1e-3, (1, N, N, N)),
             chi=rng.normal(0, 1e-3, (1, N, N, N)), B=rng.normal(0, 1e-5, (3, N, N, N)))
    o = np.argsort(pid)

    def J(e):
        p.load(pos + e * d["x"][:, pid], q + e * d["q"][:, pid], pid)
        p.kick_drift(1, dt_k, dt_d, phi + e * dev_t(d["phi"]), chi + e * dev_t(d["chi"]), B + e * dev_t(d["B"]), 3, 3,
                     a_k, a_d, a_k * a_k * N, a_d * a_d * N)
        x1, q1 = by_id(p)
        return dot(lxv, x1) + dot(lqv, q1)

    # directions above are by particle ID; the container is loaded in the order of pid
    p.load(pos, q, pid)
    lx, lq = lxv.clone(), lqv.clone()
    lphi, lchi, lB = ctx.real(1), ctx.real(1), ctx.real(3)
    p.adj_gr_kick_drift(phi, chi, B, lx, lq, lphi, lchi, lB, 3, 3, a_k, a_d, dt_k, dt_d)
    ad = (dot(lx, dev_t(d["x"])) + dot(lq, dev_t(d["q"])) + dot(lphi, dev_t(d["phi"])) + dot(lchi, dev_t(d["chi"]))
          + dot(lB, dev_t(d["B"])))
    del o
    check_fd(J, ad, 1e-6, 1e-7, "kick+drift")


def test_adjoint_deposit_matches_differences(gp):
    N, n = 16, 300
    ctx = gp.Context(N, gp.GEV_FP64)
    rng = np.random.default_rng(2)
    p, pos, pid = particles(gp, ctx, rng, n)
    q = rng.normal(0, 0.02, (3, n))
    a = 0.1
    phi = dev_t(rng.normal(0, 1e-3, (1, N, N, N)))
    lT00, lTij, lL = dev_t(rng.normal(size=(1, N, N, N))), dev_t(rng.normal(size=(6, N, N, N))), \
        dev_t(rng.normal(size=(3, N, N, N)))
    d = dict(x=rng.normal(size=(3, n)) / N, q=rng.normal(0, 0.02, (3, n)), phi=rng.normal(0, 1e-3, (1, N, N, N)))
    T00, Tij, L = ctx.real(1), ctx.real(6), ctx.real(3)

    def J(e):
        p.load(pos + e * d["x"][:, pid], q + e * d["q"][:, pid], pid)
        p.project_T00_Tij_T0i_links(T00, Tij, L, a=a, phi=phi + e * dev_t(d["phi"]), coeff=1.0)
        return dot(lT00, T00) + dot(lTij, Tij) + dot(lL, L)

    p.load(pos, q, pid)
    lx, lq, lphi = ctx.real(1).new_zeros(3, n), ctx.real(1).new_zeros(3, n), ctx.real(1)
    p.adj_gr_deposit(phi, lT00, lTij * 0.25, lL, lx, lq, lphi, a, tij_scale=4.0)  # Tij's adjoint 4 (lTij / 4)
    ad = dot(lx, dev_t(d["x"])) + dot(lq, dev_t(d["q"])) + dot(lphi, dev_t(d["phi"]))
    check_fd(J, ad, 1e-6, 1e-7, "T00+Tij+T0i-link deposits")


def test_adjoint_sources_match_differences(gp):
    N = 16
    ctx = gp.Context(N, gp.GEV_FP64)
    rng = np.random.default_rng(3)
    phi, chi = dev_t(rng.normal(0, 1e-2, (1, N, N, N))), dev_t(rng.normal(0, 1e-2, (1, N, N, N)))
    T00, Tij = dev_t(rng.normal(1, 0.3, (1, N, N, N))), dev_t(rng.normal(0, 1, (6, N, N, N)))
    lsrc, lS = dev_t(rng.normal(size=(1, N, N, N))), dev_t(rng.normal(size=(6, N, N, N)))
    dphi, dchi = dev_t(rng.normal(size=(1, N, N, N))), dev_t(rng.normal(size=(1, N, N, N)))
    dT00, dTij = dev_t(rng.normal(size=(1, N, N, N))), dev_t(rng.normal(size=(6, N, N, N)))
    c1, c2, c3, bg, cc = 0.3, 0.7, 0.2, 0.31, 0.05

    def J1(e):
        r, _, _ = ctx.prepare_source_T00(phi + e * dphi, chi + e * dchi, T00 + e * dT00, bg, c1, c2, c3)
        return dot(lsrc, r)

    lT00, lphi, lchi = ctx.real(1), ctx.real(1), ctx.real(1)
    ctx.adj_source_T00(phi, T00, lsrc, lT00, lphi, lchi, bg, c1, c2, c3)
    check_fd(J1, dot(lT00, dT00) + dot(lphi, dphi) + dot(lchi, dchi), 1e-5, 1e-8, "prepareFTsource T00")

    def J2(e):
        return dot(lS, ctx.prepare_source_Tij(phi + e * dphi, Tij + e * dTij, cc))

    lphi = ctx.real(1)
    ctx.adj_source_Tij(phi, lS, lphi)
    check_fd(J2, dot(lphi, dphi) + cc * dot(lS, dTij), 1e-5, 1e-8, "prepareFTsource Tij")


@pytest.mark.parametrize("N", [16, 32])
def test_adjoint_solves_dot_products(gp, N):
    """<y, A x> = <A^T y, x> for the chi solve (its adjoint from
    gev_adj_project_ft_scalar), and the Poisson and B solves (self-adjoint)"""
    ctx = gp.Context(N, gp.GEV_FP64)
    rng = np.random.default_rng(4)
    S, y1 = dev_t(rng.normal(size=(6, N, N, N))), dev_t(rng.normal(size=(1, N, N, N)))
    lhs = dot(y1, ctx.solve_chi(S.clone()))
    rhs = dot(ctx.adj_solve_chi(y1.clone()), S)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    x1 = dev_t(rng.normal(size=(1, N, N, N)))
    for modif in (0.0, 37.0):
        lhs = dot(y1, ctx.solve_poisson(x1.clone(), 3.0, modif))
        rhs = dot(ctx.solve_poisson(y1.clone(), 3.0, modif), x1)
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    x3, y3 = dev_t(rng.normal(size=(3, N, N, N))), dev_t(rng.normal(size=(3, N, N, N)))
    lhs = dot(y3, ctx.solve_vector(x3.clone(), 0.7, 0.0))
    rhs = dot(ctx.solve_vector(y3.clone(), 0.7, 0.0), x3)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_adjoint_link_factor_matches_differences(gp):
    N = 16
    ctx = gp.Context(N, gp.GEV_FP64)
    rng = np.random.default_rng(5)
    phi, L = dev_t(rng.normal(0, 1e-2, (1, N, N, N))), dev_t(rng.normal(size=(3, N, N, N)))
    lB = dev_t(rng.normal(size=(3, N, N, N)))
    dphi, dL = dev_t(rng.normal(size=(1, N, N, N))), dev_t(rng.normal(size=(3, N, N, N)))

    def J(e):
        B = (L + e * dL).clone()
        ctx.solve_vector_T0i_links(B, phi + e * dphi, 0.7, 0.0)
        return dot(lB, B)

    u = ctx.solve_vector(lB.clone(), 0.7, 0.0)
    lphi = ctx.real(1)
    ctx.adj_link_factor(phi, L, u, lphi)
    check_fd(J, dot(u, dL) + dot(lphi, dphi), 1e-5, 1e-8, "B solve with the link factor")


def smooth(N, seed, amp):
    r = np.random.default_rng(seed)
    F = np.fft.rfftn(r.standard_normal((N, N, N)))
    kk = np.fft.fftfreq(N, 1.0 / N)
    kz, ky, kx = np.meshgrid(kk, kk, np.arange(N // 2 + 1), indexing="ij")
    F *= np.exp(-(kx ** 2 + ky ** 2 + kz ** 2) / (2 * (N / 8.0) ** 2))
    f = np.fft.irfftn(F, s=(N, N, N), axes=(0, 1, 2))
    return amp * f / f.std()


def gr_loss_grad(gp, run, psi_np, target):
    from gevpm.adjoint import pm_density_gr
    c = run.ctx
    psi = torch.from_numpy(psi_np).to(c.rdtype).to(c.device).reshape(1, run.N, run.N, run.N).requires_grad_(True)
    rho = pm_density_gr(psi, run)
    L = ((rho.double() - target) ** 2).sum()
    L.backward()
    return float(L.detach()), psi.grad.detach().double().cpu().numpy()


@pytest.mark.parametrize("N,steps", [(16, 3), (32, 4)])
def test_gr_run_gradient_matches_central_differences(gp, N, steps):
    from gevpm.adjoint import PMRunGR
    run = PMRunGR(N, 2.0 * N, 50.0, 20.0, steps, precision=gp.GEV_FP64)
    amp = 0.3 / N ** 2
    target = run.forward(torch.from_numpy(smooth(N, 12, amp)).cuda()).double().clone()
    psi = smooth(N, 11, amp)
    L, g = gr_loss_grad(gp, run, psi, target)
    for trial in range(3):
        v = smooth(N, 100 + trial, amp)
        ad = float((g.ravel() * v.ravel()).sum())
        errs = []
        for eps in (1e-7, 3e-8):
            lp = float(((run.forward(torch.from_numpy(psi + eps * v).cuda()).double() - target) ** 2).sum())
            lm = float(((run.forward(torch.from_numpy(psi - eps * v).cuda()).double() - target) ** 2).sum())
            fd = (lp - lm) / (2 * eps)
            errs.append(abs(ad - fd) / abs(fd))
            print(f"GR run N={N} steps={steps} trial {trial} eps {eps:g}: adjoint {ad:.10e} fd {fd:.10e}", flush=True)
        assert min(errs) <= 1e-5


def test_gr_run_checkpoint_interval_invariant(gp):
    from gevpm.adjoint import PMRunGR
    N, steps = 16, 5
    amp = 0.3 / N ** 2
    psi = smooth(N, 21, amp)
    r1 = PMRunGR(N, 2.0 * N, 50.0, 0.0, steps, precision=gp.GEV_FP64, ckpt_every=1)
    r3 = PMRunGR(N, 2.0 * N, 50.0, 0.0, steps, precision=gp.GEV_FP64, ckpt_every=3)
    target = torch.full((1, N, N, N), 0.31, dtype=torch.float64, device="cuda")
    L1, g1 = gr_loss_grad(gp, r1, psi, target)
    L3, g3 = gr_loss_grad(gp, r3, psi, target)
    assert abs(L1 - L3) <= 1e-12 * abs(L1)
    assert np.abs(g1 - g3).max() <= 1e-9 * np.abs(g1).max()
    assert r3.checkpoint_bytes() < r1.checkpoint_bytes()


def test_gr_run_mixed_gradient_against_fp64(gp):
    from gevpm.adjoint import PMRunGR
    N, steps = 32, 4
    amp = 0.3 / N ** 2
    psi = smooth(N, 31, amp)
    target = torch.full((1, N, N, N), 0.31, dtype=torch.float64, device="cuda")
    r64 = PMRunGR(N, 2.0 * N, 50.0, 0.0, steps, precision=gp.GEV_FP64)
    r32 = PMRunGR(N, 2.0 * N, 50.0, 0.0, steps, precision=gp.GEV_MIXED)
    _, g64 = gr_loss_grad(gp, r64, psi, target)
    _, g32 = gr_loss_grad(gp, r32, psi, target)
    rel = float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64))
    print("GR mixed vs FP64 gradient rel L2", rel, flush=True)
    assert rel < 1e-3


def test_mixed_reverse_kernels_against_fp64(gp):
    """the FP32 reverse kick/drift and reverse deposit against their FP64 forms
    on the same particles, fields and adjoints (16^3 with crowded cells, so
    the chunked lane order and the wide tiles are both exercised): relative
    L2 of every output within 2e-6 (FP32 rounding gives about 1e-7)"""
    N, n = 16, 6000
    rng = np.random.default_rng(7)
    c = rng.integers(0, N, size=(3, n))
    c[:, :1500] = rng.integers(5, 7, size=(3, 1500))  # crowded cells
    pos = (c + rng.uniform(0.05, 0.95, size=(3, n))) / N
    q = rng.normal(0, 0.02, (3, n))
    pid = rng.permutation(n).astype(np.int64)
    fl = dict(phi=rng.normal(0, 1e-3, (1, N, N, N)), chi=rng.normal(0, 1e-3, (1, N, N, N)),
              B=rng.normal(0, 1e-5, (3, N, N, N)), lT00=rng.normal(size=(1, N, N, N)),
              lTij=rng.normal(size=(6, N, N, N)), lL=rng.normal(size=(3, N, N, N)))
    lx0, lq0 = rng.normal(size=(3, n)), rng.normal(size=(3, n))
    out = {}
    for prec in (gp.GEV_FP64, gp.GEV_MIXED):
        ctx = gp.Context(N, prec)
        dt = ctx.rdtype
        p = gp.Particles(ctx, n, 0.31 / N ** 3)
        p.load(pos, q, pid)
        f = {k: dev_t(v, dt) for k, v in fl.items()}
        lx, lq = dev_t(lx0, dt), dev_t(lq0, dt)
        lphi, lchi, lB = ctx.real(1), ctx.real(1), ctx.real(3)
        p.adj_gr_kick_drift(f["phi"], f["chi"], f["B"], lx, lq, lphi, lchi, lB, 3, 3, 0.1, 0.102, 0.015, 0.02)
        gx, gq, gphi = ctx.real(1).new_zeros(3, n), ctx.real(1).new_zeros(3, n), ctx.real(1)
        p.adj_gr_deposit(f["phi"], f["lT00"], f["lTij"], f["lL"], gx, gq, gphi, 0.1, tij_scale=0.7)
        out[prec] = dict(lx=lx, lq=lq, lphi=lphi, lchi=lchi, lB=lB, gx=gx, gq=gq, gphi=gphi)
    for k in out[gp.GEV_FP64]:
        a, b = out[gp.GEV_FP64][k].double(), out[gp.GEV_MIXED][k].double()
        rel = float(torch.linalg.norm(a - b) / torch.linalg.norm(a))
        print(f"{k}: mixed against FP64 rel L2 {rel:.2e}", flush=True)
        assert rel <= 2e-6, k
