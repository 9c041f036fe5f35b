"""Argument checks of the Python layer on the GPU, and the maxvel reduction of
the tiled kick over real particles only."""
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


def test_fields_of_another_precision_size_or_layout_are_refused(gp):
    N = 16
    cm = gp.Context(N, gp.GEV_MIXED)
    c64 = gp.Context(N, gp.GEV_FP64)
    c32 = gp.Context(32, gp.GEV_MIXED)
    p = gp.Particles(cm, 10, 1.0)
    rng = np.random.default_rng(1)
    p.load(rng.uniform(0, 1, (3, 10)), np.zeros((3, 10)), np.arange(10))
    with pytest.raises(gp.GevError, match="float32"):
        p.project_mass(c64.real(1))                      # FP64 field into a mixed context
    with pytest.raises(gp.GevError, match="elements"):
        p.project_mass(c32.real(1))                      # field of another N
    with pytest.raises(gp.GevError, match="elements"):
        p.project_Tij(cm.real(3))                        # too few components
    with pytest.raises(gp.GevError, match="contiguous"):
        p.project_mass(cm.real(1).transpose(2, 3))       # N^3 elements, not contiguous
    with pytest.raises(gp.GevError, match="complex64"):
        cm.poisson(c64.fourier(1), 1.0)
    with pytest.raises(gp.GevError, match="elements"):
        cm.solve_vector(cm.real(1))
    with pytest.raises(gp.GevError, match="multiple"):
        cm.fft_r2c(torch.zeros(N ** 3 + 1, device="cuda", dtype=torch.float32))
    with pytest.raises(gp.GevError, match="load"):
        p.load(np.zeros((3, 5)), np.zeros((3, 4)), np.arange(5))
    with pytest.raises(gp.GevError, match="load"):
        p.load(torch.zeros(3, 5, device="cuda"), torch.zeros(3, 5, device="cuda"), torch.arange(6, device="cuda"))
    with pytest.raises(gp.GevError, match="smaller"):
        p.restore(torch.zeros(8, dtype=torch.uint8, device="cuda"))
    # the container is untouched by the refused calls
    rho = cm.real(1)
    p.project_mass(rho)
    assert abs(rho.double().sum().item() - 10 * N ** 3) < 1e-3 * 10 * N ** 3


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_maxvel_over_real_particles_only(gp, O, prec):
    """a potential whose only gradient sits at the centre of each tile's first
    cell, and particles at rest where it is flat: the kick leaves them at
    rest, so maxvel is exactly 0 (the idle lanes of the tiled kernel update a
    stand-in particle there, whose value must not enter the maximum)"""
    N = 32
    ctx = gp.Context(N, gp.GEV_FP64 if prec == "fp64" else gp.GEV_MIXED)
    phi = np.zeros((N, N, N))
    for x0 in range(0, N, 8):
        for y0 in range(0, N, 8):
            for z0 in range(0, N, 4):
                phi[z0, y0, (x0 + 1) % N] = 1.0
                phi[z0, y0, (x0 - 1) % N] = -1.0
    pos = np.array([[4.5, 4.5, 2.5], [20.3, 12.6, 9.4], [12.5, 27.5, 17.5]]).T / N
    pid = np.arange(3)
    p = gp.Particles(ctx, 3, 1.0)
    p.load(pos, np.zeros((3, 3)), pid)
    phit = torch.from_numpy(phi.reshape(1, N, N, N)).to(ctx.rdtype).cuda()
    q0 = np.zeros((3, 3))
    ref = O.update_q_newton(N, pos.copy(), q0, 0.01, phi.ravel(), 0.5)
    assert ref == 0.0 and np.all(q0 == 0.0)
    assert p.update_q_newton(0.01, phit, 0.5) == 0.0
    assert p.kick_drift(0, 0.01, 0.0, phit, None, None, 1, 0, 0.5, 0.5, 1.0, 1.0) == 0.0
    zero = torch.zeros_like(phit)
    B = torch.zeros(3, N, N, N, dtype=ctx.rdtype, device="cuda")
    assert p.update_q(0.01, phit, zero, B, 3, 0.5, 0.25 * N) == 0.0


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_checkpoint_round_trip(gp, prec):
    """save, replace the particles by more of them, restore: the container
    holds the saved particles bit for bit again (count, order, cell starts);
    a buffer off 16-byte alignment is refused before anything is copied"""
    N = 16
    ctx = gp.Context(N, gp.GEV_FP64 if prec == "fp64" else gp.GEV_MIXED)
    rng = np.random.default_rng(5)
    n = 1001  # odd: the 4-byte arrays end off a 16-byte boundary
    p = gp.Particles(ctx, 2 * n, 0.31 / N ** 3)
    p.load(rng.uniform(0, 1, (3, n)), rng.normal(0, 0.01, (3, n)), rng.permutation(n).astype(np.int64))
    ref = [t.clone() for t in p.store()]
    rho_ref = ctx.real(1)
    p.project_mass(rho_ref)
    buf = p.save()
    p.load(rng.uniform(0, 1, (3, 2 * n)), rng.normal(0, 0.01, (3, 2 * n)), np.arange(2 * n, dtype=np.int64))
    p.restore(buf)
    assert p.n == n
    for a, b in zip(ref, p.store()):
        assert torch.equal(a, b)
    rho = ctx.real(1)
    p.project_mass(rho)  # the cell starts find every particle again (sums in atomic order: not bitwise)
    assert float((rho.double() - rho_ref.double()).abs().max()) <= 1e-5 * float(rho_ref.double().abs().max())
    big = torch.empty(buf.numel() + 16, dtype=torch.uint8, device="cuda")
    with pytest.raises(gp.GevError, match="aligned"):
        p.save(big[4:4 + buf.numel()])
    with pytest.raises(gp.GevError, match="smaller"):
        p.restore(buf[:buf.numel() - 256])  # the saved count needs all of it
    with pytest.raises(gp.GevError, match="elements"):
        p.save(buf[:1024])  # refused by the Python layer already
    assert p.n == n
