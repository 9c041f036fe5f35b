"""Adjoint of the Newtonian particle-mesh run (gevpm.adjoint): the gradient of a
field-level loss with respect to the initial displacement potential, from the
hand-written reverse kernels, against central finite differences of the
forward run (FP64), across checkpoint intervals, and FP32 against FP64."""
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


def smooth_field(N, seed, amp, kmax=4.0):
    """real Gaussian random field with power only at |k| < kmax (mesh units)"""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((N, N, N))
    F = np.fft.rfftn(w)
    k = np.fft.fftfreq(N, 1.0 / N)
    kz, ky, kx = np.meshgrid(k, k, np.arange(N // 2 + 1), indexing="ij")
    F *= np.exp(-(kx ** 2 + ky ** 2 + kz ** 2) / (2 * kmax ** 2))
    f = np.fft.irfftn(F, s=(N, N, N))
    return amp * f / np.abs(f).max()


def make_run(gp, N, steps, prec, ckpt=1):
    from gevpm.adjoint import PMRun
    return PMRun(N, boxsize=2.0 * N, z_in=50.0, z_end=0.0, steps=steps,
                 precision=gp.GEV_FP64 if prec == "fp64" else gp.GEV_MIXED, ckpt_every=ckpt)


def loss_and_grad(run, psi_np, target):
    from gevpm.adjoint import pm_density
    c = run.ctx
    psi = torch.from_numpy(psi_np).to(c.rdtype).to(c.device).reshape(1, run.N, run.N, run.N).requires_grad_(True)
    rho = pm_density(psi, run)
    L = ((rho.double() - target) ** 2).sum()
    L.backward()
    return float(L), psi.grad.detach().double().cpu().numpy().reshape(run.N, run.N, run.N)


def loss_only(run, psi_np, target):
    c = run.ctx
    psi = torch.from_numpy(psi_np).to(c.rdtype).to(c.device).reshape(1, run.N, run.N, run.N)
    rho = run.forward(psi)
    return float(((rho.double() - target) ** 2).sum())


@pytest.mark.parametrize("N,steps", [(16, 4), (32, 3)])
def test_adjoint_matches_central_differences(gp, N, steps):
    run = make_run(gp, N, steps, "fp64")
    # displacements of a few tenths of a cell: psi amplitude ~ 0.05 cells^2 in box units^2
    psi = smooth_field(N, 1, 0.05 / N ** 2)
    tgt_run = make_run(gp, N, steps, "fp64")
    target = tgt_run.forward(torch.from_numpy(smooth_field(N, 2, 0.05 / N ** 2)).to(tgt_run.ctx.device)).double()
    L, g = loss_and_grad(run, psi, target)
    assert L > 0 and np.isfinite(g).all()
    rng = np.random.default_rng(3)
    for trial in range(4):
        d = smooth_field(N, 10 + trial, 1.0 / N ** 2)
        eps = 1e-7  # max change of a displacement ~1e-7 cells: almost no particle crosses a cell face
        Lp, Lm = loss_only(run, psi + eps * d, target), loss_only(run, psi - eps * d, target)
        fd = (Lp - Lm) / (2 * eps)
        ad = float((g * d).sum())
        print(f"N={N} trial {trial}: adjoint {ad:.10e} fd {fd:.10e}", flush=True)
        assert abs(ad - fd) <= 1e-5 * abs(fd), (ad, fd)
    del rng


def test_adjoint_checkpoint_interval_invariant(gp):
    N, steps = 16, 5
    psi = smooth_field(N, 4, 0.05 / N ** 2)
    r1, r3 = make_run(gp, N, steps, "fp64", 1), make_run(gp, N, steps, "fp64", 3)
    target = torch.zeros(1, N, N, N, dtype=torch.float64, device=r1.ctx.device) + 0.31
    L1, g1 = loss_and_grad(r1, psi, target)
    L3, g3 = loss_and_grad(r3, psi, target)
    assert abs(L1 - L3) <= 1e-12 * abs(L1)
    assert np.abs(g1 - g3).max() <= 1e-10 * np.abs(g1).max()
    assert r3.checkpoint_bytes() < r1.checkpoint_bytes()


def test_adjoint_fp32_against_fp64(gp):
    N, steps = 32, 3
    psi = smooth_field(N, 5, 0.05 / N ** 2)
    r64, r32 = make_run(gp, N, steps, "fp64"), make_run(gp, N, steps, "mixed")
    target = torch.zeros(1, N, N, N, dtype=torch.float64, device=r64.ctx.device) + 0.31
    _, g64 = loss_and_grad(r64, psi, target)
    _, g32 = loss_and_grad(r32, psi, target)
    rel = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print("FP32 vs FP64 gradient rel L2", rel, flush=True)
    assert rel < 1e-3
