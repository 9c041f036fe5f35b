"""The GPU adjoint of the Newtonian run against the oracle's forward-mode
tangent (oracle/gev_oracle_tlm.c) on the same initial potential: the
dot-product test <grad L, v> = dL(v), which holds to rounding whatever the
number of cell-face crossings, at 32^3 and at the 128^3 / 20-cycle
differentiable workload (L = 256 Mpc/h, z = 50 -> 0, FP64 forward)."""
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


def smooth(N, seed, amp, width=16.0):
    r = np.random.default_rng(seed)
    F = np.fft.rfftn(r.standard_normal((N, N, N)))
    kk = np.fft.fftfreq(N, 1.0 / N)
    kz, ky, kx = np.meshgrid(kk, kk, np.arange(N // 2 + 1), indexing="ij")
    F *= np.exp(-(kx ** 2 + ky ** 2 + kz ** 2) / (2 * (N / width) ** 2))
    f = np.fft.irfftn(F, s=(N, N, N), axes=(0, 1, 2))
    return amp * f / f.std()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("N,steps", [(32, 8), (128, 20)])
def test_adjoint_dot_product_against_oracle_tangent(gp, O, N, steps):
    from gevpm.adjoint import PMRun, pm_density
    print(f"adjoint dot-product test N={N} steps={steps}", flush=True)
    run = PMRun(N, 2.0 * N, 50.0, 0.0, steps, precision=gp.GEV_FP64, ckpt_every=1)
    dev = run.ctx.device
    amp = 0.3 / N ** 2
    psi_np, tgt_np = smooth(N, 11, amp), smooth(N, 12, amp)
    target = run.forward(torch.from_numpy(tgt_np).to(dev)).double().clone()
    psi = torch.from_numpy(psi_np).to(dev).reshape(1, N, N, N).requires_grad_(True)
    rho = pm_density(psi, run)
    L = ((rho.double() - target) ** 2).sum()
    L.backward()
    g = psi.grad.detach().double().cpu().numpy().ravel()
    sched = np.array([[s["a"], s["a_d"], s["dtau_k"], s["dtau"]] for s in run.sched])
    tgt = target.cpu().numpy().ravel()
    for trial in range(2):
        v = smooth(N, 100 + trial, amp)
        Lo, dL = O.newton_run_tlm(N, psi_np, tgt, sched, run.fpg, run.cx, run.cq, v=v)
        ad = float(np.dot(g, v.ravel()))
        print(f"  N={N} trial {trial}: L gpu {float(L):.12e} oracle {Lo:.12e}; "
              f"<grad, v> {ad:.12e} tangent {dL:.12e} rel {abs(ad - dL) / abs(dL):.2e}", flush=True)
        assert abs(float(L) - Lo) <= 1e-11 * Lo          # same forward (summation order only)
        assert abs(ad - dL) <= 1e-8 * abs(dL)
