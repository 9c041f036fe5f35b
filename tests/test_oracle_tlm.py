"""Pins of the oracle's tangent-linear Newtonian run (oracle/gev_oracle_tlm.c),
the forward-mode derivative the GPU adjoint is checked against: central
differences of the oracle's own deterministic forward, per kernel and for
whole runs."""
import numpy as np
import pytest


def smooth(N, seed, amp):
    r = np.random.default_rng(seed)
    F = np.fft.rfftn(r.standard_normal((N, N, N)))
    kk = np.fft.fftfreq(N, 1.0 / N)
    kz, ky, kx = np.meshgrid(kk, kk, np.arange(N // 2 + 1), indexing="ij")
    F *= np.exp(-(kx ** 2 + ky ** 2 + kz ** 2) / (2 * (N / 8.0) ** 2))
    f = np.fft.irfftn(F, s=(N, N, N), axes=(0, 1, 2))
    return amp * f / f.std()


def away_from_faces(rng, N, n, margin=0.05):
    c = rng.integers(0, N, size=(3, n))
    r = rng.uniform(margin, 1 - margin, size=(3, n))
    return (c + r) / N


def test_tangent_deposit_matches_differences(O):
    N, n = 8, 40
    rng = np.random.default_rng(3)
    pos = away_from_faces(rng, N, n)
    dpos = rng.standard_normal((3, n))
    h = 1e-6 / N
    q = np.zeros((3, n))

    def rho(p):
        return O.project_T00(N, np.ascontiguousarray(p), q, 0.7 / N ** 3)

    fd = (rho(pos + h * dpos) - rho(pos - h * dpos)) / (2 * h)
    got = np.zeros(N ** 3)
    O.lib().or_tlm_deposit_mass(N, n, np.ascontiguousarray(pos), np.ascontiguousarray(dpos), 0.7, got)
    assert np.allclose(got, fd, rtol=1e-7, atol=1e-7 * np.abs(fd).max())
    assert abs(got.sum()) < 1e-12 * np.abs(got).max() * n  # mass is conserved to first order


@pytest.mark.parametrize("part", ["position", "potential"])
def test_tangent_kick_matches_differences(O, part):
    N, n, dtau, a = 8, 30, 0.03, 0.4
    rng = np.random.default_rng(4)
    pos = away_from_faces(rng, N, n)
    phi = rng.standard_normal(N ** 3)
    dphi = rng.standard_normal(N ** 3)
    dpos = rng.standard_normal((3, n)) if part == "position" else np.zeros((3, n))
    if part == "potential":
        phi_of = lambda e: phi + e * dphi  # noqa: E731
    else:
        phi_of = lambda e: phi  # noqa: E731
    h = 1e-6 / N

    def kick(e):
        q = np.zeros((3, n))
        O.update_q_newton(N, np.ascontiguousarray(pos + e * dpos), q, dtau, phi_of(e), a)
        return q

    fd = (kick(h) - kick(-h)) / (2 * h)
    dq = np.zeros((3, n))
    O.lib().or_tlm_kick_newton(N, n, np.ascontiguousarray(pos), np.ascontiguousarray(dpos), dq, phi,
                               dphi if part == "potential" else np.zeros(N ** 3), dtau * a * N)
    assert np.allclose(dq, fd, rtol=1e-6, atol=1e-7 * np.abs(fd).max())


@pytest.mark.parametrize("N,steps", [(8, 3), (16, 4)])
def test_tangent_run_matches_central_differences(O, N, steps):
    """dL of the whole run (IC map, cycles, final deposit, loss) against
    central differences of the same deterministic forward"""
    sched, fpg, cx, cq = O.newton_schedule(N, 2.0 * N, 50.0, 0.0, steps)
    amp = 0.3 / N ** 2
    psi, tgt = smooth(N, 11, amp), smooth(N, 12, amp)
    _, _, target, _, _ = O.newton_run_tlm(N, tgt, np.zeros(N ** 3), sched, fpg, cx, cq, want_state=True)
    for trial in range(3):
        v = smooth(N, 100 + trial, amp)
        L, dL = O.newton_run_tlm(N, psi, target, sched, fpg, cx, cq, v=v)
        # the particles move 3-8 cells from z = 50 to 0 (0.26 cells rms at
        # the start), so a perturbation of 1e-6 of psi already carries some
        # of them across a cell face during the run, where
        # the CIC loss has a kink; 1e-7 and 3e-8 stay on the smooth pieces
        # (and above FP64 rounding of L)
        errs = []
        for eps in (1e-7, 3e-8):
            Lp, _ = O.newton_run_tlm(N, psi + eps * v, target, sched, fpg, cx, cq)
            Lm, _ = O.newton_run_tlm(N, psi - eps * v, target, sched, fpg, cx, cq)
            fd = (Lp - Lm) / (2 * eps)
            errs.append(abs(dL - fd) / abs(fd))
            print(f"N={N} trial {trial} eps {eps:g}: L {L:.6e} dL {dL:.10e} fd {fd:.10e}")
        assert min(errs) <= 2e-6
