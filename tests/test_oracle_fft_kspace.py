"""Pins the oracle's FFT and k-space solvers (CPU only).

Each check is fixed by mathematics independent of the oracle's code: a
brute-force DFT, numpy.fft golden fixtures, Parseval, the real-space 7-point
Laplacian, and discrete longitudinal / transverse constructions.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def lap(f, N):
    """7-point Laplacian in box units (dx = 1/N) of a [z,y,x] array."""
    out = -6.0 * f
    for ax in range(3):
        out = out + np.roll(f, 1, ax) + np.roll(f, -1, ax)
    return out * N * N


def dplus(f, axis, N):
    """forward difference along lattice axis (0=x,1=y,2=z) of a [z,y,x] array"""
    ax = 2 - axis
    return (np.roll(f, -1, ax) - f) * N


def dminus(f, axis, N):
    ax = 2 - axis
    return (f - np.roll(f, 1, ax)) * N


@pytest.mark.parametrize("N", [4, 8])
def test_fft_r2c_bruteforce(O, N):
    rng = np.random.default_rng(N)
    f = rng.standard_normal((N, N, N))
    F = O.fft_r2c(f.ravel())
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    for kz in range(N):
        for ky in range(N):
            for kx in range(N // 2 + 1):
                ph = np.exp(-2j * np.pi * (kx * x + ky * y + kz * z) / N)
                ref = np.sum(f * ph)
                assert abs(F[kz, ky, kx] - ref) <= 1e-12 * (1 + abs(ref))


@pytest.mark.parametrize("N", [8, 16])
def test_fft_golden_numpy(O, N):
    d = np.load(os.path.join(GOLD, f"fft_N{N}.npz"))
    F = O.fft_r2c(d["f"].ravel())
    assert np.abs(F - d["F"]).max() <= 1e-13 * np.abs(d["F"]).max()
    g = O.fft_c2r(d["F"])
    assert np.abs(g / N ** 3 - d["f"].ravel()).max() <= 1e-13


def test_fft_parseval_roundtrip(O):
    N = 16
    rng = np.random.default_rng(3)
    f = rng.standard_normal(N ** 3)
    F = O.fft_r2c(f)
    w = np.full(N // 2 + 1, 2.0)
    w[0] = w[N // 2] = 1.0
    assert np.isclose(np.sum(w * np.abs(F) ** 2) / N ** 3, np.sum(f * f), rtol=1e-12)
    assert np.allclose(O.fft_c2r(F) / N ** 3, f, atol=1e-13)


def test_c2r_ignores_imaginary_dc_nyquist(O):
    N = 8
    F = np.zeros((N, N, N // 2 + 1), np.complex128)
    F[0, 0, 0] = 1.0 + 5.0j
    F[0, 0, N // 2] = 2.0 - 3.0j
    f = O.fft_c2r(F).reshape(N, N, N)
    x = np.arange(N)
    assert np.allclose(f, (1.0 + 2.0 * (-1.0) ** x)[None, None, :], atol=1e-14)


@pytest.mark.parametrize("modif", [0.0, 37.5])
def test_poisson_inverts_7pt_laplacian(O, modif):
    """(Lap - modif) phi = coeff s (zero-mean part when modif == 0): Poisson with the physical sign."""
    N = 16
    rng = np.random.default_rng(5)
    s = rng.standard_normal((N, N, N))
    coeff = 0.37
    F = O.fft_r2c(s.ravel())
    phi = O.fft_c2r(O.poisson(F, coeff, modif)).reshape(N, N, N)
    lhs = lap(phi, N) - modif * phi
    rhs = coeff * (s - s.mean()) if modif == 0.0 else coeff * s
    assert np.abs(lhs - rhs).max() <= 1e-10 * np.abs(rhs).max()
    if modif == 0.0:
        assert abs(phi.mean()) < 1e-14


def test_project_ft_scalar_longitudinal_and_trace(O):
    """S_ij = D+_i D+_j psi (i != j), D+_i D-_i psi (i == j)  ->  chi = N^2 (psi - <psi>);
    a pure-trace S gives chi = 0."""
    N = 16
    rng = np.random.default_rng(7)
    psi = rng.standard_normal((N, N, N))
    S = np.zeros((6, N, N, N))
    comps = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 4, (2, 2): 5}
    for (i, j), c in comps.items():
        S[c] = dplus(dminus(psi, i, N), i, N) if i == j else dplus(dplus(psi, i, N), j, N)
    SF = np.stack([O.fft_r2c(S[c].ravel()) for c in range(6)])
    chi = O.fft_c2r(O.project_ft_scalar(SF)).reshape(N, N, N)
    ref = N * N * (psi - psi.mean())
    assert np.abs(chi - ref).max() <= 1e-9 * np.abs(ref).max()
    T = np.zeros((6, N, N, N))
    f = rng.standard_normal((N, N, N))
    T[0] = T[3] = T[5] = f
    TF = np.stack([O.fft_r2c(T[c].ravel()) for c in range(6)])
    chi0 = O.fft_c2r(O.project_ft_scalar(TF))
    assert np.abs(chi0).max() <= 1e-10 * np.abs(f).max()
    # add mode accumulates
    base = O.project_ft_scalar(SF)
    twice = O.project_ft_scalar(SF, base)
    assert np.allclose(twice, 2 * base, rtol=1e-13, atol=1e-13)


def test_project_ft_vector_transverse(O):
    """longitudinal input -> 0; input transverse under D- -> (-Lap + modif) B = 4 coeff N^3 T
    (no 1/N^3 here, unlike solveModifiedPoissonFT)."""
    N = 16
    rng = np.random.default_rng(9)
    psi = rng.standard_normal((N, N, N))
    L = np.stack([dplus(psi, i, N) for i in range(3)])
    LF = np.stack([O.fft_r2c(L[i].ravel()) for i in range(3)])
    B = O.project_ft_vector(LF, 0.7, 0.0)
    assert np.abs(B).max() <= 1e-10 * np.abs(LF).max()
    A = rng.standard_normal((3, N, N, N))
    T = np.stack([dminus(A[2], 1, N) - dminus(A[1], 2, N),
                  dminus(A[0], 2, N) - dminus(A[2], 0, N),
                  dminus(A[1], 0, N) - dminus(A[0], 1, N)])
    div = sum(dminus(T[i], i, N) for i in range(3))
    assert np.abs(div).max() < 1e-9 * np.abs(T).max()
    for modif in (0.0, 11.0):
        TF = np.stack([O.fft_r2c(T[i].ravel()) for i in range(3)])
        BF = O.project_ft_vector(TF, 0.7, modif)
        Br = np.stack([O.fft_c2r(BF[i]).reshape(N, N, N) for i in range(3)])
        for i in range(3):
            lhs = -lap(Br[i], N) + modif * Br[i]
            assert np.abs(lhs - 4 * 0.7 * N ** 3 * T[i]).max() <= 1e-9 * N ** 3 * np.abs(T).max()
        # transversality of the output under D-
        divB = sum(dminus(Br[i], i, N) for i in range(3))
        assert np.abs(divB).max() <= 1e-9 * np.abs(Br).max() * N


def test_evolve_ft_vector_structure(O):
    """pure trace and longitudinal tensors leave B unchanged; the increment is D- transverse.
    (parity partially pinned: no closed form for the transverse amplitude)"""
    N = 8
    rng = np.random.default_rng(13)
    psi = rng.standard_normal((N, N, N))
    S = np.zeros((6, N, N, N))
    comps = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 4, (2, 2): 5}
    for (i, j), c in comps.items():
        S[c] = dplus(dminus(psi, i, N), i, N) if i == j else dplus(dplus(psi, i, N), j, N)
    f = rng.standard_normal((N, N, N))
    S[0] += f; S[3] += f; S[5] += f
    SF = np.stack([O.fft_r2c(S[c].ravel()) for c in range(6)])
    B0 = np.zeros((3, N, N, N // 2 + 1), np.complex128)
    B0[:, 1:, 1:, 1:3] = rng.standard_normal((3, N - 1, N - 1, 2))
    B1 = O.evolve_ft_vector(SF, B0, 0.3)
    mask = np.ones_like(B0, bool)
    mask[:, 0, 0, 0] = False
    assert np.abs((B1 - B0)[mask]).max() <= 1e-10 * np.abs(SF).max()
    R = rng.standard_normal((6, N, N, N))
    RF = np.stack([O.fft_r2c(R[c].ravel()) for c in range(6)])
    inc = O.evolve_ft_vector(RF, np.zeros_like(B0), 0.3)
    Br = np.stack([O.fft_c2r(inc[i]).reshape(N, N, N) for i in range(3)])
    divB = sum(dminus(Br[i], i, N) for i in range(3))
    assert np.abs(divB).max() <= 1e-9 * np.abs(Br).max() * N
    assert np.abs(Br).max() > 0


@pytest.mark.parametrize("mode", [(3, 1, 2), (0, 5, 0), (8, 0, 0)])
def test_power_spectrum_single_mode(O, mode):
    """extractPowerSpectrum semantics on one real Fourier mode: all power in the
    bin of |k|, equal to (weight |F|^2 |k|^3 / sinc^2) / occupation, where the
    occupation and mean |k| of every bin are brute-force counts over the full
    (unfolded) N^3 frequency grid."""
    N, nb, A = 16, 12, 0.7
    nx, ny, nz = mode
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    f = A * np.cos(2 * np.pi * (nx * x + ny * y + nz * z) / N)
    F = O.fft_r2c(f.ravel())
    kb, pw, ks, ps, occ = O.power_spectrum(F, nb, deconvolve=True)
    # full-grid brute force of the bins
    fr = np.fft.fftfreq(N, 1.0 / N)
    KZ, KY, KX = np.meshgrid(fr, fr, fr, indexing="ij")
    kk = 2 * np.pi * np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    kmax = np.sqrt(3.0) * 2 * np.pi * (N // 2)
    b = np.floor(nb * kk / kmax).astype(int)
    b[0, 0, 0] = -1
    for j in range(nb):
        sel = b == j
        assert occ[j] == sel.sum()
        if sel.sum():
            assert abs(kb[j] - kk[sel].mean()) <= 1e-12 * kk[sel].mean()
    k = 2 * np.pi * np.sqrt(nx * nx + ny * ny + nz * nz)
    jb = int(np.floor(nb * k / kmax))
    sinc = np.prod([np.sin(np.pi * n / N) * N / (np.pi * n) if n else 1.0 for n in mode])
    # a cosine puts |F| = A N^3 / 2 at +n and -n (both in the full grid: 2 entries,
    # or one self-conjugate entry of amplitude A N^3 at a Nyquist/zero plane)
    selfconj = all(n in (0, N // 2) for n in mode)
    full = (A * N ** 3) ** 2 if selfconj else 2 * (A * N ** 3 / 2) ** 2
    expect = full * k ** 3 / sinc ** 2 / occ[jb]
    assert abs(pw[jb] - expect) <= 1e-10 * expect
    others = np.delete(pw, jb)
    assert np.abs(others).max() <= 1e-20 * expect


def _rand_ft(O, rng, N, ncomp):
    return np.stack([O.fft_r2c(rng.standard_normal(N ** 3)).reshape(N, N, N // 2 + 1) for _ in range(ncomp)])


@pytest.mark.parametrize("deconv", [True, False])
def test_cross_spectrum_pins(O, deconv):
    """extractCrossSpectrum (tools.hpp:65-330): the cross spectrum of a field with
    itself is its power spectrum; power is linear in Re(F conj G), so
    cross(F, G) = (P(F + G) - P(F) - P(G)) / 2 bin by bin (polarization), with
    the same bins and occupation; comp1/comp2 pick one component of each
    field."""
    N, nb = 16, 10
    rng = np.random.default_rng(7)
    F, G = _rand_ft(O, rng, N, 1)[0], _rand_ft(O, rng, N, 1)[0]
    a = O.power_spectrum(F, nb, deconvolve=deconv)
    b = O.cross_spectrum(F, F, nb, deconvolve=deconv)
    for x, y in zip(a, b):
        assert np.allclose(x, y, rtol=1e-13, atol=0)
    c = O.cross_spectrum(F, G, nb, deconvolve=deconv)
    s = O.power_spectrum(F + G, nb, deconvolve=deconv)
    assert np.array_equal(c[4], s[4]) and np.allclose(c[0], s[0], rtol=1e-14)
    pol = 0.5 * (s[1] - a[1] - O.power_spectrum(G, nb, deconvolve=deconv)[1])
    assert np.allclose(c[1], pol, rtol=1e-10, atol=1e-10 * np.abs(s[1]).max())
    # a component pair of two 3-component fields
    F3, G3 = _rand_ft(O, rng, N, 3), _rand_ft(O, rng, N, 3)
    d = O.cross_spectrum(F3, G3, nb, deconvolve=deconv, comp1=1, comp2=2)
    e = O.cross_spectrum(F3[1], G3[2], nb, deconvolve=deconv)
    for x, y in zip(d, e):
        assert np.allclose(x, y, rtol=1e-13, atol=0)
    # a vector field sums its components
    v = O.cross_spectrum(F3, F3, nb, deconvolve=deconv)
    w = sum(np.asarray(O.power_spectrum(F3[i], nb, deconvolve=deconv)[1]) for i in range(3))
    assert np.allclose(v[1], w, rtol=1e-12)


def test_cross_spectrum_symmetric_tensor(O):
    """A symmetric tensor (xx xy xz yy yz zz) contracts as the full 3x3 tensor:
    the same spectrum as the 9 components summed without symmetry."""
    N, nb = 8, 6
    rng = np.random.default_rng(11)
    S1, S2 = _rand_ft(O, rng, N, 6), _rand_ft(O, rng, N, 6)
    idx = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
    full1 = np.stack([S1[idx[a][b]] for a in range(3) for b in range(3)])
    full2 = np.stack([S2[idx[a][b]] for a in range(3) for b in range(3)])
    s = O.cross_spectrum(S1, S2, nb, symmetric=True)
    f = O.cross_spectrum(full1, full2, nb)
    for x, y in zip(s, f):
        assert np.allclose(x, y, rtol=1e-12, atol=0)


@pytest.mark.parametrize("mode", [(3, 1, 2), (0, 5, 0), (8, 0, 0)])
def test_cross_spectrum_grid_ktype(O, mode):
    """KTYPE_GRID: |k|^2 = sum (2N sin(pi n / N))^2 (tools.hpp:96-103) and
    kmax^2 = 3 (2N)^2; one real mode's power lands in the bin of that |k| with
    weight |F|^2 |k|^3 / sinc^2 / occupation, the occupations and mean |k|
    counted by brute force over the full frequency grid."""
    N, nb, A = 16, 12, 0.7
    nx, ny, nz = mode
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    f = A * np.cos(2 * np.pi * (nx * x + ny * y + nz * z) / N)
    F = O.fft_r2c(f.ravel())
    kb, pw, ks, ps, occ = O.cross_spectrum(F, F, nb, deconvolve=True, ktype=0)
    idx = np.arange(N)
    g = (2 * N * np.sin(np.pi * idx / N)) ** 2
    KZ, KY, KX = np.meshgrid(idx, idx, idx, indexing="ij")
    kk = np.sqrt(g[KX] + g[KY] + g[KZ])
    kmax = np.sqrt(3.0) * 2 * N
    b = np.floor(nb * kk / kmax).astype(int)
    b[0, 0, 0] = -1
    for j in range(nb):
        sel = b == j
        assert occ[j] == sel.sum()
        if sel.sum():
            assert abs(kb[j] - kk[sel].mean()) <= 1e-12 * kk[sel].mean()
    k = np.sqrt(g[nx] + g[ny] + g[nz])
    jb = int(np.floor(nb * k / kmax))
    sinc = np.prod([np.sin(np.pi * n / N) * N / (np.pi * n) if n else 1.0 for n in mode])
    selfconj = all(n in (0, N // 2) for n in mode)
    full = (A * N ** 3) ** 2 if selfconj else 2 * (A * N ** 3 / 2) ** 2
    expect = full * k ** 3 / sinc ** 2 / occ[jb]
    assert abs(pw[jb] - expect) <= 1e-10 * expect
    assert np.abs(np.delete(pw, jb)).max() <= 1e-20 * expect
