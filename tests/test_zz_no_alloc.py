"""gev_power_spectrum_noalloc (static device workspace, no allocation) against
gev_power_spectrum (per-call workspace) through the C ABI: two calls in a row
on one device, the second on a smaller mesh with fewer bins, and a call whose
partials need the narrowed column grid (1024 bins)."""
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


def _spectra(gp, ctx, seed, numbins, deconvolve):
    from gevpm.spectra import power_spectrum
    N = ctx.N
    rng = np.random.default_rng(seed)
    f = torch.from_numpy(rng.standard_normal((1, N, N, N))).to(ctx.rdtype).cuda()
    F = ctx.fft_r2c(f)
    ref = power_spectrum(ctx, F, numbins, deconvolve)
    got = power_spectrum(ctx, F, numbins, deconvolve, noalloc=True)
    return ref, got


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_power_spectrum_noalloc_matches_allocating_call(gp, prec):
    p = gp.GEV_FP64 if prec == "fp64" else gp.GEV_MIXED
    # (mesh, bins, deconvolve): the first needs the narrowed grid, the second
    # is smaller than the first and runs right after it on the same workspace
    for N, nb, dec, seed in ((128, 1024, True, 1), (32, 40, False, 2), (64, 64, True, 3)):
        print(f"noalloc P(k): {prec} N={N} numbins={nb} deconvolve={dec}", flush=True)
        ctx = gp.Context(N, p)
        ref, got = _spectra(gp, ctx, seed, nb, dec)
        kb_r, pw_r, ks_r, ps_r, occ_r = ref
        kb_g, pw_g, ks_g, ps_g, occ_g = got
        assert torch.equal(occ_r, occ_g)
        assert int(occ_r.sum()) > N ** 3 // 2
        sel = occ_r > 0
        for a, b in ((kb_g, kb_r), (pw_g, pw_r)):
            # FP64 sums of the same terms in another order
            assert torch.allclose(a[sel], b[sel], rtol=1e-12, atol=1e-12 * float(b[sel].abs().max()))
        # the scatters are square roots of differences of such sums: a change
        # of eps in the sums moves them by up to sqrt(eps) of the bin's value
        for a, b, scale in ((ks_g, ks_r, kb_r), (ps_g, ps_r, pw_r)):
            assert torch.allclose(a[sel], b[sel], rtol=1e-9, atol=1e-7 * float(scale[sel].abs().max()))
        assert torch.all(pw_g[~sel] == 0) and torch.all(kb_g[~sel] == 0)


def test_power_spectrum_noalloc_refuses_bins_beyond_the_maximum(gp):
    from gevpm._lib import GevError
    from gevpm.spectra import power_spectrum
    ctx = gp.Context(16, gp.GEV_FP64)
    F = ctx.fft_r2c(torch.zeros(1, 16, 16, 16, dtype=torch.float64, device="cuda"))
    with pytest.raises(GevError, match="numbins"):
        power_spectrum(ctx, F, 1025, True, noalloc=True)
