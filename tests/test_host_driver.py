"""The C++ façade (gevolution-2.0_amd/host/gevolution_pm.hpp) and its
main.cu-shaped driver main_pm: built here, run on the GPU, final particles
compared per ID with the same cycle run through gev_step on the same IC
(gev_step itself is held to the FP64 oracle in test_gpu_parity.py)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "gevolution-2.0_amd", "lib")


def test_driver_binaries_built():
    for name in ("main_pm", "main_pm_single"):
        assert os.access(os.path.join(LIB, name), os.X_OK), f"{name} not built (run __graft_entry__.build())"


def test_driver_rejects_bad_options(tmp_path):
    """argument errors are reported without touching the GPU"""
    r = subprocess.run([os.path.join(LIB, "main_pm"), "-bogus"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "unknown option" in r.stderr
    r = subprocess.run([os.path.join(LIB, "main_pm"), "-n", "16"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "-pk" in r.stderr
    # loadHomogeneousTemplate's failures are reported before any GPU work
    r = subprocess.run([os.path.join(LIB, "main_pm"), "-n", "16", "-template", str(tmp_path / "none.dat"), "-out",
                        str(tmp_path / "x.bin")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "unable to open template file" in r.stderr
    bad = tmp_path / "bad.dat"
    bad.write_bytes(b"\x10\x00\x00\x00" + bytes(16))
    r = subprocess.run([os.path.join(LIB, "main_pm"), "-n", "16", "-template", str(bad), "-out",
                        str(tmp_path / "x.bin")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "header not recognized" in r.stderr


_PK_WRITER = r"""
#include <cstdio>
#include <cstdlib>
#include "gevolution_pm.hpp"
using namespace gevpm;
// argv: file a z_target, then numbins rows of k P sigk sigP count on stdin
int main(int argc, char **argv)
{
    const int nb = 6;
    Real kb[nb], pw[nb], ks[nb], ps[nb];
    int oc[nb];
    for (int i = 0; i < nb; i++)
        if (std::scanf("%le %le %le %le %d", &kb[i], &pw[i], &ks[i], &ps[i], &oc[i]) != 5) return 2;
    writePowerSpectrum(kb, pw, ks, ps, oc, nb, (Real)2.0, (Real)4.0, argv[1], "power spectrum of delta_N",
                       std::atof(argv[2]), std::atof(argv[3]));
    return 0;
}
"""


def _table(path):
    lines = open(path).read().splitlines()
    return lines[:3], np.array([[float(x) for x in ln.split()] for ln in lines[3:]])


def test_write_power_spectrum_exact_redshift(tmp_path):
    """the façade's writePowerSpectrum built with EXACT_OUTPUT_REDSHIFTS
    (tools.hpp:370-452), host code only: a spectrum taken past z_target is
    interpolated against the file the previous call left, weight
    (z_prev - z_target) / (z_prev - z) with z_prev as printed in that file's
    header, and labelled z_target; no file -> written as is at its own
    redshift; an unparsable header -> weight 1 (this call's values) at z_target;
    before the crossing nothing is interpolated"""
    import shutil
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    src = tmp_path / "pkw.cpp"
    src.write_text(_PK_WRITER)
    exe = str(tmp_path / "pkw")
    host = os.path.join(ROOT, "gevolution-2.0_amd", "host")
    r = subprocess.run([gxx, "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-DEXACT_OUTPUT_REDSHIFTS",
                        "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"), "-I" + host, "-o", exe,
                        str(src)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rng = np.random.default_rng(3)
    occ = np.array([3, 0, 7, 12, 0, 30])
    k = np.arange(1, 7) * 0.5
    P1, P2 = rng.uniform(1, 9, 6), rng.uniform(1, 9, 6)
    sig = rng.uniform(0.1, 1, 6)

    def write(path, a, zt, P):
        rows = "\n".join("%.17g %.17g %.17g %.17g %d" % (k[i], P[i], 0.01 * k[i], sig[i], occ[i]) for i in range(6))
        out = subprocess.run([exe, str(path), repr(a), repr(zt)], input=rows, capture_output=True, text=True,
                             timeout=60)
        assert out.returncode == 0, out.stderr
        return out.stdout

    sel = occ > 0
    zt, z1, z2 = 9.5, 10.25, 8.75
    f = tmp_path / "pk.dat"
    write(f, 1 / (1 + z1), zt, P1)          # before the crossing: as is, at its own redshift
    head, t1 = _table(f)
    assert head[1] == "# redshift z=%f" % z1
    assert np.allclose(t1[:, 1], P1[sel] / 4, rtol=1e-6) and np.array_equal(t1[:, 4], occ[sel])
    write(f, 1 / (1 + z2), zt, P2)          # past it: interpolated to z_target
    head, t2 = _table(f)
    zprev = float("%f" % z1)
    w = (zprev - zt) / (1 + zprev - (1 + z2))
    assert head[1] == "# redshift z=%f" % zt
    assert np.allclose(t2[:, 1], (1 - w) * t1[:, 1] + w * P2[sel] / 4, rtol=1e-6)
    assert np.allclose(t2[:, 0], k[sel] / 2) and np.allclose(t2[:, 3], sig[sel] / 4 / np.sqrt(occ[sel]), rtol=1e-6)
    g = tmp_path / "fresh.dat"              # no earlier file: nothing to interpolate
    write(g, 1 / (1 + z2), zt, P2)
    head, t3 = _table(g)
    assert head[1] == "# redshift z=%f" % z2 and np.allclose(t3[:, 1], P2[sel] / 4, rtol=1e-6)
    h = tmp_path / "bad.dat"                # unparsable header: weight 1, labelled z_target
    h.write_text("# x\nnot a header\n")
    out = write(h, 1 / (1 + z2), zt, P2)
    assert "error parsing power spectrum file header" in out
    head, t4 = _table(h)
    assert head[1] == "# redshift z=%f" % zt and np.allclose(t4[:, 1], P2[sel] / 4, rtol=1e-6)


def _read_out(path):
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(8), np.int64)[0])
        pid = np.frombuffer(f.read(8 * n), np.int64)
        pos = np.frombuffer(f.read(24 * n), np.float64).reshape(3, n)
        q = np.frombuffer(f.read(24 * n), np.float64).reshape(3, n)
    o = np.argsort(pid)
    return pid[o], pos[:, o], q[:, o]


def _check_files(snap, restart, pid, pos, q, L, single):
    """the façade's saveGadget2 / saveExpress output against -out: kpc/h
    positions, q / (a^1.5 GADGET_VELOCITY_CONVERSION) velocities, header
    metadata of output.hpp:418-466"""
    raw = open(snap, "rb").read()
    n = len(pid)
    assert np.frombuffer(raw, np.uint32, 1, 0)[0] == 256
    hb = raw[4:260]
    a = np.frombuffer(hb, np.float64, 1, 72)[0]
    assert np.frombuffer(hb, np.uint32, 6, 0)[1] == n and np.frombuffer(hb, np.float64, 1, 128)[0] == L * 1000
    m = 2.77459457e11 * 0.31 * L ** 3 / n / 1e10
    assert abs(np.frombuffer(hb, np.float64, 6, 24)[1] / m - 1) < 1e-12
    # records: [4|256|4] header, then [4|12n|4] positions, velocities, [4|8n|4] IDs
    fpos = np.frombuffer(raw, np.float32, 3 * n, 268).reshape(n, 3)
    fvel = np.frombuffer(raw, np.float32, 3 * n, 276 + 12 * n).reshape(n, 3)
    fid = np.frombuffer(raw, np.int64, n, 284 + 24 * n)
    f = np.argsort(fid)
    assert np.array_equal(fid[f], pid)
    assert np.allclose(fpos[f], (pos * L * 1000).T, rtol=1e-6, atol=L * 1000 * 1e-6)
    assert np.allclose(fvel[f], (q / (a ** 1.5 * 3.335640952e-6)).T, rtol=1e-6, atol=1e-3)
    rs = 4 if single else 8
    er = open(restart, "rb").read()
    assert er[:8] == b"GEXPCDM\x00" and np.frombuffer(er, np.uint32, 1, 20)[0] == rs
    eid = np.frombuffer(er, np.int64, n, 120 + 6 * n * rs)
    e = np.argsort(eid)
    assert np.array_equal(eid[e], pid)
    eq = np.frombuffer(er, np.float32 if single else np.float64, 3 * n, 120 + 3 * n * rs).reshape(n, 3)
    assert np.allclose(eq[e], q.T, rtol=1e-6, atol=1e-12)
    assert er[-256:] == hb  # the Gadget-2 header closes the express file


def _check_spectrum(path, gevpm, pid, pos, q, N, L, numbins, single):
    """the façade's extractPowerSpectrum + writePowerSpectrum (-spectra) against
    the Python binding's density spectrum of the final particles from -out:
    the text table of writePowerSpectrum, k / L, P / (N^6 2 pi^2 Omega_m^2),
    sigma(P) / sqrt(count), one line per occupied bin"""
    from gevpm.spectra import density_power
    lines = open(path).read().splitlines()
    assert lines[0] == "# power spectrum of delta_N" and lines[1].startswith("# redshift z=")
    assert lines[2].split() == ["#", "k", "Pk", "sigma(k)", "sigma(Pk)", "count"]
    tab = np.array([[float(x) for x in ln.split()] for ln in lines[3:]])
    ctx = gevpm.Context(N, gevpm.GEV_FP64)
    p = gevpm.Particles(ctx, N ** 3, 0.31 / N ** 3)
    p.load(pos, q, pid)
    kb, pw, ks, ps, occ = (t.numpy() for t in density_power(ctx, p, numbins))
    sel = occ > 0
    rescalep = float(N ** 3) ** 2 * 2 * np.pi ** 2 * 0.31 ** 2
    assert len(tab) == sel.sum() and np.array_equal(tab[:, 4].astype(np.int64), occ[sel])
    tol = 1e-4 if single else 1e-5  # %e prints 7 digits; the mixed run deposits in FP32
    assert np.allclose(tab[:, 0], kb[sel] / L, rtol=1e-5)
    assert np.allclose(tab[:, 1], pw[sel] / rescalep, rtol=tol)
    assert np.allclose(tab[:, 3], ps[sel] / rescalep / np.sqrt(occ[sel]), rtol=1e-3, atol=1e-3 * tab[:, 3].max())


def _check_exact_spectrum(path, at_start, z_pk, N, L, single):
    """-zpk (main.cu:1080-1131 with EXACT_OUTPUT_REDSHIFTS): the file holds the
    density spectra of the last cycle start before z_pk (c0) and the first past
    it, interpolated in redshift to z_pk with z_prev as printed in the header"""
    zs = [1 / a - 1 for a, _ in at_start]
    c0 = max(c for c in range(len(zs) - 1) if zs[c] >= z_pk > zs[c + 1])
    lines = open(path).read().splitlines()
    assert lines[0] == "# power spectrum of delta_N" and lines[1] == "# redshift z=%f" % z_pk
    tab = np.array([[float(x) for x in ln.split()] for ln in lines[3:]])
    rescalep = float(N ** 3) ** 2 * 2 * np.pi ** 2 * 0.31 ** 2
    kb, p0, _, _, occ = at_start[c0][1]
    p1, ps1 = at_start[c0 + 1][1][1], at_start[c0 + 1][1][3]
    sel = occ > 0
    zprev = float("%f" % zs[c0])
    w = (zprev - z_pk) / (zprev - zs[c0 + 1])
    assert 0 < w < 1
    tol = 1e-4 if single else 1e-5
    assert len(tab) == sel.sum() and np.array_equal(tab[:, 4].astype(np.int64), occ[sel])
    assert np.allclose(tab[:, 0], kb[sel] / L, rtol=1e-5)
    assert np.allclose(tab[:, 1], ((1 - w) * p0[sel] + w * p1[sel]) / rescalep, rtol=tol)
    # sigma(Pk) is the later spectrum's, not interpolated
    assert np.allclose(tab[:, 3], ps1[sel] / rescalep / np.sqrt(occ[sel]), rtol=1e-3, atol=1e-3 * tab[:, 3].max())


def _check_metric_spectra(prefix, ctx, f, a, N, L, newton, single):
    """the phi and (GR) B spectra the driver writes at the end (the façade's
    extractPowerSpectrum of a scalar and of a three-component field, no
    window) against the Python binding's spectra of the replay's fields"""
    from gevpm.spectra import cross_spectrum, power_spectrum
    cases = [("_phi.dat", f.phi, float(N ** 3) ** 2 * 2 * np.pi ** 2)]
    if not newton:
        cases.append(("_B.dat", f.Bi, a ** 4 * N ** 2 * 2 * np.pi ** 2))
    for suffix, fld, rescalep in cases:
        lines = open(prefix + suffix).read().splitlines()
        tab = np.array([[float(x) for x in ln.split()] for ln in lines[3:]])
        F = ctx.fft_r2c(fld)
        if fld.shape[0] == 1:
            res = power_spectrum(ctx, F, 20, deconvolve=False)
        else:
            res = cross_spectrum(ctx, F, F, 20, deconvolve=False)
        kb, pw, _, _, occ = (t.numpy() for t in res)
        sel = occ > 0
        assert len(tab) == sel.sum() and np.array_equal(tab[:, 4].astype(np.int64), occ[sel]), suffix
        assert np.allclose(tab[:, 0], kb[sel] / L, rtol=1e-5), suffix
        tol = 1e-3 if single else 1e-5  # the driver's and gev_step's mixed fields differ at float rounding
        assert np.allclose(tab[:, 1], pw[sel] / rescalep, rtol=tol, atol=tol * np.abs(tab[:, 1]).max()), suffix


@pytest.mark.gpu
@pytest.mark.parametrize("newton,single", [(1, False), (0, False), (0, True)])
def test_driver_matches_gev_step(tmp_path, newton, single):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gevpm
    import gevpm.cosmology as cosmo
    N, L, z_in, z_end, steps, seed = 16, 320.0, 100.0, 50.0, 4, 11
    z_pk = 75.0  # between the starts of cycles 1 (z ~ 82) and 2 (z ~ 69)
    k, P = cosmo.bbks_table(z=z_in)
    mpk = str(tmp_path / "pk.txt")
    cosmo.write_mpk(mpk, k, P)
    out = str(tmp_path / "final.bin")
    snap, restart = str(tmp_path / "snap_cdm"), str(tmp_path / "restart")
    exe = os.path.join(LIB, "main_pm_single" if single else "main_pm")
    cmd = [exe, "-n", str(N), "-L", str(L), "-zin", str(z_in), "-zend", str(z_end), "-steps", str(steps),
           "-seed", str(seed), "-pk", mpk, "-out", out, "-gadget", snap, "-express", restart,
           "-spectra", str(tmp_path / "pk"), "-numbins", "20", "-zpk", repr(z_pk)] + (["-newton"] if newton else [])
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    pid, pos, q = _read_out(out)
    assert np.array_equal(pid, np.arange(N ** 3))
    _check_files(snap, restart + ".0", pid, pos, q, L, single)
    _check_spectrum(str(tmp_path / "pk_deltaN.dat"), gevpm, pid, pos, q, N, L, 20, single)

    # the same run through gev_step (fused kick+drift, one re-binning)
    fpg = gevpm.fourpiG_of(L)
    c64 = gevpm.Context(N, gevpm.GEV_FP64)
    p64 = gevpm.Particles(c64, N ** 3, 0.31 / N ** 3)
    f64 = gevpm.Fields(c64)
    st = gevpm.StepState(0 if newton else 1, fpg, 0.31, 0.69, z_relax=z_in)
    gevpm.ic_generate(p64, f64, st, seed, L, z_in, k, P)
    if single:
        ctx = gevpm.Context(N, gevpm.GEV_MIXED)
        p = gevpm.Particles(ctx, N ** 3, 0.31 / N ** 3)
        a, b, c = p64.store()
        p.load(a, b, c)
        f = gevpm.Fields(ctx)
        f.copy_from(f64)
        st2 = gevpm.StepState(st.gr, fpg, 0.31, 0.69, z_relax=z_in)
        for key in ("a", "tau", "dtau_old", "maxvel"):
            setattr(st2, key, getattr(st, key))
        st = st2
    else:
        p, f = p64, f64
    dtau = (gevpm.particle_horizon(1 / (1 + z_end), fpg, 0.31, 0.69) - st.tau) / steps
    from gevpm.spectra import density_power
    at_start = []  # (a, density spectrum) at each cycle start, where main.cu writes spectra
    for _ in range(steps):
        st.dtau = dtau
        at_start.append((st.a, [t.cpu().numpy() for t in density_power(ctx if single else c64, p, 20)]))
        gevpm.step(p, f, st)
    _check_exact_spectrum(str(tmp_path / "pk000_deltaN.dat"), at_start, z_pk, N, L, single)
    _check_metric_spectra(str(tmp_path / "pk"), ctx if single else c64, f, st.a, N, L, newton, single)
    rp, rq, rid = (t.cpu().numpy() for t in p.store())
    o = np.argsort(rid)
    rp, rq = rp[:, o], rq[:, o]
    dpos = np.abs(((pos - rp) + 0.5) % 1.0 - 0.5).max() * N   # in cells
    dq = np.abs(q - rq).max() / np.abs(rq).max()
    # FP64: the two paths differ only in the kick/drift fusion and summation order;
    # mixed: FP32 offsets and momenta
    tol_pos, tol_q = (2e-5, 2e-5) if single else (1e-9, 1e-9)
    assert dpos < tol_pos, dpos
    assert dq < tol_q, dq
    last = [ln for ln in r.stdout.splitlines() if ln.startswith("done")]
    # RK4 in a(tau) over 4 large steps lands within ~1e-5 of the target redshift
    assert last and abs(float(last[0].split("z=")[1]) - z_end) < 1e-4


@pytest.mark.gpu
def test_driver_lightcone_output(tmp_path):
    """main_pm -lightcone: the façade's findIntersectingLightcones,
    LightconeIDBacklog and lightcone saveGadget2 inside the cycle; each file
    holds distinct IDs at distances from the observer inside that cycle's
    shell, and few IDs appear in two consecutive files"""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gevpm
    import gevpm.cosmology as cosmo
    N, L, z_in, z_end, steps = 16, 320.0, 100.0, 60.0, 8
    fpg = gevpm.fourpiG_of(L)
    k, P = cosmo.bbks_table(z=z_in)
    mpk = str(tmp_path / "pk.txt")
    cosmo.write_mpk(mpk, k, P)
    tau_in = gevpm.particle_horizon(1 / (1 + z_in), fpg, 0.31, 0.69)
    tau_end = gevpm.particle_horizon(1 / (1 + z_end), fpg, 0.31, 0.69)
    dtau = (tau_end - tau_in) / steps
    # observer whose horizon is 0.45 box beyond the initial time: shells of
    # radius <= 0.5 box meet at most the faces' periodic copies
    target = tau_in + 0.45
    lo, hi = 1 / (1 + z_in), 1.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if gevpm.particle_horizon(mid, fpg, 0.31, 0.69) < target else (lo, mid)
    lcz = 1 / lo - 1
    lcv = (0.5, 0.45, 0.55)
    prefix = str(tmp_path / "lc")
    cmd = [os.path.join(LIB, "main_pm"), "-n", str(N), "-L", str(L), "-zin", str(z_in), "-zend", str(z_end),
           "-steps", str(steps), "-pk", mpk, "-lightcone", prefix, "-lcvertex", "%g,%g,%g" % lcv, "-lcz", repr(lcz)]
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    files = sorted(f for f in os.listdir(tmp_path) if f.startswith("lc_") and f.endswith("_cdm"))
    assert len(files) >= 2
    ids = []
    for name in files:
        cycle = int(name.split("_")[1])
        raw = open(tmp_path / name, "rb").read()
        n = int(np.frombuffer(raw, np.uint32, 6, 4)[1])
        pos = np.frombuffer(raw, np.float32, 3 * n, 268).reshape(n, 3) / (L * 1000)
        pid = np.frombuffer(raw, np.int64, n, 284 + 24 * n)
        assert len(np.unique(pid)) == n
        tau = tau_in + cycle * dtau
        dist = target - tau  # d(lcz) - tau
        inner, outer = dist - 0.5 * dtau, dist + 0.55 * dtau
        d = np.linalg.norm(pos - np.array(lcv), axis=1)
        # moved to the crossing by at most |v| dtau (|v| << 1 at z > 60)
        assert (d > inner - 0.01 * dtau).all() and (d < outer + 0.01 * dtau).all()
        ids.append(set(pid.tolist()))
    for a_, b_ in zip(ids, ids[1:]):
        assert len(a_ & b_) <= 0.05 * max(len(b_), 1)


@pytest.mark.gpu
@pytest.mark.parametrize("single", [False, True])
def test_driver_template_lattice(tmp_path, single):
    """main_pm -template: loadHomogeneousTemplate, initialize(info, lat, boxSize,
    capacity, extra), initializeParticlePositions (addParticle_global per
    particle), updateRowBuffers and sampleParticles(Real*, Real*, long*, n) of
    the façade on the reference's sc1 template: the reference's IDs, particles
    at rest, and positions equal, particle for particle, to the cell-centred
    lattice gev_pcls_init_lattice builds"""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gevpm
    N = 16
    out = str(tmp_path / "lattice.bin")
    exe = os.path.join(LIB, "main_pm_single" if single else "main_pm")
    tmpl = os.path.join(ROOT, "tests", "golden", "sc1_crystal.dat")
    cmd = [exe, "-n", str(N), "-template", tmpl, "-out", out]
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    pid, pos, q = _read_out(out)
    nt = N // 4
    assert np.array_equal(pid, np.arange(N ** 3))
    assert np.all(q == 0.0)
    # the reference's ID: i + 64 (xtile + nt (ytile + nt ztile)), i the template index
    i = pid % 64
    tile = pid // 64
    xt, yt, zt = tile % nt, (tile // nt) % nt, tile // (nt * nt)
    t = np.fromfile(tmpl, np.float32, 3 * 64, offset=268).reshape(64, 3).astype(np.float64)
    expect = np.stack([(xt + t[i, 0]) / nt, (yt + t[i, 1]) / nt, (zt + t[i, 2]) / nt])
    assert np.array_equal(pos, expect)
    ctx = gevpm.Context(N, gevpm.GEV_MIXED if single else gevpm.GEV_FP64)
    p = gevpm.Particles(ctx, N ** 3, 1.0)
    p.init_lattice()
    lpos, _, lid = (x.cpu().numpy() for x in p.store())
    lattice = lpos[:, np.argsort(lid)]
    cells = np.floor(pos * N).astype(np.int64)
    key = cells[0] + N * (cells[1] + N * cells[2])
    assert np.array_equal(pos, lattice[:, key])  # the same point set, bit for bit
