"""Whole-run parity at the sizes of the step's two reference workloads.

- 256^3 particles on a 256^3 mesh, L = 640 Mpc/h, weak-field GR cycle in
  FP64, 50 cycles from z = 100 to z = 10, against the oracle's main.cu-ordered
  cycle (oracle/gev_oracle_step.c) from the same initial state.
- 512^3 particles on a 512^3 mesh, L = 1280 Mpc/h, the mixed-precision GR
  cycle against the FP64 GPU cycle from the same IC, 100 cycles from z = 100
  to z = 0: density P(k) for k < k_Nyq/2, phi / chi / B relative L2 and final
  positions.

Both compare cell keys floor(pos N) per particle ID across the two
implementations: keys must agree except for particles that lie within eps of
a cell face in either run, and the number of such particles is reported.
"""
import math

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


def rel_l2(a, b):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rel_max(a, b):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


def cell_key_check(pa, pb, N, eps):
    """pa, pb: (3, n) positions in box units of the same particles (same order).
    Returns (mismatched keys, particles within eps cells of a face in either
    run, whether every mismatch is such a particle)."""
    xa, xb = pa * N, pb * N
    ka, kb = np.floor(xa), np.floor(xb)
    mism = np.any(ka != kb, axis=0)
    near = np.zeros(pa.shape[1], bool)
    for x in (xa, xb):
        near |= np.any(np.abs(x - np.rint(x)) < eps, axis=0)
    return int(mism.sum()), int(near.sum()), bool(np.all(near[mism]))


def gpu_ic(gp, N, L, seed=42, z_in=100.0, om=0.31):
    import gevpm.cosmology as cosmo
    k, P = cosmo.bbks_table(omega_m=om, h=0.67, n_s=0.965, sigma8=0.81, z=z_in)
    ctx = gp.Context(N, gp.GEV_FP64)
    p = gp.Particles(ctx, N ** 3, om / N ** 3)
    f = gp.Fields(ctx)
    st = gp.StepState(1, gp.fourpiG_of(L), om, 1 - om, z_relax=z_in)
    gp.ic_generate(p, f, st, seed, L, z_in, k, P)
    return ctx, p, f, st


def by_id_np(p):
    pos, q, pid = p.store()
    o = torch.argsort(pid)
    return pos[:, o].cpu().numpy(), q[:, o].cpu().numpy(), pid[o].cpu().numpy()


@pytest.mark.timeout(1200)
def test_gr_fp64_256_50_cycles_vs_oracle(gp, O):
    """256^3, L = 640 Mpc/h, GR FP64, 50 cycles z = 100 -> 10 against the oracle"""
    N, L, steps, om = 256, 640.0, 50, 0.31
    M = N ** 3
    ctx, p, f, st = gpu_ic(gp, N, L)
    pos0, q0, pid = by_id_np(p)
    assert np.array_equal(pid, np.arange(M))
    ost = dict(pos=np.ascontiguousarray(pos0), q=np.ascontiguousarray(q0),
               phi=f.phi.cpu().numpy().ravel().copy(), chi=f.chi.cpu().numpy().ravel().copy(),
               Bi=f.Bi.cpu().numpy().reshape(3, -1).copy(), source=np.zeros(M), Sij=np.zeros((6, M)),
               a=st.a, tau=st.tau, dtau_old=0.0, maxvel=st.maxvel)
    del pos0, q0
    fpg = gp.fourpiG_of(L)
    dtau = (gp.particle_horizon(1 / 11.0, fpg, om, 1 - om) - st.tau) / steps
    maxvel_err = 0.0
    for cyc in range(steps):
        st.dtau = dtau
        gp.step(p, f, st)
        ost["dtau"] = dtau
        O.step(N, ost, 1, om / M, fpg, om, 1 - om, 100.0)
        maxvel_err = max(maxvel_err, abs(st.maxvel - ost["maxvel"]) / ost["maxvel"])
        assert abs(st.a - ost["a"]) <= 1e-14 * ost["a"]
        if cyc % 10 == 9:
            print(f"  256^3 GR FP64 cycle {cyc + 1}/{steps}: z = {1 / st.a - 1:.3f}", flush=True)
    # RK4 in a reaches z = 10 up to its truncation error over 50 cycles
    assert abs(1 / st.a - 11.0) < 1e-4
    gpos, gq, _ = by_id_np(p)
    d = (gpos - ost["pos"] + 0.5) % 1.0 - 0.5
    dpos_max = float(np.abs(d).max() * N)
    dpos_rms = float(np.sqrt(np.mean(np.sum(d * d, 0))) * N)
    res = dict(dpos_max_cells=dpos_max, dpos_rms_cells=dpos_rms,
               q_rel_l2=rel_l2(gq, ost["q"]), q_rel_max=rel_max(gq, ost["q"]), maxvel_rel=maxvel_err)
    for nm, ref in (("phi", ost["phi"]), ("chi", ost["chi"]), ("Bi", ost["Bi"])):
        g = getattr(f, nm).cpu().numpy().reshape(ref.shape)
        res[nm + "_rel_l2"] = rel_l2(g, ref)
        res[nm + "_rel_max"] = rel_max(g, ref)
    eps = 1e-9
    mism, near, ok = cell_key_check(gpos, ost["pos"], N, eps)
    res.update(cell_key_mismatch=mism, within_eps_of_face=near, eps_cells=eps)
    print("  256^3 GR FP64 vs oracle after 50 cycles:", res, flush=True)
    # the runs differ by summation order only (atomic deposits, FFT order)
    assert dpos_max < 1e-10 and dpos_rms < 1e-11
    assert res["q_rel_max"] < 1e-9 and res["q_rel_l2"] < 1e-10
    assert res["phi_rel_max"] < 1e-9 and res["phi_rel_l2"] < 1e-10
    assert res["chi_rel_max"] < 1e-7 and res["chi_rel_l2"] < 1e-8
    assert res["Bi_rel_max"] < 1e-7 and res["Bi_rel_l2"] < 1e-8
    assert maxvel_err < 1e-9
    assert ok, f"{mism} cell keys differ, not all within {eps} cells of a face"
    assert near < 100


@pytest.mark.timeout(1200)
def test_gr_mixed_512_100_cycles_vs_fp64(gp):
    """512^3, L = 1280 Mpc/h: the mixed form against FP64 from the same IC over
    100 cycles z = 100 -> 0 (P(k) below k_Nyq/2, fields, positions, cell keys)"""
    from gevpm.spectra import density_power
    N, L, total, om = 512, 1280.0, 100, 0.31
    c64, p64, f64, st64 = gpu_ic(gp, N, L)
    cm = gp.Context(N, gp.GEV_MIXED)
    pm = gp.Particles(cm, N ** 3, om / N ** 3)
    pos, q, pid = p64.store()
    pm.load(pos, q, pid)
    del pos, q, pid
    fm = gp.Fields(cm)
    fm.copy_from(f64)
    stm = gp.StepState(1, st64.fourpiG, om, 1 - om, z_relax=100.0)
    for key in ("a", "tau", "dtau_old", "maxvel"):
        setattr(stm, key, getattr(st64, key))
    dtau = (gp.particle_horizon(1.0, st64.fourpiG, om, 1 - om) - st64.tau) / total
    for cyc in range(total):
        st64.dtau = dtau
        stm.dtau = dtau
        gp.step(pm, fm, stm)
        gp.step(p64, f64, st64)
        if cyc % 25 == 24:
            torch.cuda.synchronize()
            print(f"  512^3 mixed + FP64 cycle {cyc + 1}/{total}: z = {1 / st64.a - 1:.3f}", flush=True)
    assert abs(stm.a - st64.a) <= 1e-12 and abs(st64.a - 1.0) < 1e-4
    nb = 64
    kb, pw64, _, _, occ = density_power(c64, p64, nb, work=f64.source, F=f64.scalarFT)
    _, pwm, _, _, _ = density_power(cm, pm, nb, work=fm.source, F=fm.scalarFT)
    sel = (occ > 0) & (kb < 0.5 * math.pi * N) & (kb > 0)
    pk_err = float(((pwm[sel] - pw64[sel]).abs() / pw64[sel].abs()).max())
    res = dict(pk_rel_err_max=pk_err, pk_bins=int(sel.sum()))
    for nm in ("phi", "chi", "Bi"):
        a, b = getattr(fm, nm).double(), getattr(f64, nm).double()
        res[nm + "_rel_l2"] = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
        res[nm + "_rel_max"] = float((a - b).abs().max() / b.abs().max())
    del fm, f64
    torch.cuda.empty_cache()
    pa, _, ia = p64.store()
    oa = torch.argsort(ia)
    pa = pa[:, oa]
    del ia, oa
    pb, _, ib = pm.store()
    ob = torch.argsort(ib)
    pb = pb[:, ob]
    del ib, ob
    d = torch.remainder(pb - pa + 0.5, 1.0) - 0.5
    res["pos_rms_cells"] = float(d.pow(2).sum(0).mean().sqrt() * N)
    res["pos_max_cells"] = float(d.abs().max() * N)
    del d
    eps = 1e-3
    xa, xb = pa * N, pb * N
    mism = (torch.floor(xa) != torch.floor(xb)).any(0)
    near = ((xa - torch.round(xa)).abs() < eps).any(0) | ((xb - torch.round(xb)).abs() < eps).any(0)
    res.update(cell_key_mismatch=int(mism.sum()), within_eps_of_face=int(near.sum()), eps_cells=eps,
               mismatch_outside_eps=int((mism & ~near).sum()))
    print("  512^3 mixed vs FP64 after 100 cycles:", res, flush=True)
    assert sel.sum() >= 16
    assert pk_err < 1e-4
    assert res["phi_rel_l2"] < 1e-3 and res["chi_rel_l2"] < 1e-4 and res["Bi_rel_l2"] < 1e-4
    assert res["pos_rms_cells"] < 1e-4 and res["pos_max_cells"] < 1e-3
    assert res["mismatch_outside_eps"] == 0
