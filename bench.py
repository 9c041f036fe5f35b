#!/usr/bin/env python3
"""Throughput of the gevolution particle-mesh cycle on one MI355X.

Workload (default): the weak-field GR cycle of main.cu (T00/Tij/T0i CIC
projections, phi/chi/B solves, kick, drift + re-binning) with 512^3
particles on a 512^3 mesh, L = 1280 Mpc/h, Zel'dovich ICs from a BBKS LCDM
spectrum (Omega_m=0.31, h=0.67, n_s=0.965, sigma_8=0.81) at z=100, 100
cycles uniform in conformal time to z=0, in the mixed-precision form.  The
same run is repeated in FP64 from the same IC to report the FP64 rate and
the P(k) / field errors of the mixed form.

Prints one JSON line.  --gpus N > 1 is rejected: this build runs on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gevolution-2.0_amd", "python"))

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s at 512^3 (FP64 vs mixed, P(k) err); adjoint/forward time ratio"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=98, help="timed cycles")
    ap.add_argument("--warmup", type=int, default=2, help="untimed cycles before the timed ones")
    ap.add_argument("--mesh", type=int, default=512)
    ap.add_argument("--boxsize", type=float, default=1280.0)
    ap.add_argument("--z-in", type=float, default=100.0)
    ap.add_argument("--z-end", type=float, default=0.0)
    ap.add_argument("--total-steps", type=int, default=100, help="cycles spanning z_in -> z_end")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 leg (no P(k) error)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--adj-mesh", type=int, default=128)
    ap.add_argument("--adj-steps", type=int, default=20)
    ap.add_argument("--adj-modes", type=int, default=64)
    ap.add_argument("--cpu-mesh", type=int, default=512)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only time the mixed-form adjoint workloads (no FP64 gradient, no finite differences), "
                         "print their JSON and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB fill and three 1 GiB copies (the reference kernels that scale the "
                         "FETCH_SIZE / WRITE_SIZE counter readings to bytes), then exit")
    return ap.parse_args()


def copy_calibration():
    """a 1 GiB device fill, then three 1 GiB device-to-device copies"""
    import torch
    x = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
    x.fill_(1.0)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    log("copy calibration: 1 GiB fill + 3 x 1 GiB copy done")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs the PM step on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is designed but not built.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.adjoint_only:
        out = adjoint_leg(args, checks=False)
        out["gr"] = adjoint_gr_leg(args, checks=False)
        print(json.dumps(out), flush=True)
        return
    import numpy as np
    import torch

    import gevpm
    import gevpm.cosmology as cosmo
    from gevpm.spectra import density_power

    N, L = args.mesh, args.boxsize
    om, ol = 0.31, 0.69
    fpg = gevpm.fourpiG_of(L)
    k, P = cosmo.bbks_table(omega_m=om, h=0.67, n_s=0.965, sigma8=0.81, z=args.z_in)
    total = args.total_steps
    assert args.warmup + args.steps <= total, "warmup + steps must not exceed --total-steps"

    # ---------------------------------------------------------------- IC (FP64)
    t0 = time.time()
    c64 = gevpm.Context(N, gevpm.GEV_FP64)
    p64 = gevpm.Particles(c64, N ** 3, om / N ** 3)
    f64 = gevpm.Fields(c64)
    st64 = gevpm.StepState(1, fpg, om, ol, z_relax=args.z_in)
    gevpm.ic_generate(p64, f64, st64, 42, L, args.z_in, k, P)
    a_end = 1.0 / (1.0 + args.z_end)
    dtau = (gevpm.particle_horizon(a_end, fpg, om, ol) - st64.tau) / total
    torch.cuda.synchronize()
    ic_s = time.time() - t0
    log(f"IC {N}^3 generated in {ic_s:.1f}s, a={st64.a:.5f} tau={st64.tau:.4f} dtau={dtau:.5f}")

    # ------------------------------------------------------- mixed-form copy
    cm = gevpm.Context(N, gevpm.GEV_MIXED)
    pm = gevpm.Particles(cm, N ** 3, om / N ** 3)
    pos, q, pid = p64.store()
    pm.load(pos, q, pid)
    del pos, q, pid
    fm = gevpm.Fields(cm)
    fm.copy_from(f64)
    stm = gevpm.StepState(1, fpg, om, ol, z_relax=args.z_in)
    for key in ("a", "tau", "dtau_old", "maxvel"):
        setattr(stm, key, getattr(st64, key))

    def run(ctx, pcls, fields, st, label):
        """warmup cycles, the timed cycles, then the rest of the workload's
        cycles; each segment is bracketed by events on the library stream's
        device (synchronised), so the whole run's average is reported beside
        the timed window's.  Phase times come from the library's own events
        on its streams, every cycle."""
        gevpm.step_timing(ctx, True)
        phases, phases_all = {}, {}

        def segment(ncyc, acc):
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            w0 = time.time()
            for _ in range(ncyc):
                st.dtau = dtau
                gevpm.step(pcls, fields, st)
                for kname, v in gevpm.step_timings(ctx).items():
                    for a_ in acc:
                        a_[kname] = a_.get(kname, 0.0) + v
            e1.record()
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) if ncyc else 0.0, time.time() - w0

        z0 = 1.0 / st.a - 1.0
        ms_w, _ = segment(args.warmup, [phases_all])
        z1 = 1.0 / st.a - 1.0
        ms_t, wall = segment(args.steps, [phases, phases_all])
        z2 = 1.0 / st.a - 1.0
        ms_r, _ = segment(total - args.warmup - args.steps, [phases_all])
        ms = ms_t / args.steps
        log(f"{label}: {ms:.2f} ms/cycle over cycles {args.warmup + 1}-{args.warmup + args.steps} "
            f"(wall {1e3 * wall / args.steps:.2f}), all {total} cycles {(ms_w + ms_t + ms_r) / total:.2f} ms/cycle, "
            f"z_final={1 / st.a - 1:.4f}")
        window = {"timed_cycles": [args.warmup + 1, args.warmup + args.steps], "timed_z": [z1, z2],
                  "all_cycles": {"cycles": total, "z": [z0, 1.0 / st.a - 1.0],
                                 "ms_per_step": (ms_w + ms_t + ms_r) / total,
                                 "value": N ** 3 * 1e3 * total / (ms_w + ms_t + ms_r),
                                 "phases_ms": {kk: v / total for kk, v in phases_all.items()}}}
        return ms, {kk: v / args.steps for kk, v in phases.items()}, window

    ms_m, ph_m, win_m = run(cm, pm, fm, stm, "mixed")
    result = {"ic_fp64_s": ic_s, "rows_next": rows_next(cm, pm, fm, stm, dtau, L, om, ol, fpg)}
    if not args.no_fp64:
        ms_64, ph_64, win_64 = run(c64, p64, f64, st64, "fp64")
        # ---------------------------------------------- errors of the mixed form
        nb = 64
        kb, pw64, _, _, occ = density_power(c64, p64, nb, work=f64.source, F=f64.scalarFT)
        _, pwm, _, _, _ = density_power(cm, pm, nb, work=fm.source, F=fm.scalarFT)
        knyq = math.pi * N  # in units of 1/box (k L)
        sel = (occ > 0) & (kb < 0.5 * knyq) & (kb > 0)
        pk_err = float(((pwm[sel] - pw64[sel]).abs() / pw64[sel].abs()).max())

        def rel_l2(a, b):
            a, b = a.double(), b.double()
            return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))

        def rel_max(a, b):
            a, b = a.double(), b.double()
            return float((a - b).abs().max() / b.abs().max())

        fields_err = {nm: {"rel_l2": rel_l2(getattr(fm, nm), getattr(f64, nm)),
                           "rel_max": rel_max(getattr(fm, nm), getattr(f64, nm))} for nm in ("phi", "chi", "Bi")}
        pa, _, ia = p64.store()
        pb, _, ib = pm.store()
        oa, ob = torch.argsort(ia), torch.argsort(ib)
        d = torch.remainder(pb[:, ob] - pa[:, oa] + 0.5, 1.0) - 0.5
        pos_err = {"rms_cells": float(d.pow(2).sum(0).mean().sqrt() * N), "max_cells": float(d.abs().max() * N)}
        del pa, pb, ia, ib, d
        result.update(fp64={"value": N ** 3 * 1e3 / ms_64, "ms_per_step": ms_64, "phases_ms": ph_64,
                            "all_cycles": win_64["all_cycles"]},
                      mixed_over_fp64_speedup=ms_64 / ms_m,
                      mixed_over_fp64_speedup_all_cycles=win_64["all_cycles"]["ms_per_step"] /
                      win_m["all_cycles"]["ms_per_step"],
                      pk_rel_err_max=pk_err,
                      pk_bins_compared=int(sel.sum()), field_errors=fields_err, position_errors=pos_err)
    torch.cuda.empty_cache()

    # ----------------------------------------------------- dominant kernel
    roof = kernel_roofline(cm, pm.n, win_m["all_cycles"]["phases_ms"], total)
    cpu = None if args.no_cpu else cpu_baseline(args, om, ol, k, P)
    if not args.no_adjoint:
        torch.cuda.empty_cache()
        result["adjoint"] = adjoint_leg(args)
        torch.cuda.empty_cache()
        result["adjoint"]["gr"] = adjoint_gr_leg(args)
    value = N ** 3 * 1e3 / ms_m
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_m,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"GR weak-field PM cycle (main.cu order), {N}^3 particles on {N}^3 mesh, "
                               f"L={L:g} Mpc/h, BBKS Zel'dovich IC z={args.z_in:g}, {total} cycles to "
                               f"z={args.z_end:g}, mixed precision (FP32 fields/FFT/deposit, FP64 k=0, "
                               "chi combination, sums)",
                   "mesh": N, "particles": N ** 3, "boxsize_mpc_h": L, "cycles": total,
                   "timed_cycles": win_m["timed_cycles"], "timed_z": win_m["timed_z"]},
        "all_cycles": win_m["all_cycles"],
        "phases_ms": ph_m,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    line.update(result)
    print(json.dumps(line), flush=True)


def rows_next(cm, pm, fm, st, dtau, L, om, ol, fpg):
    """The callers beside the step, timed once on the final mixed state (each
    is a whole library call, so host work and PCIe copies are included where
    the call has them): P(k) of the CIC density (deposit, r2c, binning), a
    Gadget-2 snapshot of every 64th particle and the particle lightcone of a
    full-sky observer at the box centre whose shell (radius 0.4 box) lies
    inside the box, every 8th particle, written to a temporary directory."""
    import tempfile
    import torch
    import gevpm
    from gevpm import lightcone as lcm
    from gevpm.spectra import density_power

    def timed(fn):
        torch.cuda.synchronize()
        t = time.time()
        r = fn()
        torch.cuda.synchronize()
        return 1e3 * (time.time() - t), r

    out = {}
    density_power(cm, pm, 64, work=fm.source, F=fm.scalarFT)  # warm
    out["pk_ms"], _ = timed(lambda: density_power(cm, pm, 64, work=fm.source, F=fm.scalarFT))
    with tempfile.TemporaryDirectory() as d:
        hdr = gevpm.gev_gadget2_header()
        hdr.num_files = 1
        hdr.time, hdr.BoxSize = st.a, L * 1000.0
        ids = pm.store()[2]
        sel = int((ids % 64 == 0).sum())
        del ids
        hdr.npart[1] = hdr.npartTotal[1] = sel
        out["gadget2_ms"], _ = timed(lambda: pm.save_gadget2(os.path.join(d, "snap"), hdr, tracer_factor=64,
                                                             phi=fm.phi))
        out["gadget2_particles"] = sel
        geo = lcm.Geometry(vertex=(0.5, 0.5, 0.5))
        dist = 0.4
        nv, verts = lcm.vertices(geo, dist + (0.5 + lcm.IDCHECK_ZONE) * dtau, dist - 0.5 * dtau,
                                 lcm.box_domain(cm.N))
        log_ = lcm.IDLog(cm)
        h2 = gevpm.gev_gadget2_header()
        h2.num_files = 1
        h2.time, h2.BoxSize = st.a, L * 1000.0
        dadtau = st.a * gevpm.hconf(st.a, fpg, om, ol)
        out["lightcone_ms"], w = timed(lambda: pm.save_gadget2_lightcone(
            os.path.join(d, "lc"), h2, geo, dist, dtau, dtau, dadtau, verts, log_, fm.phi, tracer_factor=8))
        out["lightcone_particles"] = w
        out["lightcone_vertices"] = nv
    return out


# Algorithmic HBM bytes per launch of the kernels that can dominate the mixed
# cycle (R = 4-byte reals, n particles, M cells): each input read once, each
# output written once, no halo or re-read counted.
def _bytes_models(M, n, R=4, links=True):
    # links: the GR step sums T0i on the links in the T00+Tij pass (128 <= N <= 1024)
    nc = 10 if links else 7
    return {
        "kick_drift": (n * (7 * R + 8 * R + 8) + M * 5 * R,
                       "k_update_tiled (fused GR kick + drift): per particle 3 offsets + 3 momenta + cell read, "
                       "3 offsets + 3 momenta + cell + rank written, cell counter read + written; per cell phi, "
                       "chi, B0..B2 read"),
        "deposit_T00_Tij": (n * 7 * R + M * (4 + R + nc * R),
                            f"k_deposit T00+Tij{'+T0i links' if links else ''}: per particle 3 offsets + 3 momenta "
                            f"+ cell read; per cell start + phi read, {nc} components written"),
        "deposit_T0i": (n * 7 * R + M * (4 + R + 3 * R),
                        "k_deposit T0i: per particle 3 offsets + 3 momenta + cell read; per cell start + phi read, "
                        "3 components written"),
    }


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_gr_mixed_512.txt")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) >= 3 and parts[0] == "phase" and parts[1] == kernel:
                kv = dict(zip(parts[2::2], parts[3::2]))
                if "traffic_bytes" in kv:
                    return float(kv["traffic_bytes"])
    return None


# the library kernel each single-kernel phase brackets (mixed form)
# (the names a rocprofv3 kernel trace shows)
_PHASE_KERNEL = {
    "deposit_T00_Tij": "gev::k_deposit<float, 5, 16, 4, 4>",
    "kick_drift": "gev::k_update_tiled<float, 16, 8, 8, 5, 512>",
    "deposit_T0i": "gev::k_deposit<float, 1, 32, 4, 4>",
}


def kernel_roofline(ctx, n, phases, cycles):
    """Roofline of the single kernel with the largest time over all cycles of
    the workload.  Its launch duration is the library's own HIP-event timing
    of that phase on the library stream (the phase brackets that one kernel
    and a one-word fill), averaged over every cycle of the run, so it can be
    compared with the kernel's average duration in a rocprofv3 kernel trace
    of the same command."""
    models = _bytes_models(ctx.M, n, links=128 <= ctx.N <= 1024)
    name = max(models, key=lambda kk: phases.get(kk, 0.0))
    bytes_per_launch, desc = models[name]
    ms = phases[name]
    gbs = bytes_per_launch / (ms * 1e-3) / 1e9
    return {"kernel": name, "hip_kernel": _PHASE_KERNEL.get(name), "bound": "hbm", "achieved": gbs,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS, "traffic": measured_traffic(name),
            "avg_launch_ms": ms, "launches_averaged": cycles, "algorithmic_bytes_per_launch": bytes_per_launch,
            "bytes_model": desc}


def adjoint_gr_leg(args, checks=True):
    """The same two differentiable workloads with the weak-field GR cycle
    (gevpm.adjoint.PMRunGR: T00/Tij/T0i deposits, phi/chi/B solves, GR kick
    and drift): adj-mesh^3 for adj-steps cycles (ckpt every cycle) and twice
    the mesh for twice the cycles with a checkpoint every 4 cycles, both in the
    mixed form; backward / forward time ratios, peak device memory of the
    checkpointed run, and its gradient against the FP64 form's on the eight
    half-size crops."""
    import numpy as np
    import torch
    import gevpm
    from gevpm.adjoint import PMRunGR

    def smooth(seed, amp, N):
        r = np.random.default_rng(seed)
        F = np.fft.rfftn(r.standard_normal((N, N, N)))
        kk = np.fft.fftfreq(N, 1.0 / N)
        kz, ky, kx = np.meshgrid(kk, kk, np.arange(N // 2 + 1), indexing="ij")
        F *= np.exp(-(kx ** 2 + ky ** 2 + kz ** 2) / (2 * (N / 16.0) ** 2))
        f = np.fft.irfftn(F, s=(N, N, N), axes=(0, 1, 2))
        return amp * f / f.std()

    def used_bytes():
        free, total = torch.cuda.mem_get_info()
        return total - free

    def timed(N, steps, prec, ckpt, peak=None):
        amp = 0.3 / N ** 2
        run = PMRunGR(N, 2.0 * N, 50.0, 0.0, steps, precision=prec, ckpt_every=ckpt)
        dev = run.ctx.device
        t = run.forward(torch.from_numpy(smooth(12, amp, N)).to(dev)).double().clone()
        psi = torch.from_numpy(smooth(11, amp, N)).to(dev).reshape(1, N, N, N)
        run.backward(2 * (run.forward(psi).double() - t))  # warm-up
        torch.cuda.synchronize()
        if peak is not None:
            run.mem_probe = lambda: peak.__setitem__(0, max(peak[0], used_bytes()))
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record()
        rho = run.forward(psi)
        e[1].record()
        g = run.backward(2 * (rho.double() - t))
        e[2].record()
        torch.cuda.synchronize()
        res = {"forward_ms": e[0].elapsed_time(e[1]), "backward_ms": e[1].elapsed_time(e[2]),
               "adjoint_over_forward": e[1].elapsed_time(e[2]) / e[0].elapsed_time(e[1]),
               "checkpoint_bytes": run.checkpoint_bytes()}
        g = g.double().cpu().numpy().reshape(N, N, N)
        del run, t, psi, rho
        torch.cuda.empty_cache()
        return res, g

    N, steps = args.adj_mesh, args.adj_steps
    out = {"workload": f"weak-field GR PM (elliptic B), {N}^3 particles on {N}^3 mesh, L={2 * N} Mpc/h, {steps} "
                       "cycles z=50->0, mixed form, loss sum (rho_final - rho_target)^2, gradient w.r.t. the "
                       "initial displacement potential"}
    res, _ = timed(N, steps, gevpm.GEV_MIXED, 1)
    out.update(res)
    N5, steps5 = 2 * N, 2 * steps
    torch.cuda.empty_cache()
    base = used_bytes()
    peak = [0]
    res5, g5 = timed(N5, steps5, gevpm.GEV_MIXED, 4, peak)
    res5.update(workload=f"weak-field GR PM, {N5}^3 particles on {N5}^3 mesh, L={2 * N5} Mpc/h, {steps5} cycles "
                         "z=50->0, mixed form, a checkpoint every 4 cycles",
                peak_device_bytes=peak[0] - base)
    if checks:
        _, g5d = timed(N5, steps5, gevpm.GEV_FP64, 4)
        h = N5 // 2
        crops = [float(np.linalg.norm(g5[z:z + h, y:y + h, x:x + h] - g5d[z:z + h, y:y + h, x:x + h]) /
                       np.linalg.norm(g5d[z:z + h, y:y + h, x:x + h]))
                 for z in (0, h) for y in (0, h) for x in (0, h)]
        res5.update(mixed_vs_fp64_grad_rel_l2=float(np.linalg.norm(g5 - g5d) / np.linalg.norm(g5d)), crop_size=h,
                    crops_rel_l2_max=max(crops), crops_rel_l2_median=float(np.median(crops)))
    out["checkpointed"] = res5
    log(f"GR adjoint: {out}")
    return out


def adjoint_leg(args, checks=True):
    """Differentiable Newtonian run at adj-mesh^3, L = 2 adj-mesh Mpc/h, adj-steps
    cycles z=50 -> 0, loss = sum (rho_final - rho_target)^2 with the target from
    another seed.  Reports the backward / forward time ratio of the mixed form
    (backward includes the per-cycle recompute), the FP32 gradient against the
    FP64 one, and the FP64 gradient against central differences on adj-modes
    random single Fourier modes of psi (eps = 1e-3 of the mode's own scale)."""
    import numpy as np
    import torch
    import gevpm
    from gevpm.adjoint import PMRun, pm_density
    N, steps = args.adj_mesh, args.adj_steps
    rng = np.random.default_rng(7)

    def smooth(seed, amp, N=N):
        r = np.random.default_rng(seed)
        F = np.fft.rfftn(r.standard_normal((N, N, N)))
        kk = np.fft.fftfreq(N, 1.0 / N)
        kz, ky, kx = np.meshgrid(kk, kk, np.arange(N // 2 + 1), indexing="ij")
        F *= np.exp(-(kx ** 2 + ky ** 2 + kz ** 2) / (2 * (N / 16.0) ** 2))
        f = np.fft.irfftn(F, s=(N, N, N), axes=(0, 1, 2))
        return amp * f / f.std()

    amp = 0.3 / N ** 2  # rms displacement ~ a few tenths of a cell at z = 50
    psi_np, tgt_np = smooth(11, amp), smooth(12, amp)
    out = {"workload": f"Newtonian PM, {N}^3 particles on {N}^3 mesh, L={2 * N} Mpc/h, {steps} cycles z=50->0, "
                       "loss sum (rho_final - rho_target)^2, gradient w.r.t. the initial displacement potential"}

    def run_grad(prec):
        run = PMRun(N, 2.0 * N, 50.0, 0.0, steps, precision=prec, ckpt_every=1)
        dev = run.ctx.device
        target = run.forward(torch.from_numpy(tgt_np).to(run.ctx.rdtype).to(dev)).double().clone()
        psi = torch.from_numpy(psi_np).to(run.ctx.rdtype).to(dev).reshape(1, N, N, N)
        # warm-up, then timed forward and backward
        rho = run.forward(psi)
        run.backward(2 * (rho.double() - target))
        torch.cuda.synchronize()
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record()
        rho = run.forward(psi)
        e[1].record()
        g = run.backward(2 * (rho.double() - target))
        e[2].record()
        torch.cuda.synchronize()
        return run, target, g.double().cpu().numpy().reshape(N, N, N), e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2])

    run32, _, g32, fwd_ms, bwd_ms = run_grad(gevpm.GEV_MIXED)
    out.update(forward_ms=fwd_ms, backward_ms=bwd_ms, adjoint_over_forward=bwd_ms / fwd_ms,
               checkpoint_bytes=run32.checkpoint_bytes())
    del run32
    torch.cuda.empty_cache()

    # the larger adjoint workload: twice the mesh, twice the cycles, a checkpoint
    # every 4 cycles (the backward replays up to 3 cycles per segment)
    N5, steps5 = 2 * N, 2 * steps
    amp5 = 0.3 / N5 ** 2
    def used_bytes():
        free, total = torch.cuda.mem_get_info()
        return total - free

    torch.cuda.empty_cache()
    base5 = used_bytes()
    run5 = PMRun(N5, 2.0 * N5, 50.0, 0.0, steps5, precision=gevpm.GEV_MIXED, ckpt_every=4)
    dev5 = run5.ctx.device
    t5 = run5.forward(torch.from_numpy(smooth(12, amp5, N5)).to(run5.ctx.rdtype).to(dev5)).double().clone()
    psi5 = torch.from_numpy(smooth(11, amp5, N5)).to(run5.ctx.rdtype).to(dev5).reshape(1, N5, N5, N5)
    run5.backward(2 * (run5.forward(psi5).double() - t5))
    torch.cuda.synchronize()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    peak = [0]
    run5.mem_probe = lambda: peak.__setitem__(0, max(peak[0], used_bytes()))
    e[0].record()
    rho5 = run5.forward(psi5)
    e[1].record()
    g5 = run5.backward(2 * (rho5.double() - t5))
    e[2].record()
    torch.cuda.synchronize()
    g5 = g5.double().cpu().numpy().reshape(N5, N5, N5)
    out["checkpointed"] = {
        "workload": f"Newtonian PM, {N5}^3 particles on {N5}^3 mesh, L={2 * N5} Mpc/h, {steps5} cycles z=50->0, "
                    "mixed form, a checkpoint every 4 cycles",
        "forward_ms": e[0].elapsed_time(e[1]), "backward_ms": e[1].elapsed_time(e[2]),
        "adjoint_over_forward": e[1].elapsed_time(e[2]) / e[0].elapsed_time(e[1]),
        "checkpoint_bytes": run5.checkpoint_bytes(),
        "peak_device_bytes": peak[0] - base5,
        "peak_note": "device memory in use (hipMemGetInfo) above what was in use before the run was built, sampled "
                     "after the forward and after each backward segment's replay"}
    log(f"adjoint {N5}^3 checkpointed: {out['checkpointed']}")
    del run5, rho5, t5
    torch.cuda.empty_cache()
    if checks:
        # the same gradient from the FP64 form, compared on the eight 128^3 crops
        run5d = PMRun(N5, 2.0 * N5, 50.0, 0.0, steps5, precision=gevpm.GEV_FP64, ckpt_every=4)
        t5d = run5d.forward(torch.from_numpy(smooth(12, amp5, N5)).to(dev5)).double().clone()
        g5d = run5d.backward(2 * (run5d.forward(psi5.double()).double() - t5d)).double().cpu().numpy()
        g5d = g5d.reshape(N5, N5, N5)
        h = N5 // 2
        crops = []
        for cz in (0, h):
            for cy in (0, h):
                for cx_ in (0, h):
                    a_ = g5[cz:cz + h, cy:cy + h, cx_:cx_ + h]
                    b_ = g5d[cz:cz + h, cy:cy + h, cx_:cx_ + h]
                    crops.append(float(np.linalg.norm(a_ - b_) / np.linalg.norm(b_)))
        out["checkpointed"].update(mixed_vs_fp64_grad_rel_l2=float(np.linalg.norm(g5 - g5d) / np.linalg.norm(g5d)),
                                   crop_size=h, crops_rel_l2_max=max(crops),
                                   crops_rel_l2_median=float(np.median(crops)))
        del run5d, t5d, g5d
        torch.cuda.empty_cache()
    del psi5, g5
    if not checks:
        return out
    run64, target, g64, f64_ms, b64_ms = run_grad(gevpm.GEV_FP64)
    out.update(fp64_forward_ms=f64_ms, fp64_backward_ms=b64_ms,
               fp32_vs_fp64_grad_rel_l2=float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64)))
    # central differences on random single modes, at three step sizes; for
    # each, the particles whose final cell differs between the + and - runs
    # (a lower bound on the cell-face crossings that put kinks of the CIC loss
    # inside the difference interval)
    x = np.arange(N)
    dev = run64.ctx.device

    def loss(p):
        rho = run64.forward(torch.from_numpy(p).to(dev).reshape(1, N, N, N)).double()
        pos, _, pid = run64.pcls.store()
        cells = torch.floor(pos[:, torch.argsort(pid)] * N)
        return float(((rho - target) ** 2).sum()), cells

    eps_list = (1e-3, 1e-6, 1e-9)
    errs = {r: [] for r in eps_list}
    cross = {r: [] for r in eps_list}
    for _ in range(args.adj_modes):
        kv = rng.integers(-4, 5, size=3)
        ph = rng.uniform(0, 2 * np.pi)
        mode = np.cos(2 * np.pi * (kv[0] * x[None, None, :] + kv[1] * x[None, :, None] + kv[2] * x[:, None, None]) / N
                      + ph)
        ad = float((g64 * mode).sum())
        for rel in eps_list:
            eps = rel * amp
            (lp, cp), (lm, cm_) = loss(psi_np + eps * mode), loss(psi_np - eps * mode)
            fd = (lp - lm) / (2 * eps)
            errs[rel].append(abs(ad - fd) / max(abs(fd), 1e-300))
            cross[rel].append(int((cp != cm_).any(0).sum()))
    out.update(fd_modes=int(args.adj_modes), fd={
        f"{rel:g}": {"rel_err_median": float(np.median(errs[rel])), "rel_err_max": float(np.max(errs[rel])),
                     "frac_within_1e5": float(np.mean(np.array(errs[rel]) <= 1e-5)),
                     "final_cell_changes_median": float(np.median(cross[rel])),
                     "modes_without_final_cell_change": int(np.sum(np.array(cross[rel]) == 0)),
                     "rel_err_median_without_cell_change": (float(np.median(
                         [e_ for e_, c_ in zip(errs[rel], cross[rel]) if c_ == 0]))
                         if any(c_ == 0 for c_ in cross[rel]) else None)}
        for rel in eps_list},
        fd_note="eps is relative to the rms of psi; the exact check of the gradient is the dot-product test "
                "against the oracle's forward-mode tangent (tests/test_pm_adjoint_dot.py)")
    return out


def cpu_baseline(args, om, ol, k, P):
    """The FP64 oracle (oracle/, a C + OpenMP restatement of the reference
    step) timed on this host: cpu-steps GR cycles of the cpu-mesh^3 workload
    (by default the benchmark's own 512^3 mesh) from the same IC.  The
    reference CPU build (LATfield2 + FFTW + MPI) cannot be built here."""
    import numpy as np
    import torch
    import gevpm
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    O.build()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.set_threads(threads)
    N, L = args.cpu_mesh, args.boxsize * args.cpu_mesh / args.mesh
    fpg = gevpm.fourpiG_of(L)
    c = gevpm.Context(N, gevpm.GEV_FP64)
    p = gevpm.Particles(c, N ** 3, om / N ** 3)
    f = gevpm.Fields(c)
    st = gevpm.StepState(1, fpg, om, ol, z_relax=args.z_in)
    gevpm.ic_generate(p, f, st, 42, L, args.z_in, k, P)
    pos, q, _ = (t.cpu().numpy() for t in p.store())
    M = N ** 3
    ost = dict(pos=np.ascontiguousarray(pos), q=np.ascontiguousarray(q), phi=f.phi.cpu().numpy().ravel().copy(),
               chi=f.chi.cpu().numpy().ravel().copy(), Bi=f.Bi.cpu().numpy().reshape(3, -1).copy(),
               source=np.zeros(M), Sij=np.zeros((6, M)), a=st.a, tau=st.tau)
    del p, f, c
    torch.cuda.empty_cache()
    dtau = (gevpm.particle_horizon(1.0, fpg, om, ol) - st.tau) / args.total_steps
    ost["dtau_old"] = dtau  # a regular cycle (the first cycle of a run skips the phi solve)
    t0 = time.time()
    for _ in range(args.cpu_steps):
        ost["dtau"] = dtau
        O.step(N, ost, 1, om / M, fpg, om, ol, args.z_in)
    dt = time.time() - t0
    return {"value": M * args.cpu_steps / dt, "unit": "particle-steps/s", "cores": O.get_threads(),
            "kind": "port", "sample": f"{args.cpu_steps} GR cycles from the z={args.z_in:g} IC, {N}^3 particles "
                                      f"on {N}^3 mesh, L={L:g} Mpc/h, FP64 oracle, {dt:.1f}s"}


if __name__ == "__main__":
    main()
