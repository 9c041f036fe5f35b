"""ctypes binding of the CPU FP64 oracle (oracle/libgevoracle).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU baseline,
never as part of the product path.

Array conventions (numpy):
  real field     float64 [N^3]  (x fastest), components stacked [c, N^3]
  Fourier field  complex128 [N, N, N//2+1] indexed [kz, ky, kx] (numpy rfftn
                 layout of an array indexed [z, y, x]); components stacked first
  particles      pos, q float64 [3, np]; id int64 [np]
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libgevoracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_cp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _opt(a):
    return None if a is None else np.ascontiguousarray(a, np.float64)


def _c(a):
    return np.ascontiguousarray(a, np.float64)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    I, D, Lg, V = C.c_int, C.c_double, C.c_long, None
    vp = C.c_void_p
    sig = {
        "or_version": (I, []),
        "or_set_threads": (V, [I]),
        "or_get_threads": (I, []),
        "or_fft_r2c": (V, [I, _dp, _cp]),
        "or_fft_c2r": (V, [I, _cp, _dp]),
        "or_poisson": (V, [I, _cp, _cp, D, D]),
        "or_project_ft_scalar": (V, [I, _cp, _cp, I]),
        "or_project_ft_vector": (V, [I, _cp, _cp, D, D]),
        "or_evolve_ft_vector": (V, [I, _cp, _cp, D]),
        "or_prepare_source_T00": (D, [I, _dp, _dp, _dp, D, _dp, D, D, D]),
        "or_prepare_source_Tij": (V, [I, _dp, _dp, _dp, D]),
        "or_project_T00": (V, [I, Lg, _dp, _dp, D, _dp, D, vp, D]),
        "or_project_T0i": (V, [I, Lg, _dp, _dp, D, _dp, vp, D]),
        "or_project_Tij": (V, [I, Lg, _dp, _dp, D, _dp, D, vp, D]),
        "or_update_q": (D, [I, Lg, _dp, _dp, D, _dp, vp, vp, I, D, D]),
        "or_update_q_newton": (D, [I, Lg, _dp, _dp, D, _dp, vp, I, D]),
        "or_update_pos": (V, [I, Lg, _dp, _dp, D, vp, vp, vp, I, D, D]),
        "or_update_pos_newton": (V, [I, Lg, _dp, _dp, D, D]),
        "or_displace_particles": (D, [I, Lg, _dp, _dp, D]),
        "or_initialize_q": (D, [I, Lg, _dp, _dp, _dp, D]),
        "or_hconf": (D, [D, D, D, D, D]),
        "or_rk4bg": (D, [D, D, D, D, D, D]),
        "or_particle_horizon": (D, [D, D, D, D, D]),
        "or_philox4x32_10": (V, [_u32p, _u32p, _u32p]),
        "or_spline_eval": (D, [I, _dp, _dp, D]),
        "or_ic_displacement_ft": (V, [I, C.c_uint32, I, D, I, _dp, _dp, _cp, _cp]),
        "or_pk_table": (V, [I, _dp, _dp, D, _dp, _dp]),
        "or_ic_generate": (D, [I, C.c_uint32, I, I, D, D, D, D, D, I, _dp, _dp, _dp, _dp, _ip, _dp, _dp, _dp,
                               _dp, _dp, C.POINTER(C.c_double)]),
        "or_step": (V, [I, Lg, I, D, D, D, D, D, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "or_power_spectrum": (V, [I, _cp, I, I, _dp, _dp, _dp, _dp, _i32p]),
        "or_cross_spectrum": (V, [I, I, _cp, _cp, I, I, I, I, I, I, _dp, _dp, _dp, _dp, _i32p]),
        "or_tlm_deposit_mass": (V, [I, Lg, _dp, _dp, D, _dp]),
        "or_tlm_kick_newton": (V, [I, Lg, _dp, _dp, _dp, _dp, _dp, D]),
        "or_newton_run_tlm": (D, [I, _dp, vp, _dp, I, _dp, D, D, D, D, vp, vp, vp, C.POINTER(C.c_double)]),
        "or_lightcone_vertices": (I, [_dp, _dp, D, D, D, _dp, _dp, I]),
        "or_lightcone_select": (Lg, [I, Lg, _dp, _dp, _ip, vp, I, _dp, _dp, D, D, D, D, D, D, D, _dp, I, vp, Lg,
                                     _f32p, _f32p, _ip, _u8p, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def set_threads(n):
    lib().or_set_threads(int(n))


def get_threads():
    return lib().or_get_threads()


def fft_r2c(f):
    N = round(len(f) ** (1 / 3))
    out = np.zeros((N, N, N // 2 + 1), np.complex128)
    lib().or_fft_r2c(N, np.ascontiguousarray(f, np.float64), out)
    return out


def fft_c2r(F):
    N = F.shape[0]
    out = np.zeros(N ** 3, np.float64)
    lib().or_fft_c2r(N, np.ascontiguousarray(F, np.complex128), out)
    return out


def poisson(F, coeff, modif=0.0):
    out = np.empty_like(F)
    lib().or_poisson(F.shape[0], np.ascontiguousarray(F), out, coeff, modif)
    return out


def project_ft_scalar(S, chi=None):
    N = S.shape[1]
    add = chi is not None
    out = np.ascontiguousarray(chi.copy()) if add else np.zeros(S.shape[1:], np.complex128)
    lib().or_project_ft_scalar(N, np.ascontiguousarray(S), out, int(add))
    return out


def project_ft_vector(S, coeff=1.0, modif=0.0):
    out = np.empty_like(S)
    lib().or_project_ft_vector(S.shape[1], np.ascontiguousarray(S), out, coeff, modif)
    return out


def evolve_ft_vector(S, B, a2dtau):
    out = np.ascontiguousarray(B.copy())
    lib().or_evolve_ft_vector(S.shape[1], np.ascontiguousarray(S), out, a2dtau)
    return out


def prepare_source_T00(phi, chi, source, bg, c1, c2, c3):
    N = round(len(phi) ** (1 / 3))
    out = np.empty_like(source)
    s = lib().or_prepare_source_T00(N, phi, chi, source, bg, out, c1, c2, c3)
    return out, s


def prepare_source_Tij(phi, Tij, coeff):
    N = round(len(phi) ** (1 / 3))
    out = np.empty_like(Tij)
    lib().or_prepare_source_Tij(N, phi, np.ascontiguousarray(Tij), out, coeff)
    return out


def project_T00(N, pos, q, pmass, a=1.0, phi=None, coeff=1.0, out=None):
    out = np.zeros(N ** 3) if out is None else out
    phi = _opt(phi)
    lib().or_project_T00(N, pos.shape[1], _c(pos), _c(q), pmass, out, a, _ptr(phi), coeff)
    return out


def project_T0i(N, pos, q, pmass, phi=None, coeff=1.0, out=None):
    out = np.zeros((3, N ** 3)) if out is None else out
    phi = _opt(phi)
    lib().or_project_T0i(N, pos.shape[1], _c(pos), _c(q), pmass, out, _ptr(phi), coeff)
    return out


def project_Tij(N, pos, q, pmass, a=1.0, phi=None, coeff=1.0, out=None):
    out = np.zeros((6, N ** 3)) if out is None else out
    phi = _opt(phi)
    lib().or_project_Tij(N, pos.shape[1], _c(pos), _c(q), pmass, out, a, _ptr(phi), coeff)
    return out


def update_q(N, pos, q, dtau, phi, chi, Bi, nfield, a, bcoeff):
    phi, chi, Bi = _opt(phi), _opt(chi), _opt(Bi)
    return lib().or_update_q(N, pos.shape[1], pos, q, dtau, phi, _ptr(chi), _ptr(Bi), nfield, a, bcoeff)


def update_q_newton(N, pos, q, dtau, psi, a, chi=None, nfield=1):
    chi = _opt(chi)
    return lib().or_update_q_newton(N, pos.shape[1], pos, q, dtau, psi, _ptr(chi), nfield, a)


def update_pos(N, pos, q, dtau, phi, chi, Bi, nfield, a, bcoeff):
    phi, chi, Bi = _opt(phi), _opt(chi), _opt(Bi)
    lib().or_update_pos(N, pos.shape[1], pos, q, dtau, _ptr(phi), _ptr(chi), _ptr(Bi), nfield, a, bcoeff)


def update_pos_newton(N, pos, q, dtau, a):
    lib().or_update_pos_newton(N, pos.shape[1], pos, q, dtau, a)


def displace_particles(N, pos, xi, coeff=1.0):
    return lib().or_displace_particles(N, pos.shape[1], pos, xi, coeff)


def initialize_q(N, pos, q, pot, coeff):
    return lib().or_initialize_q(N, pos.shape[1], pos, q, pot, coeff)


def hconf(a, fourpiG, om, ol, orad=0.0):
    return lib().or_hconf(a, fourpiG, om, ol, orad)


def rk4bg(a, fourpiG, om, ol, dtau, orad=0.0):
    return lib().or_rk4bg(a, fourpiG, om, ol, orad, dtau)


def particle_horizon(a, fourpiG, om, ol, orad=0.0):
    return lib().or_particle_horizon(a, fourpiG, om, ol, orad)


def philox4x32_10(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def spline_eval(x, y, xv):
    return lib().or_spline_eval(len(x), np.asarray(x, np.float64), np.asarray(y, np.float64), float(xv))


def ic_generate(N, seed, k_hmpc, P_mpch3, boxsize, fourpiG, om, ol, a, gr=0, ksphere=1):
    M = N ** 3
    pos = np.zeros((3, M)); q = np.zeros((3, M)); pid = np.zeros(M, np.int64)
    phi = np.zeros(M); chi = np.zeros(M); Bi = np.zeros((3, M)); src = np.zeros(M); Sij = np.zeros((6, M))
    md = C.c_double(0.0)
    mv = lib().or_ic_generate(N, seed, ksphere, gr, boxsize, fourpiG, om, ol, a, len(k_hmpc),
                              np.ascontiguousarray(k_hmpc, np.float64), np.ascontiguousarray(P_mpch3, np.float64),
                              pos, q, pid, phi, chi, Bi, src, Sij, C.byref(md))
    return dict(pos=pos, q=q, id=pid, phi=phi, chi=chi, Bi=Bi, source=src, Sij=Sij, maxvel=mv, max_disp=md.value)


def step(N, st, gr, pmass, fourpiG, om, ol, z_relax):
    """One main-loop cycle; st is a dict with arrays pos,q,phi,chi,Bi,source,Sij and
    scalars a,tau,dtau,dtau_old (updated in place)."""
    sc = np.array([st["a"], st["tau"], st["dtau"], st["dtau_old"], st.get("maxvel", 0.0), 0.0, 0.0])
    lib().or_step(N, st["pos"].shape[1], gr, pmass, fourpiG, om, ol, z_relax, st["pos"], st["q"], st["phi"],
                  st["chi"], st["Bi"], st["source"], st["Sij"], sc)
    st["a"], st["tau"], st["dtau_old"], st["maxvel"], st["T00hom"], st["phi_hom"] = sc[0], sc[1], sc[3], sc[4], sc[5], sc[6]
    return st


def power_spectrum(FT, numbins, deconvolve=True):
    N = FT.shape[0]
    kb, pw, ks, ps = (np.zeros(numbins) for _ in range(4))
    occ = np.zeros(numbins, np.int32)
    lib().or_power_spectrum(N, np.ascontiguousarray(FT), numbins, int(deconvolve), kb, pw, ks, ps, occ)
    return kb, pw, ks, ps, occ


def cross_spectrum(F1, F2, numbins, deconvolve=True, ktype=1, comp1=-1, comp2=-1, symmetric=False):
    """extractCrossSpectrum (tools.hpp:65-330).  F1, F2: (ncomp, N, N, N/2+1) or
    (N, N, N/2+1) complex arrays in the [kz, ky, kx] layout; ktype 0 grid, 1 linear."""
    F1 = np.ascontiguousarray(F1, np.complex128)
    F2 = np.ascontiguousarray(F2, np.complex128)
    if F1.ndim == 3:
        F1, F2 = F1[None], F2[None]
    ncomp, N = F1.shape[0], F1.shape[1]
    kb, pw, ks, ps = (np.zeros(numbins) for _ in range(4))
    occ = np.zeros(numbins, np.int32)
    lib().or_cross_spectrum(N, ncomp, F1, F2, int(symmetric), comp1, comp2, numbins, int(deconvolve), ktype, kb, pw,
                            ks, ps, occ)
    return kb, pw, ks, ps, occ


MAX_INTERSECTS = 12
IDCHECK_ZONE = 0.05


def lightcone_vertices(lc_vertex, direction, opening, outer, inner, domain, cap=MAX_INTERSECTS):
    """findIntersectingLightcones: (count, vertices[min(count, cap), 3])"""
    out = np.zeros((max(cap, 1), 3))
    n = lib().or_lightcone_vertices(_c(lc_vertex), _c(direction), float(opening), float(outer), float(inner),
                                    _c(domain), out, int(cap))
    return n, out[:min(n, cap)].copy()


def lightcone_select(N, pos, q, pid, phi, tracer_factor, lc_vertex, direction, opening, dist, dtau, dtau_old,
                     dadtau, a, boxsize, vertices, backlog=None):
    """lightcone form of saveGadget2 for one rank: (pos [n,3] f32, vel [n,3] f32,
    id [n], prelog flag [n] bool, margin [np]) in particle order"""
    pos, q = _c(pos), _c(q)
    np_ = pos.shape[1]
    pid = np.ascontiguousarray(pid, np.int64)
    vertices = _c(np.asarray(vertices, np.float64).reshape(-1, 3))
    bl = None if backlog is None or len(backlog) == 0 else np.ascontiguousarray(backlog, np.int64)
    po = np.zeros((max(np_, 1), 3), np.float32)
    vo = np.zeros((max(np_, 1), 3), np.float32)
    io = np.zeros(max(np_, 1), np.int64)
    pl = np.zeros(max(np_, 1), np.uint8)
    margin = np.zeros(max(np_, 1))
    ph = None if phi is None else _c(phi).reshape(-1)
    n = lib().or_lightcone_select(N, np_, pos, q, pid, _ptr(ph), int(tracer_factor), _c(lc_vertex), _c(direction),
                                  float(opening), float(dist), float(dtau), float(dtau_old), float(dadtau), float(a),
                                  float(boxsize), vertices, len(vertices), _ptr(bl), 0 if bl is None else len(bl),
                                  po, vo, io, pl, margin)
    return po[:n].copy(), vo[:n].copy(), io[:n].copy(), pl[:n].astype(bool), margin[:np_].copy()


# ------------------------------------------------ tangent-linear Newton run
def newton_schedule(N, boxsize, z_in, z_end, steps, om=0.31, ol=0.69):
    """(sched [steps, 4] = a, a_d, dtau_k, dtau per cycle; 4 pi G; cx; cq) of a
    Newtonian run uniform in conformal time from z_in to z_end, the lattice
    Zel'dovich map x = x_lat + cx N grad1(psi), q = cq N grad1(psi) with
    cx = 1 and cq = a f(a) H(a), f = Omega_m(a)^0.55"""
    fpg = 1.5 * boxsize * boxsize / 2997.92458 / 2997.92458
    a = 1.0 / (1.0 + z_in)
    tau0 = particle_horizon(a, fpg, om, ol)
    dtau = (particle_horizon(1.0 / (1.0 + z_end), fpg, om, ol) - tau0) / steps
    om_a = om / (om + ol * a ** 3)
    cq = a * om_a ** 0.55 * hconf(a, fpg, om, ol)
    sched = np.zeros((steps, 4))
    dtau_old = 0.0
    for k in range(steps):
        a_d = rk4bg(a, fpg, om, ol, 0.5 * dtau)
        sched[k] = (a, a_d, 0.5 * (dtau + dtau_old), dtau)
        a = rk4bg(a_d, fpg, om, ol, 0.5 * dtau)
        dtau_old = dtau
    return sched, fpg, 1.0, cq


def newton_run_tlm(N, psi, target, sched, fpg, cx, cq, v=None, om=0.31, want_state=False):
    """L (and dL = <grad L, v> when v is given) of the Newtonian run; with
    want_state also the final rho and particle pos, q indexed by ID"""
    M = N ** 3
    psi = _c(np.asarray(psi).reshape(M))
    target = _c(np.asarray(target).reshape(M))
    vv = None if v is None else _c(np.asarray(v).reshape(M))
    rho = np.zeros(M) if want_state else None
    pos = np.zeros((3, M)) if want_state else None
    q = np.zeros((3, M)) if want_state else None
    dL = C.c_double(0.0)
    L = lib().or_newton_run_tlm(N, psi, _ptr(vv), target, len(sched), _c(sched), fpg, om, cx, cq, _ptr(rho),
                                _ptr(pos), _ptr(q), C.byref(dL))
    out = (L, dL.value if v is not None else None)
    return out + ((rho, pos, q) if want_state else ())
