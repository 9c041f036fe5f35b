"""ctypes binding of libgevpm.so (the C ABI declared in include/gev_pm.h).

The library is loaded from the in-tree build (gevolution-2.0_amd/lib); there is
no CPU fallback: if the shared object is missing or fails to load, importing
this module raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "..", "lib", "libgevpm.so"))

GEV_OK, GEV_ERR_ARG, GEV_ERR_HIP, GEV_ERR_STATE, GEV_ERR_NOMEM = 0, -1, -2, -3, -4
GEV_FP64, GEV_MIXED = 0, 1


class GevError(RuntimeError):
    pass


class gev_pcls_view(C.Structure):
    _fields_ = [("n", C.c_int64), ("precision", C.c_int), ("pos", C.c_void_p * 3), ("q", C.c_void_p * 3),
                ("id", C.c_void_p), ("cell", C.c_void_p), ("cell_start", C.c_void_p)]


class gev_fields(C.Structure):
    _fields_ = [("phi", C.c_void_p), ("chi", C.c_void_p), ("Bi", C.c_void_p), ("source", C.c_void_p),
                ("Sij", C.c_void_p), ("scalarFT", C.c_void_p), ("SijFT", C.c_void_p), ("BiFT", C.c_void_p)]


class gev_gadget2_header(C.Structure):
    """gadget2_header (metadata.hpp:159-178), 256 bytes"""
    _fields_ = [("npart", C.c_uint32 * 6), ("mass", C.c_double * 6), ("time", C.c_double), ("redshift", C.c_double),
                ("flag_sfr", C.c_int32), ("flag_feedback", C.c_int32), ("npartTotal", C.c_uint32 * 6),
                ("flag_cooling", C.c_int32), ("num_files", C.c_int32), ("BoxSize", C.c_double),
                ("Omega0", C.c_double), ("OmegaLambda", C.c_double), ("HubbleParam", C.c_double),
                ("flag_age", C.c_int32), ("flag_metals", C.c_int32), ("npartTotalHW", C.c_uint32 * 6),
                ("fill", C.c_char * 64)]


class gev_lightcone_geometry(C.Structure):
    """lightcone_geometry (metadata.hpp:205-212)"""
    _fields_ = [("vertex", C.c_double * 3), ("z", C.c_double), ("direction", C.c_double * 3),
                ("opening", C.c_double), ("distance", C.c_double * 2)]


class gev_step_state(C.Structure):
    _fields_ = [("gr", C.c_int), ("vector_elliptic", C.c_int), ("fourpiG", C.c_double), ("Omega_m", C.c_double),
                ("Omega_Lambda", C.c_double), ("Omega_rad", C.c_double), ("z_relax", C.c_double),
                ("a", C.c_double), ("tau", C.c_double), ("dtau", C.c_double), ("dtau_old", C.c_double),
                ("maxvel", C.c_double), ("T00hom", C.c_double), ("phi_hom", C.c_double)]


# name -> (restype, argtypes); every symbol declared in include/gev_pm.h
_vp, _i, _d, _i64, _u32 = C.c_void_p, C.c_int, C.c_double, C.c_int64, C.c_uint32
_dp = C.POINTER(C.c_double)
SIGNATURES = {
    "gev_last_error": (C.c_char_p, []),
    "gev_version": (_i, []),
    "gev_ctx_create": (_i, [_i, _i, _vp, C.POINTER(_vp)]),
    "gev_ctx_destroy": (_i, [_vp]),
    "gev_ctx_sync": (_i, [_vp]),
    "gev_ctx_kpitch": (_i, [_vp]),
    "gev_ctx_set_stream": (_i, [_vp, _vp]),
    "gev_pcls_create": (_i, [_vp, _i64, _d, C.POINTER(_vp)]),
    "gev_pcls_destroy": (_i, [_vp]),
    "gev_pcls_load": (_i, [_vp, _i64, _vp, _vp, _vp, _i]),
    "gev_pcls_store": (_i, [_vp, _vp, _vp, _vp, _i]),
    "gev_pcls_view_get": (_i, [_vp, C.POINTER(gev_pcls_view)]),
    "gev_pcls_sort": (_i, [_vp]),
    "gev_pcls_init_lattice": (_i, [_vp]),
    "gev_pcls_copy_index": (_i, [_vp, _vp, _vp]),
    "gev_project_T00": (_i, [_vp, _vp, _d, _vp, _d, _i]),
    "gev_project_mass": (_i, [_vp, _vp, _d, _i]),
    "gev_project_T0i": (_i, [_vp, _vp, _vp, _d, _i]),
    "gev_project_Tij": (_i, [_vp, _vp, _d, _vp, _d, _i]),
    "gev_project_T00_Tij": (_i, [_vp, _vp, _vp, _d, _vp, _d, _i]),
    "gev_project_T00_Tij_T0i_links": (_i, [_vp, _vp, _vp, _vp, _d, _vp, _d]),
    "gev_prepare_source_T00": (_i, [_vp, _vp, _vp, _vp, _d, _vp, _d, _d, _d, _dp, _dp]),
    "gev_prepare_source_Tij": (_i, [_vp, _vp, _vp, _vp, _d]),
    "gev_fft_r2c": (_i, [_vp, _vp, _vp, _i]),
    "gev_fft_c2r": (_i, [_vp, _vp, _vp, _i]),
    "gev_poisson": (_i, [_vp, _vp, _vp, _d, _d]),
    "gev_project_ft_scalar": (_i, [_vp, _vp, _vp, _i]),
    "gev_project_ft_vector": (_i, [_vp, _vp, _vp, _d, _d]),
    "gev_evolve_ft_vector": (_i, [_vp, _vp, _vp, _d]),
    "gev_set_k0": (_i, [_vp, _vp, _d, _d]),
    "gev_solve_poisson": (_i, [_vp, _vp, _vp, _vp, _d, _d, _i, _d]),
    "gev_solve_chi": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "gev_solve_chi_Tij": (_i, [_vp, _vp, _vp, _d, _vp, _vp, _vp]),
    "gev_solve_vector": (_i, [_vp, _vp, _vp, _d, _d]),
    "gev_solve_vector_T0i_links": (_i, [_vp, _vp, _vp, _vp, _d, _d]),
    "gev_solve_poisson_T00": (_i, [_vp, _vp, _vp, _vp, _d, _d, _d, _d, _vp, _vp, _d, _d, _i, _dp]),
    "gev_ctx_sums": (_i, [_vp, _dp]),
    "gev_update_q": (_i, [_vp, _d, _vp, _vp, _vp, _i, _d, _d, _dp]),
    "gev_update_q_newton": (_i, [_vp, _d, _vp, _vp, _i, _d, _dp]),
    "gev_update_pos": (_i, [_vp, _d, _vp, _vp, _vp, _i, _d, _d]),
    "gev_update_pos_newton": (_i, [_vp, _d, _d]),
    "gev_kick_drift": (_i, [_vp, _i, _d, _d, _vp, _vp, _vp, _i, _i, _d, _d, _d, _d, _dp, _i]),
    "gev_pcls_rebin_counted": (_i, [_vp]),
    "gev_pcls_state_bytes": (_i, [_vp, C.POINTER(C.c_int64)]),
    "gev_pcls_save": (_i, [_vp, _vp, C.c_int64]),
    "gev_pcls_restore": (_i, [_vp, _vp, C.c_int64]),
    "gev_pcls_save_gadget2": (_i, [_vp, C.c_char_p, C.POINTER(gev_gadget2_header), _i, _d, _d, _vp]),
    "gev_pcls_load_gadget2": (_i, [_vp, C.c_char_p, C.POINTER(gev_gadget2_header)]),
    "gev_pcls_save_express": (_i, [_vp, C.c_char_p, C.POINTER(gev_gadget2_header)]),
    "gev_pcls_load_express": (_i, [_vp, C.c_char_p, C.POINTER(gev_gadget2_header)]),
    "gev_lightcone_vertices": (_i, [C.POINTER(gev_lightcone_geometry), _d, _d, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), _i, C.POINTER(C.c_int)]),
    "gev_idlog_create": (_i, [_vp, C.POINTER(_vp)]),
    "gev_idlog_destroy": (_i, [_vp]),
    "gev_idlog_sizes": (_i, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "gev_idlog_get": (_i, [_vp, _i, C.POINTER(C.c_int64)]),
    "gev_idlog_set_backlog": (_i, [_vp, C.POINTER(C.c_int64), C.c_int64]),
    "gev_idlog_commit": (_i, [_vp]),
    "gev_pcls_save_gadget2_lightcone": (_i, [_vp, C.c_char_p, C.POINTER(gev_gadget2_header),
                                             C.POINTER(gev_lightcone_geometry), _d, _d, _d, _d,
                                             C.POINTER(C.c_double), _i, _vp, _vp, _i, C.POINTER(C.c_int64)]),
    "gev_power_spectrum": (_i, [_vp, _vp, _i, _i, _dp, _dp, _dp, _dp, C.POINTER(C.c_int64)]),
    "gev_power_spectrum_noalloc": (_i, [_vp, _vp, _i, _i, _dp, _dp, _dp, _dp, C.POINTER(C.c_int64)]),
    "gev_cross_spectrum": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _dp, _dp, _dp, _dp,
                                C.POINTER(C.c_int64)]),
    "gev_adj_deposit": (_i, [_vp, _vp, _vp]),
    "gev_adj_divergence": (_i, [_vp, _vp, _vp, _d]),
    "gev_adj_gather": (_i, [_vp, _vp, _vp, _vp, _vp, _d, _d, _d]),
    "gev_adj_ic_lattice": (_i, [_vp, _vp, _vp, _d, _d, _vp]),
    "gev_adj_gr_kick_drift": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _d, _d]),
    "gev_adj_gr_deposit": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d]),
    "gev_adj_source_T00": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _d, _d]),
    "gev_adj_source_Tij": (_i, [_vp, _vp, _vp, _vp]),
    "gev_adj_link_factor": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "gev_adj_project_ft_scalar": (_i, [_vp, _vp, _vp]),
    "gev_ic_displacement_ft": (_i, [_vp, _u32, _i, _d, _i, _dp, _dp, _vp, _vp]),
    "gev_displace_particles": (_i, [_vp, _vp, _d, _dp]),
    "gev_initialize_q": (_i, [_vp, _vp, _d, _dp]),
    "gev_hconf": (_d, [_d, _d, _d, _d, _d]),
    "gev_rk4bg": (_d, [_d, _d, _d, _d, _d, _d]),
    "gev_particle_horizon": (_d, [_d, _d, _d, _d, _d]),
    "gev_step": (_i, [_vp, C.POINTER(gev_fields), C.POINTER(gev_step_state)]),
    "gev_step_timing": (_i, [_vp, _i]),
    "gev_step_timings": (_i, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_float), _i]),
    "gev_ic_generate": (_i, [_vp, C.POINTER(gev_fields), C.POINTER(gev_step_state), _u32, _i, _d, _d, _i, _dp, _dp,
                             _dp]),
    "gev_field_sum": (_i, [_vp, _vp, _i, _dp]),
    "gev_field_fill": (_i, [_vp, _vp, _i64, _d]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libgevpm.so not built ({LIB_PATH}); run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def check(rc, what=""):
    if rc != GEV_OK:
        msg = lib.gev_last_error().decode(errors="replace")
        raise GevError(f"{what} failed ({rc}): {msg}")
    return rc
