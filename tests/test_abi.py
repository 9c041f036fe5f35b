"""The C-ABI library builds, loads and exports every function include/gev_pm.h
declares; host-only logic (cosmology table, argument checks that need no
device) runs on the CPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gev_pm.h")
LIB = os.path.join(ROOT, "gevolution-2.0_amd", "lib", "libgevpm.so")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gev_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "run __graft_entry__.build() first"
    lib = ctypes.CDLL(LIB)
    names = declared()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (gev_\w+)", out))
    assert set(names) <= exported


def test_python_binding_covers_header():
    from gevpm import _lib
    assert set(declared()) <= set(_lib.SIGNATURES)


def test_background_host_functions_match_oracle(O):
    from gevpm import _lib
    for a in (0.01, 0.3, 1.0):
        args = (a, 0.7, 0.31, 0.69, 0.0)
        assert _lib.lib.gev_hconf(*args) == O.hconf(a, 0.7, 0.31, 0.69)
        assert np.isclose(_lib.lib.gev_particle_horizon(*args), O.particle_horizon(a, 0.7, 0.31, 0.69), rtol=1e-13)
        assert _lib.lib.gev_rk4bg(a, 0.7, 0.31, 0.69, 0.0, 0.01) == O.rk4bg(a, 0.7, 0.31, 0.69, 0.01)


def test_bad_arguments_fail_loudly_without_gpu():
    from gevpm import _lib
    h = ctypes.c_void_p()
    rc = _lib.lib.gev_ctx_create(12, 0, None, ctypes.byref(h))  # not a power of two
    assert rc == _lib.GEV_ERR_ARG
    assert b"power of two" in _lib.lib.gev_last_error()
    rc = _lib.lib.gev_ctx_create(16, 7, None, ctypes.byref(h))
    assert rc == _lib.GEV_ERR_ARG


def test_bbks_table_normalisation():
    import gevpm.cosmology as cosmo
    k, P = cosmo.bbks_table(z=0.0)
    assert np.isclose(cosmo.sigma_R(k, P, 8.0), 0.81, rtol=1e-10)
    # EdS growth is D = a
    assert np.isclose(cosmo.growth_factor(0.25, 1.0, 0.0), 0.25, rtol=1e-10)
    k2, P2 = cosmo.bbks_table(z=100.0)
    D = cosmo.growth_factor(1 / 101.0, 0.31, 0.69)
    assert np.allclose(P2, P * D * D, rtol=1e-12)


def test_checked_pointers_refuse_mismatches():
    """the Python layer checks dtype, contiguity, element count and device of
    every tensor before its pointer reaches a kernel"""
    import torch
    from gevpm import GevError, _checked
    good = torch.zeros(2, 8, dtype=torch.float32)
    with pytest.raises(GevError, match="expected torch.float32"):
        _checked(good.double(), "f", torch.float32, 16, None)
    with pytest.raises(GevError, match="contiguous"):
        _checked(good.t(), "f", torch.float32, 16, None)
    with pytest.raises(GevError, match="expected 24 elements"):
        _checked(good, "f", torch.float32, 24, None)
    with pytest.raises(GevError, match="at least 24"):
        _checked(good, "f", torch.float32, 24, None, at_least=True)
    with pytest.raises(GevError, match="a tensor is required"):
        _checked(None, "f", torch.float32, 16, None)
    with pytest.raises(GevError, match="torch tensor"):
        _checked(np.zeros(16, np.float32), "f", torch.float32, 16, None)
    with pytest.raises(GevError, match="must live on"):
        _checked(good, "f", torch.float32, 16, None)  # right in every way but on the host
    assert _checked(None, "f", torch.float32, 16, None, optional=True) is None
