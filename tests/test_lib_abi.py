"""The C-ABI library: it builds, loads, exports every entry point include/ctprj.h
declares, and its argument / state errors come back as status codes (no GPU needed)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "ctprj.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*[\s\*]+([a-z_][a-z0-9_]*)\s*\(",
                       txt, flags=re.M)
    return sorted(set(n for n in names if n.startswith(("prj_", "ctprj_"))))


@pytest.fixture(scope="module")
def L():
    from imgrecsrc_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from imgrecsrc_amd.build import build
        build()
    return _lib.lib()


def test_header_lists_entry_points():
    names = declared_functions()
    assert "prj_setup" in names and "prj_run" in names and "ctprj_forward" in names
    assert len(names) >= 20


def test_every_declared_symbol_is_exported(L):
    from imgrecsrc_amd import _lib
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (\w+)$", out, flags=re.M))
    for name in declared_functions():
        assert name in exported, name
        assert hasattr(L, name)
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert bound == set(declared_functions())


def test_version_and_state_errors(L):
    from imgrecsrc_amd import _lib
    assert b"gfx950" in L.ctprj_version()
    img = np.zeros(4, np.float32)
    rc = L.prj_run(img.ctypes.data, img.ctypes.data, _lib.FWD_BIT)
    assert rc in (_lib.CTPRJ_ESTATE,)
    assert b"prj_setup" in L.ctprj_last_error()


def test_argument_errors_without_device(L):
    from imgrecsrc_amd import _lib
    kn, _h = _lib.dbl_array([1.0, 2.0, 3.0])
    # E out of range is rejected before any device work
    assert L.ctprj_b1_iout(kn, 0, None, None, 0, None, None, None, None) == _lib.CTPRJ_EINVAL
    assert L.ctprj_b1_iout(kn, 33, None, None, 0, None, None, None, None) == _lib.CTPRJ_EINVAL
    bad, _h2 = _lib.dbl_array([1.0, 1.0, 3.0])
    assert L.ctprj_b1_iout(bad, 1, None, None, 0, None, None, None, None) == _lib.CTPRJ_EINVAL
    assert L.ctprj_plan_create(None, 1, None) == _lib.CTPRJ_EINVAL
    g = _lib.Geom(64, 95, 36, 72, 1.0, 1.0, 20.0)  # fan source inside the image
    h = ctypes.c_void_p()
    assert L.ctprj_plan_create(ctypes.byref(g), 1, ctypes.byref(h)) == _lib.CTPRJ_EINVAL
    assert b"inside the image" in L.ctprj_last_error()
    g = _lib.Geom(64, 95, 36, 72, 0.0, 1.0, 0.0)
    assert L.ctprj_plan_create(ctypes.byref(g), 1, ctypes.byref(h)) == _lib.CTPRJ_EINVAL


def test_failure_injection_hook(L):
    from imgrecsrc_amd import _lib
    assert L.ctprj_debug_fail_alloc(-1) == _lib.CTPRJ_EINVAL
    assert L.ctprj_debug_fail_alloc(0) == 0


def test_no_device_is_reported(L):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from imgrecsrc_amd import _lib
    g = _lib.Geom(16, 23, 8, 16, 1.0, 1.0, 0.0)
    h = ctypes.c_void_p()
    assert L.ctprj_plan_create(ctypes.byref(g), 1, ctypes.byref(h)) == _lib.CTPRJ_ENODEV


def test_mex_argument_checks(L):
    from imgrecsrc_amd import mex
    with pytest.raises(ValueError, match="number of input arguments"):
        mex.gpuPrj(0, {})
    with pytest.raises(ValueError, match="no field"):
        mex.gpuPrj(0, {"n": 4}, "config")


def test_geometry_helpers():
    from imgrecsrc_amd.bhc import basis_at_zero, kappa_logspan, normalize_ie, spline_knots
    k = spline_knots(np.logspace(np.log10(0.03), np.log10(30), 20))
    assert k.size == 22 and np.isclose(k[1] / k[0], k[2] / k[1])
    ie = normalize_ie(np.ones(20), k)
    assert basis_at_zero(k) @ ie == pytest.approx(1.0)
    kl = kappa_logspan(17, 3.0)
    assert kl.size == 17 and np.all(np.diff(np.log(kl)) > 0)
