"""ctypes binding of libctprj.so (include/ctprj.h).

The library is the only compute path: loading fails loudly when it has not been built,
and every call raises ``CtprjError`` on a non-zero status.  ``torch`` is imported first
so that the library shares torch's HIP runtime (both carry soname libamdhip64.so.7).
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (load torch's HIP runtime before libctprj)

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libctprj.so")

FWD_BIT = 1 << 1
BWD_BIT = 1 << 2
FBP_BIT = 1 << 3
MAX_E = 32

CTPRJ_OK = 0
CTPRJ_EINVAL = -1
CTPRJ_ENODEV = -2
CTPRJ_ENOMEM = -3
CTPRJ_EHIP = -4
CTPRJ_ESTATE = -5

_NAMES = {CTPRJ_EINVAL: "EINVAL", CTPRJ_ENODEV: "ENODEV", CTPRJ_ENOMEM: "ENOMEM",
          CTPRJ_EHIP: "EHIP", CTPRJ_ESTATE: "ESTATE"}


class CtprjError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ctprj {_NAMES.get(code, code)}: {msg}")
        self.code = code


class Geom(ctypes.Structure):
    """struct ctprj_geom: the fields of the reference's struct prjConf."""
    _fields_ = [("n", ctypes.c_int), ("prjWidth", ctypes.c_int), ("np", ctypes.c_int),
                ("prjFull", ctypes.c_int), ("dSize", ctypes.c_float),
                ("effectiveRate", ctypes.c_float), ("d", ctypes.c_float)]

    def as_tuple(self):
        return (self.n, self.prjWidth, self.np, self.prjFull, self.dSize,
                self.effectiveRate, self.d)


class PnpgInfo(ctypes.Structure):
    _fields_ = [("cost", ctypes.c_double), ("t", ctypes.c_double), ("theta", ctypes.c_double),
                ("iterations", ctypes.c_int), ("trials", ctypes.c_int), ("rejects", ctypes.c_int),
                ("restarts", ctypes.c_int), ("last_accepted", ctypes.c_int),
                ("difAlpha", ctypes.c_double), ("admmTol", ctypes.c_double),
                ("innerSearch", ctypes.c_int), ("maxInnerItr", ctypes.c_int),
                ("t_reset", ctypes.c_double), ("resets", ctypes.c_int)]


class IeInfo(ctypes.Structure):
    _fields_ = [("cost", ctypes.c_double), ("decrement", ctypes.c_double),
                ("iterations", ctypes.c_int), ("passes", ctypes.c_int),
                ("converged", ctypes.c_int)]


class Tune(ctypes.Structure):
    _fields_ = [("fwd_block", ctypes.c_int), ("adj_tile", ctypes.c_int),
                ("flags", ctypes.c_int), ("adj_seg", ctypes.c_int), ("fwd_seg", ctypes.c_int),
                ("adj_split", ctypes.c_int), ("fwd_split", ctypes.c_int)]


# (name, restype, argtypes) of every entry point declared in include/ctprj.h
_vp, _i, _f, _ll = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_longlong
_dp = ctypes.POINTER(ctypes.c_double)
_llp = ctypes.POINTER(ctypes.c_longlong)
SIGNATURES = [
    ("prj_setup", _i, [_i, _i, _i, _i, _f, _f, _f]),
    ("prj_show", _i, []),
    ("prj_run", _i, [_vp, _vp, _i]),
    ("prj_get_geom", _i, [ctypes.POINTER(Geom)]),
    ("ctprj_plan_create", _i, [ctypes.POINTER(Geom), _i, ctypes.POINTER(_vp)]),
    ("ctprj_plan_destroy", _i, [_vp]),
    ("ctprj_plan_layout", _i, [_vp, _llp, _llp, _llp, _llp]),
    ("ctprj_plan_nnz", _i, [_vp, _llp, _vp]),
    ("ctprj_plan_modes", _i, [_vp, ctypes.POINTER(ctypes.c_int)]),
    ("ctprj_plan_get_tune", _i, [_vp, ctypes.POINTER(Tune)]),
    ("ctprj_plan_set_tune", _i, [_vp, ctypes.POINTER(Tune)]),
    ("ctprj_plan_autotune", _i, [_vp, _i, ctypes.c_char_p, _vp]),
    ("ctprj_forward", _i, [_vp, _vp, _vp, _vp]),
    ("ctprj_adjoint", _i, [_vp, _vp, _vp, _vp]),
    ("ctprj_fbp", _i, [_vp, _vp, _vp, _vp]),
    ("ctprj_poly_iout", _i, [_i, _dp, _i, _dp, _vp, _ll, _vp, _vp, _vp, _vp]),
    ("ctprj_b1_iout", _i, [_dp, _i, _dp, _vp, _ll, _vp, _vp, _vp, _vp]),
    ("ctprj_plan_set_bhc_model", _i, [_vp, _i, _i]),
    ("ctprj_plan_get_bhc_model", _i, [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    ("ctprj_bhc_grad_alpha", _i, [_vp, _vp, _vp, _dp, _i, _dp, _f, _vp, _vp, _vp, _vp]),
    ("ctprj_bhc_grad_alpha_noalloc", _i, [_vp, _vp, _vp, _dp, _i, _dp, _f, _vp, _vp, _vp, _vp]),
    ("ctprj_bhc_alpha_rays", _i, [_vp, _vp, _vp, _dp, _i, _dp, _f, _vp, _vp, _vp, _vp]),
    ("ctprj_bhc_grad_ie", _i, [_vp, _vp, _vp, _vp, _dp, _i, _dp, _f, _vp, _vp, _vp, _vp]),
    ("ctprj_ie_solve", _i, [_vp, _vp, _vp, _vp, _dp, _i, _dp, _f, _i, ctypes.c_double,
                            ctypes.POINTER(IeInfo), _vp]),
    ("ctprj_poly_measure", _i, [_vp, _vp, _i, _dp, _dp, _dp, _dp, _i, ctypes.c_double, _vp,
                                _dp, _dp, _vp]),
    ("ctprj_gen_beam_harden", _i, [_vp, _vp, _i, _dp, _dp, _dp, _dp, _i, ctypes.c_double, _vp,
                                   _dp, _dp, _vp]),
    ("ctprj_wavelet_create", _i, [_i, _i, _i, _dp, _i, _i, ctypes.POINTER(_vp)]),
    ("ctprj_wavelet_destroy", _i, [_vp]),
    ("ctprj_wavelet_set_masks", _i, [_vp, _vp, _vp]),
    ("ctprj_dwt", _i, [_vp, _vp, _vp, _vp]),
    ("ctprj_idwt", _i, [_vp, _vp, _vp, _vp]),
    ("ctprj_wavelet_l1", _i, [_vp, _vp, _dp, _vp]),
    ("ctprj_prox_wavelet_admm", _i, [_vp, _vp, ctypes.c_double, ctypes.c_double, _i, _vp, _vp,
                                     ctypes.POINTER(ctypes.c_int), _vp]),
    ("ctprj_tv_create", _i, [_i, _i, _i, _i, ctypes.POINTER(_vp)]),
    ("ctprj_tv_destroy", _i, [_vp]),
    ("ctprj_tv_set_mask", _i, [_vp, _vp]),
    ("ctprj_tv_value", _i, [_vp, _vp, _dp, _vp]),
    ("ctprj_tv_denoise", _i, [_vp, _vp, ctypes.c_double, ctypes.c_double, _i, _vp,
                              ctypes.POINTER(ctypes.c_int), _vp]),
    ("ctprj_pnpg_create", _i, [_vp, _vp, _dp, _i, _dp, _f, _vp, ctypes.c_double,
                               ctypes.POINTER(_vp)]),
    ("ctprj_pnpg_step", _i, [_vp, _vp, _vp]),
    ("ctprj_pnpg_set_method", _i, [_vp, _i, _i]),
    ("ctprj_pnpg_set_mask", _i, [_vp, _vp]),
    ("ctprj_pnpg_set_wavelet", _i, [_vp, _vp, ctypes.c_double]),
    ("ctprj_pnpg_set_tv", _i, [_vp, _vp, ctypes.c_double]),
    ("ctprj_pnpg_set_linesearch", _i, [_vp, _i]),
    ("ctprj_pnpg_stats", _i, [_vp, ctypes.POINTER(PnpgInfo), _vp]),
    ("ctprj_pnpg_destroy", _i, [_vp]),
    ("ctprj_pnpg_snapshot_bytes", _i, [_vp, ctypes.POINTER(ctypes.c_longlong)]),
    ("ctprj_pnpg_save", _i, [_vp, _vp, ctypes.c_longlong, _vp]),
    ("ctprj_pnpg_load", _i, [_vp, _vp, ctypes.c_longlong, _vp]),
    ("ctprj_last_error", ctypes.c_char_p, []),
    ("ctprj_device_arch", _i, [ctypes.c_char_p, _i]),
    ("ctprj_version", ctypes.c_char_p, []),
    ("ctprj_debug_fail_alloc", _i, [_i]),
    ("ctprj_debug_pnpg_decide", _i, [_vp, _i, _i, _vp, ctypes.POINTER(ctypes.c_int)]),
]

_lib = None


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m imgrecsrc_amd.build` "
                "(there is no fallback implementation)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().ctprj_last_error()
        raise CtprjError(rc, msg.decode() if msg else "")


def dbl_array(values):
    arr = (ctypes.c_double * len(values))(*[float(v) for v in values])
    return ctypes.cast(arr, _dp), arr
