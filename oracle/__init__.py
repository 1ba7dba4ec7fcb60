"""CPU checkers for the CT projector and the beam-hardening likelihood.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the checker;
the product path (``imgrecsrc_amd``) never imports it.

Contents
--------
* ``liboracle.so`` (``ctprj_oracle.c``): FP64 restatement of the reference projector,
  ``prj/cpuPrj.c:46-806`` + ``prj/utils.h:12-151``, evaluated ray by ray with no angular
  symmetry.
* ``_ref/libcpuprj_ref.so``: the reference's own ``prj/cpuPrj.c`` compiled from
  ``/root/reference`` (``oracle/Makefile``), FP32, 64 pthreads.
* ``bhc_oracle``: FP64 NumPy restatement of ``bhc/Spline.m`` (``b1Iout``, ``b0Iout``,
  ``disIout``), ``bhc/poissLAlpha.m`` and ``bhc/poissLI.m``.

Pinning (see the module headers): golden vectors from the compiled reference, exact
polygon-clipping pixel/strip areas, analytic ellipse line integrals, dense-matrix
transposition on tiny images, quadrature of the spline Laplace transforms and central
finite differences of the likelihoods.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_ORACLE_SO = os.path.join(HERE, "liboracle.so")
_REF_SO = os.path.join(HERE, "_ref", "libcpuprj_ref.so")

FWD_BIT = 1 << 1   # prj/prj.h:30
BWD_BIT = 1 << 2   # prj/prj.h:31
FBP_BIT = 1 << 3   # prj/prj.h:32


class OrcConf(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("prjWidth", ctypes.c_int), ("np", ctypes.c_int),
        ("prjFull", ctypes.c_int), ("dSize", ctypes.c_double),
        ("effectiveRate", ctypes.c_double), ("d", ctypes.c_double),
        ("angle_mode", ctypes.c_int), ("support_mode", ctypes.c_int),
    ]


def build(force: bool = False) -> None:
    """Compile liboracle.so (and _ref when /root/reference exists)."""
    if force or not os.path.exists(_ORACLE_SO):
        subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def build_ref() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_ORACLE_SO):
            build()
        L = ctypes.CDLL(_ORACLE_SO)
        P = ctypes.POINTER(OrcConf)
        dp = ctypes.c_void_p
        L.orc_get_weight.restype = ctypes.c_double
        L.orc_get_weight.argtypes = [ctypes.c_double] * 4
        L.orc_strip_area.restype = ctypes.c_double
        L.orc_strip_area.argtypes = [ctypes.c_double] * 4
        for f in (L.orc_forward, L.orc_adjoint):
            f.restype = ctypes.c_longlong
            f.argtypes = [P, dp, dp, ctypes.c_int]
        L.orc_nnz.restype = ctypes.c_longlong
        L.orc_nnz.argtypes = [P, ctypes.c_int]
        L.orc_dense.restype = ctypes.c_int
        L.orc_dense.argtypes = [P, dp]
        L.orc_fbp.restype = ctypes.c_int
        L.orc_fbp.argtypes = [P, dp, dp, ctypes.c_int]
        L.orc_ramp_filter.restype = None
        L.orc_ramp_filter.argtypes = [dp, ctypes.c_int]
        L.orc_csr_counts.restype = ctypes.c_int
        L.orc_csr_counts.argtypes = [P, dp, ctypes.c_int]
        L.orc_csr_fill.restype = ctypes.c_int
        L.orc_csr_fill.argtypes = [P, dp, dp, dp, ctypes.c_int]
        L.orc_csr_apply.restype = ctypes.c_int
        L.orc_csr_apply.argtypes = [ctypes.c_longlong, ctypes.c_longlong, dp, dp, dp, dp, dp,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


def conf(n, prjWidth, np_, prjFull, dSize=1.0, effectiveRate=1.0, d=0.0,
         angle_mode=1, support_mode=0) -> OrcConf:
    return OrcConf(int(n), int(prjWidth), int(np_), int(prjFull), float(dSize),
                   float(effectiveRate), float(d), int(angle_mode), int(support_mode))


def _nth():
    """Worker threads: the cores this process may use, at most 16."""
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(16, cores or 1))


def get_weight(dist, bw, c, s) -> float:
    return lib().orc_get_weight(float(dist), float(bw), float(c), float(s))


def strip_area(dist, bw, c, s) -> float:
    return lib().orc_strip_area(float(dist), float(bw), float(c), float(s))


def forward(cf: OrcConf, img: np.ndarray):
    """sino[np, prjWidth] (float64) and nnz(A)."""
    img = np.ascontiguousarray(img, dtype=np.float64).reshape(cf.n, cf.n)
    sino = np.zeros((cf.np, cf.prjWidth), np.float64)
    nnz = lib().orc_forward(ctypes.byref(cf), img.ctypes.data, sino.ctypes.data, _nth())
    return sino, nnz


def adjoint(cf: OrcConf, sino: np.ndarray):
    sino = np.ascontiguousarray(sino, dtype=np.float64).reshape(cf.np, cf.prjWidth)
    img = np.zeros((cf.n, cf.n), np.float64)
    nnz = lib().orc_adjoint(ctypes.byref(cf), sino.ctypes.data, img.ctypes.data, _nth())
    return img, nnz


class Matrix:
    """The FP64 system matrix A of one geometry, stored in compressed rows (12 bytes per
    nonzero), for checking many images of the same geometry: one visit of every ray
    builds it, after which forward / adjoint of K stacked images cost a sparse product.
    The stored forward adds each ray's terms in the order orc_forward does, so it equals
    `forward` bit for bit; the stored adjoint equals `adjoint` up to FP64 rounding."""

    def __init__(self, cf: OrcConf):
        self.cf = cf
        self.rows = cf.np * cf.prjWidth
        self.ncols = cf.n * cf.n
        cnt = np.zeros(self.rows, np.int64)
        L = lib()
        L.orc_csr_counts(ctypes.byref(cf), cnt.ctypes.data, _nth())
        self.rowptr = np.zeros(self.rows + 1, np.int64)
        np.cumsum(cnt, out=self.rowptr[1:])
        self.nnz = int(self.rowptr[-1])
        self.col = np.empty(self.nnz, np.int32)
        self.val = np.empty(self.nnz, np.float64)
        L.orc_csr_fill(ctypes.byref(cf), self.rowptr.ctypes.data, self.col.ctypes.data,
                       self.val.ctypes.data, _nth())

    def _apply(self, x, trans):
        n_in, n_out = (self.rows, self.ncols) if trans else (self.ncols, self.rows)
        x = np.ascontiguousarray(x, dtype=np.float64)
        K = x.size // n_in
        assert x.size == K * n_in, "input size is not a whole number of vectors"
        out = np.empty(K * n_out, np.float64)
        rc = lib().orc_csr_apply(self.rows, self.ncols, self.rowptr.ctypes.data,
                                 self.col.ctypes.data, self.val.ctypes.data, x.ctypes.data,
                                 out.ctypes.data, K, int(trans), _nth())
        if rc != 0:
            raise MemoryError("orc_csr_apply: no memory for the per-thread images")
        return out

    def forward(self, img):
        """Sinograms (K, np, prjWidth) of K stacked images (K, n, n) or one (n, n)."""
        out = self._apply(img, False)
        return out.reshape(-1, self.cf.np, self.cf.prjWidth)

    def adjoint(self, sino):
        """Images (K, n, n) of K stacked sinograms."""
        out = self._apply(sino, True)
        return out.reshape(-1, self.cf.n, self.cf.n)


def nnz(cf: OrcConf) -> int:
    return int(lib().orc_nnz(ctypes.byref(cf), _nth()))


def dense(cf: OrcConf) -> np.ndarray:
    A = np.zeros((cf.np * cf.prjWidth, cf.n * cf.n), np.float64)
    lib().orc_dense(ctypes.byref(cf), A.ctypes.data)
    return A


def fbp(cf: OrcConf, sino: np.ndarray) -> np.ndarray:
    sino = np.ascontiguousarray(sino, dtype=np.float64).reshape(cf.np, cf.prjWidth)
    img = np.zeros((cf.n, cf.n), np.float64)
    lib().orc_fbp(ctypes.byref(cf), sino.ctypes.data, img.ctypes.data, _nth())
    return img


def ramp_filter(signal: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(signal, dtype=np.float64).copy()
    lib().orc_ramp_filter(s.ctypes.data, s.size)
    return s


# ----------------------------------------------------------------------------------
# The reference's own cpuPrj (FP32), compiled from /root/reference into _ref/.
# ----------------------------------------------------------------------------------
_ref = None


def ref_available() -> bool:
    return os.path.exists(_REF_SO)


def ref_lib():
    """ctypes handle on the compiled reference (setup/cpuPrj, prj/prj.h:59-70)."""
    global _ref
    if _ref is None:
        L = ctypes.CDLL(_REF_SO)
        L.setup.restype = None
        L.setup.argtypes = [ctypes.c_int] * 4 + [ctypes.c_float] * 3
        L.cpuPrj.restype = ctypes.c_int
        L.cpuPrj.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char]
        _ref = L
    return _ref


def ref_run(n, prjWidth, np_, prjFull, dSize, effectiveRate, d, data, cmd):
    """Run the reference cpuPrj on `data` (image for FWD, sinogram for BWD/FBP).

    Returns the float32 output in the reference's row-major layout."""
    L = ref_lib()
    L.setup(int(n), int(prjWidth), int(np_), int(prjFull), float(dSize),
            float(effectiveRate), float(d))
    if cmd == FWD_BIT:
        img = np.ascontiguousarray(data, dtype=np.float32).reshape(n, n).copy()
        sino = np.zeros((np_, prjWidth), np.float32)
        L.cpuPrj(img.ctypes.data, sino.ctypes.data, bytes([cmd]))
        return sino
    sino = np.ascontiguousarray(data, dtype=np.float32).reshape(np_, prjWidth).copy()
    img = np.zeros((n, n), np.float32)
    L.cpuPrj(img.ctypes.data, sino.ctypes.data, bytes([cmd]))
    return img
