"""The reference's MEX operator API, ``gpuPrj(arg0, conf, cmd)`` (prj/mPrj.c:37-173).

Same three-argument call, same command strings and the same marshalling as the MEX
gateway, on top of the C ABI's prj_setup / prj_run (host arrays, synchronous):

=====================================  =====================================================
``gpuPrj(0, conf, 'config')``          conf fields n, prjWidth, np, prjFull, dSize,
                                       effectiveRate, d (prj/mPrj.c:49-73); returns None
``s = gpuPrj(img, 0, 'forward')``      img n x n (MATLAB matrix semantics: element (i, j)
                                       is image row i, column j; a flat vector is taken in
                                       column-major order); returns a float64 column
                                       vector of prjWidth*np, view-major (prj/mPrj.c:75-111)
``x = gpuPrj(sino, 0, 'backward')``    sino with prjWidth*np elements, MATLAB column-major
                                       order (a prjWidth x np matrix or its flat vector);
                                       returns float64 n*n x 1, column-major image
                                       (prj/mPrj.c:112-136)
``x = gpuPrj(sino, 0, 'FBP')``         as 'backward' (prj/mPrj.c:137-159)
any other string                       prints the configuration (prj/mPrj.c:160-170)
=====================================  =====================================================

Differences from the MEX, by design: a wrong argument count or dtype raises
``ValueError``/``TypeError`` instead of printing and returning nothing
(prj/mPrj.c:43-46,101-104), and device errors raise ``CtprjError`` instead of calling
exit() (prj/gpuPrj.cu:60-68).  ``cpuPrj`` is the same entry (ConfigCT falls back from
gpuPrj to cpuPrj by name only, tomography/ConfigCT.m:29-40).
"""
from __future__ import annotations

import numpy as np

from ._lib import BWD_BIT, FBP_BIT, FWD_BIT, Geom, check, lib

_FIELDS = ("n", "prjWidth", "np", "prjFull", "dSize", "effectiveRate", "d")


def _field(conf, name):
    if isinstance(conf, dict):
        if name not in conf:
            raise ValueError(f"config struct has no field '{name}'")
        return conf[name]
    if not hasattr(conf, name):
        raise ValueError(f"config struct has no field '{name}'")
    return getattr(conf, name)


def _current():
    g = Geom()
    check(lib().prj_get_geom(g))
    return g


def _numeric(a, what):
    a = np.asarray(a)
    if a.dtype == np.float64 or a.dtype == np.float32:
        return a
    raise TypeError(f"unknown input parameter type in mPrj {what}: {a.dtype}")


def gpuPrj(*args):
    if len(args) != 3:
        raise ValueError(f"number of input arguments are: {len(args)} (need 3)")
    arg0, conf, cmd = args
    L = lib()
    if cmd == "config":
        n = int(_field(conf, "n"))
        W = int(_field(conf, "prjWidth"))
        npj = int(_field(conf, "np"))
        pf = int(_field(conf, "prjFull"))
        ds = float(_field(conf, "dSize"))
        er = float(_field(conf, "effectiveRate"))
        d = float(_field(conf, "d"))
        check(L.prj_setup(n, W, npj, pf, ds, er, d))
        return None
    if cmd == "forward":
        g = _current()
        a = _numeric(arg0, "forward")
        if a.size != g.n * g.n:
            raise ValueError(f"forward: image has {a.size} elements, config needs {g.n * g.n}")
        img = a.reshape(g.n, g.n) if a.ndim == 2 else a.reshape(g.n, g.n, order="F")
        img = np.ascontiguousarray(img, dtype=np.float32)  # row-major, same orientation
        sino = np.empty(g.np * g.prjWidth, np.float32)
        check(L.prj_run(img.ctypes.data, sino.ctypes.data, FWD_BIT))
        return sino.astype(np.float64).reshape(-1, 1)
    if cmd in ("backward", "FBP"):
        g = _current()
        a = _numeric(arg0, cmd)
        if a.size != g.np * g.prjWidth:
            raise ValueError(f"{cmd}: sinogram has {a.size} elements, config needs "
                             f"{g.np * g.prjWidth}")
        sino = np.ascontiguousarray(a.reshape(-1, order="F"), dtype=np.float32)
        img = np.empty((g.n, g.n), np.float32)
        check(L.prj_run(img.ctypes.data, sino.ctypes.data,
                        BWD_BIT if cmd == "backward" else FBP_BIT))
        return img.astype(np.float64).reshape(-1, order="F").reshape(-1, 1)
    check(L.prj_show())
    return None


cpuPrj = gpuPrj


def fan_par_ops(conf, Ts: float = 1.0, mask=None):
    """(Phi, Phit, FBP) closures of tomography/ConfigCT.m:109-135 over gpuPrj.

    ``conf`` carries ConfigCT's fields (imgSize, prjWidth, prjNum, prjFull, dSize,
    effectiveRate, dist) or the prjConf names.  ``mask`` is an optional boolean n x n
    image: Phi takes the masked vector (MATLAB find(mask) order, utils/maskFunc.m) and
    Phit returns it."""
    def get(*names):
        for nm in names:
            if isinstance(conf, dict) and nm in conf:
                return conf[nm]
            if not isinstance(conf, dict) and hasattr(conf, nm):
                return getattr(conf, nm)
        raise ValueError(f"conf lacks {names}")

    c = {"n": get("imgSize", "n"), "prjWidth": get("prjWidth"), "np": get("prjNum", "np"),
         "prjFull": get("prjFull"), "dSize": get("dSize"),
         "effectiveRate": get("effectiveRate"), "d": get("dist", "d")}
    gpuPrj(0, c, "config")
    n = int(c["n"])
    if mask is not None:
        m = np.asarray(mask).reshape(-1, order="F") != 0
        idx = np.flatnonzero(m)  # column-major linear indices, as find(mask)

        def mb(s):  # maskFunc(s, maskIdx, n): scatter the masked vector into n x n
            x = np.zeros(n * n)
            x[idx] = np.asarray(s).reshape(-1)
            return x.reshape(n, n, order="F")

        def ma(x):  # gather
            return np.asarray(x).reshape(-1, order="F")[idx].reshape(-1, 1)
    else:
        def mb(s):
            s = np.asarray(s)
            return s if s.ndim == 2 else s.reshape(n, n, order="F")

        def ma(x):
            return x

    Phi = lambda s: gpuPrj(mb(s), 0, "forward") * Ts  # noqa: E731
    Phit = lambda s: ma(gpuPrj(s, 0, "backward")) * Ts  # noqa: E731
    FBP = lambda s: gpuPrj(s, 0, "FBP") / Ts  # noqa: E731
    return Phi, Phit, FBP
