"""FP64 NumPy restatement of the wavelet proximal step -- TEST INFRASTRUCTURE ONLY.

Follows, line by line:
  * ``mdwt``  -- rwt/mdwt_r.c:75-169 (MDWT: L levels of periodised two-channel analysis,
                 rows then columns of the top-left block, lowpass first)
  * ``midwt`` -- rwt/midwt_r.c:68-182 (MIDWT: the synthesis, columns then rows, coarsest
                 level first)
  * ``admm``  -- npg/admm.m:1-57 (ADMM for 0.5||x - a||^2 + I(x >= 0) + u||W x||_1 with
                 the adaptive penalty rho), with Utils.softThresh (utils/Utils.m:13-20)
  * ``l1_penalty`` -- pNorm(Psit(x), 1) (bhc/beamhardenSpline.m:208, utils/pNorm.m)
  * ``masked_ops`` -- Psi / Psit with supports (tomography/ConfigCT.m:85-86,
                      utils/maskFunc.m); ``wvlt_mask`` -- utils/wvltMask.m

Images are 2-D arrays whose row index is the MATLAB row index (the C row-major layout of
the same matrix), so ``mdwt(x)`` here equals ``mdwt(x)`` of the MATLAB matrix.

Pinned by: the reference's own MDWT / MIDWT compiled from rwt/*.c into _ref/ (bit for
bit), perfect reconstruction and energy preservation (the periodised transform is
orthonormal for the daubcqf filters), and for ``admm`` the closed form
W^T soft(W a, u) wherever that is non-negative, plus the prox optimality conditions.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np


def _filters(h):
    h = np.asarray(h, np.float64).ravel()
    lh = h.size
    if lh < 2 or lh % 2:
        raise ValueError("filter length must be even")
    lo = h.copy()                                            # h0[lhm1-j] = h[j]
    hi = np.array([(-1.0) ** j * h[lh - 1 - j] for j in range(lh)])  # h1[lhm1-j]
    g0 = h.copy()
    g1 = np.array([(-1.0) ** k * h[lh - 1 - k] for k in range(lh)])
    return lo, hi, g0, g1


def _analysis(x, lo, hi):
    """One level along axis 1 of x: lowpass half then highpass half (fpsconv)."""
    n = x.shape[1]
    lh = lo.size
    p2 = np.arange(0, n, 2)
    al = np.zeros((x.shape[0], n // 2))
    ah = np.zeros((x.shape[0], n // 2))
    for j in range(lh):
        col = x[:, (p2 + j) % n]
        al = al + col * lo[j]
        ah = ah + col * hi[j]
    return np.concatenate([al, ah], axis=1)


def _synthesis(x, g0, g1):
    """Inverse of one level along axis 1 (bpsconv): x = [lowpass | highpass]."""
    n = x.shape[1]
    M = n // 2
    lh = g0.size
    lhm1, lhhm1 = lh - 1, lh // 2 - 1
    lo, hi = x[:, :M], x[:, M:]
    i = np.arange(M)
    e = np.zeros((x.shape[0], M))
    o = np.zeros((x.shape[0], M))
    for j in range(lhhm1 + 1):
        k = (i + j - lhhm1) % M
        e = e + lo[:, k] * g0[lhm1 - 1 - 2 * j] + hi[:, k] * g1[lhm1 - 1 - 2 * j]
        o = o + lo[:, k] * g0[lhm1 - 2 * j] + hi[:, k] * g1[lhm1 - 2 * j]
    out = np.empty_like(x)
    out[:, 0::2] = e
    out[:, 1::2] = o
    return out


def mdwt(x, h, L):
    """rwt/mdwt_r.c MDWT of a 2-D array (rows and columns divisible by 2^L)."""
    y = np.array(x, np.float64, copy=True)
    R, C = y.shape
    lo, hi, _, _ = _filters(h)
    for lev in range(L):
        bm, bn = R >> lev, C >> lev
        y[:bm, :bn] = _analysis(y[:bm, :bn], lo, hi)          # rows
        y[:bm, :bn] = _analysis(y[:bm, :bn].T, lo, hi).T      # columns
    return y


def midwt(y, h, L):
    """rwt/midwt_r.c MIDWT of a 2-D array."""
    x = np.array(y, np.float64, copy=True)
    R, C = x.shape
    _, _, g0, g1 = _filters(h)
    for lev in range(L - 1, -1, -1):
        bm, bn = R >> lev, C >> lev
        x[:bm, :bn] = _synthesis(x[:bm, :bn].T, g0, g1).T     # columns
        x[:bm, :bn] = _synthesis(x[:bm, :bn], g0, g1)         # rows
    return x


def soft_thresh(x, thresh):
    """utils/Utils.m:13-20."""
    t = abs(thresh)
    y = np.zeros_like(x)
    y[x > t] = x[x > t] - t
    y[x < -t] = x[x < -t] + t
    return y


def l1_penalty(x, h, L):
    return float(np.abs(mdwt(x, h, L)).sum())


def masked_ops(h, L, mask, maskk):
    """Psi / Psit of tomography/ConfigCT.m:85-86 on the vectors of the supports mask
    (pixels) and maskk (coefficients), with utils/maskFunc.m's embedding."""
    mask = np.asarray(mask, bool)
    maskk = np.asarray(maskk, bool)

    def embed(v, m):
        x = np.zeros(m.shape)
        x[m] = v
        return x

    Psi = lambda s: midwt(embed(s, maskk), h, L)[mask]   # noqa: E731
    Psit = lambda x: mdwt(embed(x, mask), h, L)[maskk]   # noqa: E731
    return Psi, Psit


def wvlt_mask(mask, h, L):
    """utils/wvltMask.m by its definition: the union over the mask's pixels i of the
    nonzero coefficients of mdwt(e_i)."""
    mask = np.asarray(mask) != 0
    out = np.zeros(mask.shape, bool)
    e = np.zeros(mask.shape)
    for i in zip(*np.nonzero(mask)):
        e[i] = 1.0
        out |= mdwt(e, h, L) != 0
        e[i] = 0.0
    return out


def admm(a, u, h, L, relative_tol=1e-6, max_itr=1000, init=None, ops=None):
    """npg/admm.m: argmin 0.5||x-a||^2 + I(x>=0) + u||Psit(x)||_1; returns (x, iterations).
    ops = (Psi, Psit) (e.g. masked_ops), default the full transforms."""
    a = np.asarray(a, np.float64)
    init = a if init is None else np.asarray(init, np.float64)
    relative_tol = min(1e-3, relative_tol)
    scale = float(np.sqrt((a * a).sum()))
    if scale == 0:
        return np.zeros_like(a), 0
    init = init / scale
    a = a / scale
    u = u / scale
    if ops is None:
        Psi = lambda s: midwt(s, h, L)   # noqa: E731
        Psit = lambda x: mdwt(x, h, L)   # noqa: E731
    else:
        Psi, Psit = ops
    nu = 0.0
    rho = 1.0
    cnt = 0
    preS = Psit(init)
    s = preS
    itr = 0
    while True:
        itr += 1
        cnt += 1
        alpha = np.maximum((a + rho * Psi(s + nu)) / (1 + rho), 0)
        psit_alpha = Psit(alpha)
        s = soft_thresh(psit_alpha - nu, u / rho)
        nu = nu + s - psit_alpha
        difS = float(np.sqrt(((s - preS) ** 2).sum()))
        preS = s
        residual = float(np.sqrt(((s - psit_alpha) ** 2).sum()))
        if itr > max_itr:
            break
        if difS <= relative_tol and residual <= relative_tol:
            break
        if cnt > 20:
            if difS > 10 * residual:
                rho = rho / 2
                nu = nu * 2
                cnt = 0
            elif difS < residual / 10:
                rho = rho * 2
                nu = nu / 2
                cnt = 0
    alpha = np.maximum((a + rho * Psi(s + nu)) / (1 + rho), 0) * scale
    return alpha, itr


# ---------------------------------------------------- the reference's own transforms ---

_REF_SO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "librwt_ref.so")
_rwt = None


def ref_available() -> bool:
    return os.path.exists(_REF_SO)


def _ref():
    global _rwt
    if _rwt is None:
        L = ctypes.CDLL(_REF_SO)
        dp, i = ctypes.POINTER(ctypes.c_double), ctypes.c_int
        for name in ("MDWT", "MIDWT"):
            f = getattr(L, name)
            f.restype = None
            f.argtypes = [dp, i, i, dp, i, i, dp]
        _rwt = L
    return _rwt


def ref_mdwt(x, h, L):
    """y = MDWT(x) of rwt/mdwt_r.c (MDWT(x_in, m, n, h, lh, L, y_out))."""
    # MATLAB matrices are column-major with m = rows: the Fortran-ordered copy
    xf = np.asfortranarray(np.array(x, np.float64))
    hf = np.ascontiguousarray(np.asarray(h, np.float64).ravel())
    yf = np.zeros_like(xf, order="F")
    dp = ctypes.POINTER(ctypes.c_double)
    m, n = xf.shape
    _ref().MDWT(xf.ctypes.data_as(dp), m, n, hf.ctypes.data_as(dp), hf.size, L,
                yf.ctypes.data_as(dp))
    return np.array(yf)


def ref_midwt(y, h, L):
    """x = MIDWT(y) of rwt/midwt_r.c (its signature is MIDWT(x_out, m, n, h, lh, L, y_in))."""
    yf = np.asfortranarray(np.array(y, np.float64))
    hf = np.ascontiguousarray(np.asarray(h, np.float64).ravel())
    xf = np.zeros_like(yf, order="F")
    dp = ctypes.POINTER(ctypes.c_double)
    m, n = yf.shape
    _ref().MIDWT(xf.ctypes.data_as(dp), m, n, hf.ctypes.data_as(dp), hf.size, L,
                 yf.ctypes.data_as(dp))
    return np.array(xf)
