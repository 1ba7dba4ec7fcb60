"""Orthogonal wavelet transforms and the wavelet-l1 proximal step on the device.

Mirrors of the reference's MATLAB contract, computed by libctprj:

* ``daubcqf(N, TYPE)`` -> (h_0, h_1): Daubechies scaling / wavelet filters
  (rwt/daubcqf.m), computed on the host
* ``Wavelet(rows, cols, h, L).mdwt(x)`` / ``.midwt(y)``: the periodised 2-D transforms of
  rwt/mdwt_r.c and rwt/midwt_r.c (Psit and Psi of tomography/ConfigCT.m:82-86)
* ``Wavelet.admm(a, u, relativeTol, maxItr, init)`` -> (x, iterations): npg/admm.m,
  argmin 0.5||x - a||^2 + I(x >= 0) + u ||mdwt(x)||_1
* ``Wavelet.penalty(x)``: ||mdwt(x)||_1 (the wvltADMM penalty of
  bhc/beamhardenSpline.m:208)
* ``Wavelet.set_masks(mask, coef_mask)`` and ``wvlt_mask`` (utils/wvltMask.m): the
  supports of ConfigCT's Psi / Psit

Images are float32 CUDA tensors (a stack of ``batch`` images of rows x cols, row-major
with the MATLAB row index as row); the transforms and the ADMM state are FP64 inside.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import check, dbl_array, lib
from .projector import _ptr, _stream_ptr


def daubcqf(N: int, TYPE: str = "min"):
    """rwt/daubcqf.m: Daubechies' filters of length N (even), normalised to sqrt(2):
    the product of (1 + z^-1)^(N/2) with the spectral factor whose zeros are the K-1
    roots inside the unit circle ('min'), reversed for 'max'.  Returns (h_0, h_1)."""
    if N % 2:
        raise ValueError("No Daubechies filter exists for ODD length")
    if TYPE not in ("min", "max"):
        raise ValueError("TYPE must be 'min' or 'max'")
    K = N // 2
    a = 1.0
    p = np.array([1.0])
    q = np.array([1.0])
    h0 = np.array([1.0, 1.0])
    for j in range(1, K):
        a = -a * 0.25 * (j + K - 1) / j
        h0 = np.concatenate([[0.0], h0]) + np.concatenate([h0, [0.0]])
        p = np.concatenate([[0.0], -p]) + np.concatenate([p, [0.0]])
        p = np.concatenate([[0.0], -p]) + np.concatenate([p, [0.0]])
        q = np.concatenate([[0.0], q, [0.0]]) + a * p
    if K > 1:
        r = np.roots(q)
        # MATLAB sorts complex values by magnitude, then phase angle
        r = r[np.lexsort((np.angle(r), np.abs(r)))]
        h0 = np.convolve(h0, np.real(np.poly(r[:K - 1])))
    h0 = np.sqrt(2.0) * h0 / h0.sum()
    if TYPE == "max":
        h0 = h0[::-1].copy()
    if abs((h0 ** 2).sum()) - 1 > 1e-4:
        raise ValueError('Numerically unstable for this value of "N".')
    h1 = h0[::-1].copy()
    h1[0::2] = -h1[0::2]
    return h0, h1


def _as_f32(x, device) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return (x != 0).to(device=device, dtype=torch.float32).reshape(-1)
    return torch.from_numpy((np.asarray(x) != 0).astype(np.float32).ravel()).to(device)


def wvlt_mask(mask, h, L) -> np.ndarray:
    """The coefficients that depend on a pixel of ``mask`` (utils/wvltMask.m: the union
    of the supports of mdwt(e_i) over the mask's pixels), from one transform of the
    mask with every tap replaced by its magnitude, so that no contribution cancels (the
    reference's test `Wt(e_i) ~= 0` differs only where contributions cancel exactly)."""
    m = (np.asarray(mask) != 0).astype(np.float64)
    R, C = m.shape
    a = np.abs(np.asarray(h, np.float64).ravel())
    lh = a.size
    lo, hi = a, a[::-1]  # tap magnitudes of the lowpass and the highpass analysis filters

    def analysis(x):
        n = x.shape[1]
        p2 = np.arange(0, n, 2)
        al = np.zeros((x.shape[0], n // 2))
        ah = np.zeros((x.shape[0], n // 2))
        for j in range(lh):
            col = x[:, (p2 + j) % n]
            al = al + col * lo[j]
            ah = ah + col * hi[j]
        return np.concatenate([al, ah], axis=1)

    y = m.copy()
    for lev in range(L):
        bm, bn = R >> lev, C >> lev
        y[:bm, :bn] = analysis(y[:bm, :bn])
        y[:bm, :bn] = analysis(y[:bm, :bn].T).T
    return y > 0


class Wavelet:
    """Device wavelet context: rows x cols images (both divisible by 2^L), ``batch``
    images per call, scaling filter h (even length <= 32), L levels."""

    def __init__(self, rows: int, cols: int, h, L: int, batch: int = 1, device=None):
        self.rows, self.cols, self.L, self.batch = int(rows), int(cols), int(L), int(batch)
        self.h = np.asarray(h, np.float64).ravel()
        self.device = torch.device(device if device is not None else "cuda")
        hp, _hh = dbl_array(self.h)
        out = ctypes.c_void_p()
        check(lib().ctprj_wavelet_create(self.rows, self.cols, self.batch, hp, self.h.size,
                                         self.L, ctypes.byref(out)))
        self._h = out

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("wavelet context is closed")
        return self._h

    @property
    def elems(self) -> int:
        return self.rows * self.cols * self.batch

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().ctprj_wavelet_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_masks(self, mask=None, coef_mask=None):
        """Supports of Psi / Psit (tomography/ConfigCT.m:85-86): ``mask`` the image pixels
        (opt.mask), ``coef_mask`` the coefficients (maskk, see ``wvlt_mask``); arrays of
        the context's size, nonzero = inside; None = no mask."""
        m = None if mask is None else self._check(_as_f32(mask, self.device), "mask")
        c = None if coef_mask is None else self._check(_as_f32(coef_mask, self.device),
                                                       "coef_mask")
        check(lib().ctprj_wavelet_set_masks(self.handle, _ptr(m) if m is not None else None,
                                            _ptr(c) if c is not None else None))

    def _check(self, t: torch.Tensor, what: str) -> torch.Tensor:
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda" or t.dtype != torch.float32:
            raise TypeError(f"{what} must be a float32 CUDA tensor")
        if t.numel() != self.elems:
            raise ValueError(f"{what} has {t.numel()} elements, expected {self.elems}")
        return t.contiguous()

    def mdwt(self, x: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        x = self._check(x, "x")
        out = torch.empty_like(x) if out is None else self._check(out, "out")
        check(lib().ctprj_dwt(self.handle, _ptr(x), _ptr(out), _stream_ptr()))
        return out

    def midwt(self, y: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        y = self._check(y, "y")
        out = torch.empty_like(y) if out is None else self._check(out, "out")
        check(lib().ctprj_idwt(self.handle, _ptr(y), _ptr(out), _stream_ptr()))
        return out

    def penalty(self, x: torch.Tensor) -> float:
        x = self._check(x, "x")
        v = ctypes.c_double()
        check(lib().ctprj_wavelet_l1(self.handle, _ptr(x), ctypes.byref(v), _stream_ptr()))
        return v.value

    def admm(self, a: torch.Tensor, u: float, relativeTol: float = 1e-6, maxItr: int = 1000,
             init: torch.Tensor | None = None):
        a = self._check(a, "a")
        ip = _ptr(self._check(init, "init")) if init is not None else None
        out = torch.empty_like(a)
        itr = ctypes.c_int()
        check(lib().ctprj_prox_wavelet_admm(self.handle, _ptr(a), float(u), float(relativeTol),
                                            int(maxItr), ip, _ptr(out), ctypes.byref(itr),
                                            _stream_ptr()))
        return out, itr.value


__all__ = ["daubcqf", "wvlt_mask", "Wavelet"]
