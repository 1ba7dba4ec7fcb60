"""Total-variation proximal step on the device (utils/TV.m denoise ->
utils/denoise_bound_mod.m, utils/tlv.m): ``TV(rows, cols, tv='iso'|'l1').denoise(x, u,
innerThresh, maxInnerItr)`` -> (x, iterations) solves argmin 0.5||y - x||^2 + u TV(y)
over y >= 0; ``.tlv(x)`` is the TV value.  Float32 CUDA images (a stack of ``batch``),
FP64 state inside."""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import check, lib
from .projector import _ptr, _stream_ptr


class TV:
    def __init__(self, rows: int, cols: int, tv: str = "iso", batch: int = 1):
        if tv not in ("iso", "l1"):
            raise ValueError("tv must be 'iso' or 'l1'")
        self.rows, self.cols, self.batch, self.tv = int(rows), int(cols), int(batch), tv
        out = ctypes.c_void_p()
        check(lib().ctprj_tv_create(self.rows, self.cols, self.batch, 1 if tv == "iso" else 0,
                                    ctypes.byref(out)))
        self._h = out

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("TV context is closed")
        return self._h

    @property
    def elems(self) -> int:
        return self.rows * self.cols * self.batch

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().ctprj_tv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, t: torch.Tensor, what: str) -> torch.Tensor:
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda" or t.dtype != torch.float32:
            raise TypeError(f"{what} must be a float32 CUDA tensor")
        if t.numel() != self.elems:
            raise ValueError(f"{what} has {t.numel()} elements, expected {self.elems}")
        return t.contiguous()

    def set_mask(self, mask=None):
        """Support (opt.mask of TV.denoise); nonzero = inside; None = no mask."""
        if mask is None:
            check(lib().ctprj_tv_set_mask(self.handle, None))
            return
        m = mask if isinstance(mask, torch.Tensor) else torch.from_numpy(np.asarray(mask))
        m = self._check((m != 0).to(device="cuda", dtype=torch.float32).reshape(-1), "mask")
        check(lib().ctprj_tv_set_mask(self.handle, _ptr(m)))

    def tlv(self, x: torch.Tensor) -> float:
        v = ctypes.c_double()
        check(lib().ctprj_tv_value(self.handle, _ptr(self._check(x, "x")), ctypes.byref(v),
                                   _stream_ptr()))
        return v.value

    def denoise(self, x: torch.Tensor, u: float, innerThresh: float = 1e-4,
                maxInnerItr: int = 100):
        x = self._check(x, "x")
        out = torch.empty_like(x)
        itr = ctypes.c_int()
        check(lib().ctprj_tv_denoise(self.handle, _ptr(x), float(u), float(innerThresh),
                                     int(maxInnerItr), _ptr(out), ctypes.byref(itr),
                                     _stream_ptr()))
        return out, itr.value


__all__ = ["TV"]
