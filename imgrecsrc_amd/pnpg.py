"""On-device PNPG alpha step (npg/PNPG.m:50-167, non-negativity prox) over a projector
plan and the polychromatic objective of bhc/poissLAlpha.m (or bhc/gaussLAlpha.m), in the
b1 / b0 / dis spectrum basis.

``PNPG.step(alpha)`` enqueues one line-search trial with no host synchronisation, so a
run of trials can be captured once in a HIP graph (``torch.cuda.CUDAGraph``) and
replayed; ``stats()`` reads the device scalars.  ``PNPG.set_wavelet(w, u)`` switches to
the reference's default prox ('wvltADMM': non-negativity plus u ||W alpha||_1 solved by
npg/admm.m), ``PNPG.set_tv(tv, u)`` to 'tvl1' / 'tviso' (utils/TV.m); their trials
synchronise.  ``PNPG.save(path, alpha)`` / ``PNPG.load(path, alpha)`` snapshot a run and
resume it in a new solver bit for bit (bhc/doBHC.m:85-91 saves and warm-starts between
penalty levels).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import PnpgInfo, check, dbl_array, lib
from .bhc import _knots_ie, gaussLAlpha, poissLAlpha, set_model
from .projector import Projector, _ptr, _stream_ptr


def bb_step(proj: Projector, Imea, Ie, alpha, Ts, knots, rel=1e-5, basis: str = "b1",
            noise: str = "poisson") -> float:
    """Barzilai-Borwein estimate of the inverse step (npg/Methods.m:100-107):
    t = |<g1 - g2, d>| / ||d||^2 with d = rel * g1/||g1||, g2 = grad(alpha - d)."""
    obj = gaussLAlpha if noise.lower() == "gaussian" else poissLAlpha
    _, g1 = obj(Imea, Ie, alpha, proj, Ts, knots, basis=basis)
    d = rel * g1 / torch.linalg.vector_norm(g1.double()).float()
    _, g2 = obj(Imea, Ie, alpha - d, proj, Ts, knots, basis=basis)
    num = torch.dot((g1 - g2).double(), d.double()).abs().item()
    den = torch.dot(d.double(), d.double()).item()
    t = num / den if den > 0 else 1.0
    return t if np.isfinite(t) and t > 0 else 1.0


class PNPG:
    def __init__(self, proj: Projector, Imea: torch.Tensor, knots, Ie, Ts: float,
                 alpha0: torch.Tensor, t0: float | None = None, basis: str = "b1",
                 noise: str = "poisson"):
        self.proj = proj
        self.Imea = Imea.reshape(-1).contiguous()
        self.knots = np.asarray(knots, np.float64)
        self.Ie = np.asarray(Ie, np.float64)
        self.Ts = float(Ts)
        a0 = alpha0.reshape(-1).contiguous()
        if t0 is None:
            t0 = bb_step(proj, self.Imea, self.Ie, a0, self.Ts, self.knots, basis=basis,
                         noise=noise)
        self.t0 = float(t0)
        E, kp, self._kh, ip, self._ih = _knots_ie(self.knots, self.Ie, basis)
        set_model(proj, basis, noise)  # the solver keeps the model it is created with
        h = ctypes.c_void_p()
        check(lib().ctprj_pnpg_create(proj.handle, _ptr(self.Imea), kp, E, ip, self.Ts, _ptr(a0),
                                      ctypes.c_double(self.t0), ctypes.byref(h)))
        self._h = h

    def set_method(self, method: str = "PNPG", preSteps: int = 2):
        """'PNPG' (npg/PNPG.m) or 'NPG' (npg/NPG.m, opt.alphaStep = 'NPG'); before the
        first step."""
        if method.upper() not in ("PNPG", "NPG"):
            raise ValueError("method must be 'PNPG' or 'NPG'")
        check(lib().ctprj_pnpg_set_method(self._h, 1 if method.upper() == "NPG" else 0,
                                          int(preSteps)))

    def set_mask(self, mask=None):
        """Support of alpha (opt.mask): pixels outside stay zero; None = no mask.  Give a
        wavelet or TV prior the same mask."""
        if mask is None:
            check(lib().ctprj_pnpg_set_mask(self._h, None))
            return
        m = mask if isinstance(mask, torch.Tensor) else torch.from_numpy(np.asarray(mask))
        m = (m != 0).to(device="cuda", dtype=torch.float32).reshape(-1).contiguous()
        if m.numel() != self.proj.img_elems:
            raise ValueError("mask must have the plan's image size")
        self._mask = m
        check(lib().ctprj_pnpg_set_mask(self._h, _ptr(m)))

    def set_wavelet(self, wavelet, u: float):
        """Wavelet l1 prox u ||mdwt(alpha)||_1 (opt.proximal = 'wvltADMM'); ``wavelet`` is an
        imgrecsrc_amd.wavelet.Wavelet over the plan's images (None: non-negativity only).
        Steps then synchronise and cannot be graph-captured."""
        self._wavelet = wavelet  # keeps the context alive as long as the solver
        check(lib().ctprj_pnpg_set_wavelet(self._h, wavelet.handle if wavelet is not None
                                           else None, ctypes.c_double(float(u))))

    def set_tv(self, tv, u: float):
        """TV prox u TV(alpha) (opt.proximal = 'tvl1' / 'tviso'); ``tv`` is an
        imgrecsrc_amd.tv.TV over the plan's images (None: non-negativity only)."""
        self._tv = tv
        check(lib().ctprj_pnpg_set_tv(self._h, tv.handle if tv is not None else None,
                                      ctypes.c_double(float(u))))

    def set_linesearch(self, max_shrinks: int = 20):
        """Failed tests per iteration that shrink the step before the line search gives up
        (the reference's 20)."""
        check(lib().ctprj_pnpg_set_linesearch(self._h, int(max_shrinks)))

    def step(self, alpha: torch.Tensor, stream=None):
        check(lib().ctprj_pnpg_step(self._h, _ptr(alpha), _stream_ptr(stream)))

    def stats(self, stream=None) -> dict:
        info = PnpgInfo()
        check(lib().ctprj_pnpg_stats(self._h, ctypes.byref(info), _stream_ptr(stream)))
        return {f: getattr(info, f) for f, _ in PnpgInfo._fields_}

    def save_state(self, stream=None) -> np.ndarray:
        """The solver's iteration state after every queued trial, as an opaque byte array
        (ctprj_pnpg_save); alpha is not in it."""
        n = ctypes.c_longlong()
        check(lib().ctprj_pnpg_snapshot_bytes(self._h, ctypes.byref(n)))
        blob = np.empty(n.value, np.uint8)
        check(lib().ctprj_pnpg_save(self._h, blob.ctypes.data_as(ctypes.c_void_p), n.value,
                                    _stream_ptr(stream)))
        return blob

    def load_state(self, blob, stream=None):
        """Restores a save_state() blob into this solver (created with the same plan,
        measurements, spectrum, Ts and prox): its next step continues the saved run."""
        blob = np.ascontiguousarray(np.asarray(blob, np.uint8).reshape(-1))
        check(lib().ctprj_pnpg_load(self._h, blob.ctypes.data_as(ctypes.c_void_p), blob.size,
                                    _stream_ptr(stream)))

    def save(self, path, alpha: torch.Tensor):
        """Snapshot of a run to an .npz file: the solver state, alpha, Ie, knots and Ts
        (bhc/doBHC.m:85 saves the workspace after each penalty level)."""
        np.savez(path, state=self.save_state(), alpha=alpha.detach().reshape(-1).cpu().numpy(),
                 Ie=self.Ie, knots=self.knots, Ts=np.float64(self.Ts))

    def load(self, path, alpha: torch.Tensor):
        """Resumes from save(): the solver state is restored and alpha (the plan's image
        vector on the device) overwritten with the saved iterate."""
        with np.load(path, allow_pickle=False) as z:
            a = z["alpha"]
            if a.size != alpha.numel():
                raise ValueError(f"snapshot alpha has {a.size} pixels, alpha {alpha.numel()}")
            if not (np.array_equal(z["Ie"], self.Ie) and np.array_equal(z["knots"], self.knots)
                    and float(z["Ts"]) == self.Ts):
                raise ValueError("snapshot of another spectrum (Ie, knots) or Ts")
            self.load_state(z["state"])
            alpha.copy_(torch.from_numpy(a).to(alpha.device).reshape(alpha.shape))

    def close(self):
        if getattr(self, "_h", None):
            lib().ctprj_pnpg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["PNPG", "bb_step"]
