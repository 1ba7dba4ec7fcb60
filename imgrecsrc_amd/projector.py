"""Device-resident projector plans over one or many slices (libctprj plan API).

``Projector`` owns a ``ctprj_plan``: slices of ragged size and mixed parallel/fan
geometry are packed back to back (images n_s*n_s each, sinograms np_s*prjWidth_s each)
and every call runs all slices in one launch on torch's current HIP stream.

Geometry fields are those of the reference's struct prjConf (prj/prj.h:42-57) and the
``conf`` struct that tomography/ConfigCT.m:110-116 passes to the MEX 'config' command:
n, prjWidth, np, prjFull, dSize, effectiveRate, d (d <= 0: parallel beam).
"""
from __future__ import annotations

import ctypes
from typing import Iterable, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import Geom, Tune, check, lib

GEOM_FIELDS = ("n", "prjWidth", "np", "prjFull", "dSize", "effectiveRate", "d")


def make_geom(g) -> Geom:
    """Geom from a dict with the prjConf field names, a Geom, or a 7-tuple."""
    if isinstance(g, Geom):
        return g
    if isinstance(g, dict):
        vals = [g[k] for k in GEOM_FIELDS]
    else:
        vals = list(g)
    n, W, npj, pf, ds, er, d = vals
    return Geom(int(n), int(W), int(npj), int(pf), float(ds), float(er), float(d))


def _stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t: torch.Tensor):
    return ctypes.c_void_p(t.data_ptr())


class Projector:
    """sino = A img, img = A^T sino, FBP -- for every slice of the plan at once."""

    def __init__(self, geoms: Iterable, device=None, symmetric: bool = True):
        geoms = [make_geom(g) for g in (geoms if not isinstance(geoms, (dict, Geom)) else [geoms])]
        if not geoms:
            raise ValueError("need at least one slice geometry")
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device(self.device.type, torch.cuda.current_device())
        self.geoms = geoms
        arr = (Geom * len(geoms))(*geoms)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(lib().ctprj_plan_create(arr, len(geoms), ctypes.byref(h)))
        self._h = h
        ie, se = ctypes.c_longlong(), ctypes.c_longlong()
        io = (ctypes.c_longlong * len(geoms))()
        so = (ctypes.c_longlong * len(geoms))()
        check(lib().ctprj_plan_layout(h, ctypes.byref(ie), ctypes.byref(se), io, so))
        self.img_elems, self.sino_elems = ie.value, se.value
        self.img_offsets = list(io)
        self.sino_offsets = list(so)
        if not symmetric:
            self.set_tune(flags=0)

    # ------------------------------------------------------------------ lifetime
    def close(self):
        if getattr(self, "_h", None):
            lib().ctprj_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        if not self._h:
            raise RuntimeError("projector plan has been closed")
        return self._h

    # ------------------------------------------------------------------ layout
    @property
    def nslices(self) -> int:
        return len(self.geoms)

    def img_shape(self, s: int = 0):
        return (self.geoms[s].n, self.geoms[s].n)

    def sino_shape(self, s: int = 0):
        return (self.geoms[s].np, self.geoms[s].prjWidth)

    def pack_images(self, imgs: Sequence[torch.Tensor]) -> torch.Tensor:
        return torch.cat([i.reshape(-1).to(torch.float32) for i in imgs]).to(self.device)

    def unpack_images(self, flat: torch.Tensor):
        return [flat[o:o + g.n * g.n].view(g.n, g.n) for o, g in zip(self.img_offsets, self.geoms)]

    def pack_sinos(self, sinos: Sequence[torch.Tensor]) -> torch.Tensor:
        return torch.cat([s.reshape(-1).to(torch.float32) for s in sinos]).to(self.device)

    def unpack_sinos(self, flat: torch.Tensor):
        return [flat[o:o + g.np * g.prjWidth].view(g.np, g.prjWidth)
                for o, g in zip(self.sino_offsets, self.geoms)]

    def _check(self, t: torch.Tensor, elems: int, what: str) -> torch.Tensor:
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
            raise TypeError(f"{what} must be a CUDA (HIP) tensor")
        if t.device != self.device:
            raise ValueError(f"{what} is on {t.device}, the plan on {self.device}")
        if t.dtype != torch.float32:
            raise TypeError(f"{what} must be float32, got {t.dtype}")
        if t.numel() != elems:
            raise ValueError(f"{what} has {t.numel()} elements, plan needs {elems}")
        return t.contiguous()

    def _out(self, out: torch.Tensor | None, elems: int, like: torch.Tensor) -> torch.Tensor:
        """The output buffer: a new one, or the caller's after the same checks as an input
        plus contiguity (the kernels write `elems` floats from its first element; a
        contiguous copy would silently drop the result)."""
        if out is None:
            return torch.empty(elems, dtype=torch.float32, device=like.device)
        self._check(out, elems, "out")
        if not out.is_contiguous():
            raise ValueError("out must be contiguous")
        return out

    # ------------------------------------------------------------------ operators
    def forward(self, img: torch.Tensor, out: torch.Tensor | None = None, stream=None):
        img = self._check(img, self.img_elems, "image")
        out = self._out(out, self.sino_elems, img)
        check(lib().ctprj_forward(self.handle, _ptr(img), _ptr(out), _stream_ptr(stream)))
        return out

    def adjoint(self, sino: torch.Tensor, out: torch.Tensor | None = None, stream=None):
        sino = self._check(sino, self.sino_elems, "sinogram")
        out = self._out(out, self.img_elems, sino)
        check(lib().ctprj_adjoint(self.handle, _ptr(sino), _ptr(out), _stream_ptr(stream)))
        return out

    def fbp(self, sino: torch.Tensor, out: torch.Tensor | None = None, stream=None):
        sino = self._check(sino, self.sino_elems, "sinogram")
        out = self._out(out, self.img_elems, sino)
        check(lib().ctprj_fbp(self.handle, _ptr(sino), _ptr(out), _stream_ptr(stream)))
        return out

    def nnz(self, stream=None) -> int:
        v = ctypes.c_longlong()
        check(lib().ctprj_plan_nnz(self.handle, ctypes.byref(v), _stream_ptr(stream)))
        return v.value

    def modes(self):
        """Per slice: 0 general kernels; symmetric kernels under the 8-fold (1), 4-fold (2)
        or 2-fold (3) symmetry group of its view set."""
        m = (ctypes.c_int * self.nslices)()
        check(lib().ctprj_plan_modes(self.handle, m))
        return list(m)

    # ------------------------------------------------------------------ tuning
    def get_tune(self) -> dict:
        t = Tune()
        check(lib().ctprj_plan_get_tune(self.handle, ctypes.byref(t)))
        return {"fwd_block": t.fwd_block, "adj_tile": t.adj_tile, "flags": t.flags,
                "adj_seg": t.adj_seg, "fwd_seg": t.fwd_seg, "adj_split": t.adj_split,
                "fwd_split": t.fwd_split}

    def set_tune(self, fwd_block=None, adj_tile=None, flags=None, adj_seg=None, fwd_seg=None,
                 adj_split=None, fwd_split=None):
        cur = self.get_tune()
        given = {"fwd_block": fwd_block, "adj_tile": adj_tile, "flags": flags,
                 "adj_seg": adj_seg, "fwd_seg": fwd_seg, "adj_split": adj_split,
                 "fwd_split": fwd_split}
        t = Tune(*[int(cur[k] if v is None else v) for k, v in given.items()])
        check(lib().ctprj_plan_set_tune(self.handle, ctypes.byref(t)))

    def autotune(self, reps: int = 3, cache_path: str | None = None, stream=None) -> dict:
        cp = cache_path.encode() if cache_path else None
        check(lib().ctprj_plan_autotune(self.handle, int(reps), cp, _stream_ptr(stream)))
        return self.get_tune()


def device_arch() -> str:
    buf = ctypes.create_string_buffer(64)
    check(lib().ctprj_device_arch(buf, 64))
    return buf.value.decode()


def version() -> str:
    return lib().ctprj_version().decode()


__all__ = ["Projector", "make_geom", "device_arch", "version", "GEOM_FIELDS", "_lib"]
