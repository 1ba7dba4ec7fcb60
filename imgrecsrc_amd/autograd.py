"""torch.autograd.Function bindings over the ctypes-level C ABI.

* ``project(img, proj)`` -- sino = A img; backward: A^T (the hand-written adjoint kernel).
* ``backproject(sino, proj)`` -- img = A^T sino; backward: A.
Outputs are flat (packed) arrays; gradients take the shape of their input (an n x n
image, an np x prjWidth sinogram or the packed vector).
* ``poisson_nll(alpha, Ie, proj, Imea, knots, Ts)`` -- the polychromatic Poisson NLL of
  bhc/poissLAlpha.m (FP64 scalar).  Forward keeps only the projection A*alpha; backward
  recomputes the per-ray spectral exponentials from it (no stored exp or ray x basis
  matrix) and returns dNLL/dalpha = Ts*A^T[(1-Imea/BL) sBL] and
  dNLL/dIe = A_s^T (1 - Imea/Ir) (bhc/poissLI.m, FP64).
"""
from __future__ import annotations

import torch

from . import bhc
from .projector import Projector


class _Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, img, proj: Projector):
        ctx.proj, ctx.shape = proj, img.shape
        return proj.forward(img.detach().reshape(-1))

    @staticmethod
    def backward(ctx, g):
        return ctx.proj.adjoint(g.contiguous().reshape(-1).float()).view(ctx.shape), None


class _Backproject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sino, proj: Projector):
        ctx.proj, ctx.shape = proj, sino.shape
        return proj.adjoint(sino.detach().reshape(-1))

    @staticmethod
    def backward(ctx, g):
        return ctx.proj.forward(g.contiguous().reshape(-1).float()).view(ctx.shape), None


class _PoissonNLL(torch.autograd.Function):
    @staticmethod
    def forward(ctx, alpha, Ie, proj: Projector, Imea, knots, Ts: float):
        sraw = proj.forward(alpha.detach().reshape(-1).float())
        ie = Ie.detach().double().cpu().numpy()
        f, _ = bhc.alpha_rays(proj, sraw, Imea, ie, knots, Ts, want_w=False)
        ctx.save_for_backward(sraw, Imea)
        ctx.proj, ctx.knots, ctx.Ts, ctx.ie = proj, knots, Ts, ie
        ctx.shape = alpha.shape
        ctx.ie_dtype, ctx.ie_device = Ie.dtype, Ie.device
        return f

    @staticmethod
    def backward(ctx, gf):
        sraw, Imea = ctx.saved_tensors
        proj = ctx.proj
        ga = gi = None
        if ctx.needs_input_grad[0]:
            _, w = bhc.alpha_rays(proj, sraw, Imea, ctx.ie, ctx.knots, ctx.Ts, want_w=True)
            ga = (proj.adjoint(w) * gf.to(torch.float32)).view(ctx.shape)
        if ctx.needs_input_grad[1]:
            _, g, _ = bhc.poissLI(Imea, ctx.ie, proj, ctx.Ts, ctx.knots, sraw=sraw)
            gi = (g * gf).to(ctx.ie_dtype).to(ctx.ie_device)
        return ga, gi, None, None, None, None


def project(img: torch.Tensor, proj: Projector) -> torch.Tensor:
    return _Project.apply(img, proj)


def backproject(sino: torch.Tensor, proj: Projector) -> torch.Tensor:
    return _Backproject.apply(sino, proj)


def poisson_nll(alpha: torch.Tensor, Ie: torch.Tensor, proj: Projector, Imea: torch.Tensor,
                knots, Ts: float = 1.0) -> torch.Tensor:
    return _PoissonNLL.apply(alpha, Ie, proj, Imea, knots, Ts)


__all__ = ["project", "backproject", "poisson_nll"]
