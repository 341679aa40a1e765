"""torch.autograd bindings over the C ABI.

``project(vol, g)`` is the Siddon forward projector A (ct_forward).  Its
backward is the matched adjoint A^T (ct_adjoint): an exact transpose of the
forward's intersection lengths, recomputed by re-traversing the rays, so
nothing per ray is stored between forward and backward.  ``tv_eps(vol, eps)`` is
TV_eps (ct_tv_eps) with its analytic gradient (ct_tv_eps_grad).

The reference has no differentiation; the closest routine is the matched
pixel-driven backprojector (CBCT/CBCT_1024V_1024P_multiGPU_v6.3_20170823/
pixel_driven_backprj/kernel_IterativeRecon_CBCT.cu:563), which the adjoint
follows.
"""
from __future__ import annotations

import torch

from .core import Context
from .geometry import Geometry

_DEFAULT: dict[int, Context] = {}


def default_context(device: int = 0) -> Context:
    if device not in _DEFAULT:
        _DEFAULT[device] = Context(device)
    return _DEFAULT[device]


class _Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vol: torch.Tensor, g: Geometry, ct: Context, view0: int, nviews: int,
                out_dtype: torch.dtype | None):
        ctx.g, ctx.ct, ctx.view0, ctx.nviews, ctx.vdtype = g, ct, view0, nviews, vol.dtype
        return ct.forward(g, vol.contiguous(), view0=view0, nviews=nviews, out_dtype=out_dtype)

    @staticmethod
    def backward(ctx, grad: torch.Tensor):
        gv = torch.empty(ctx.g.vol_shape, device=grad.device, dtype=ctx.vdtype)
        ctx.ct.adjoint(ctx.g, grad.contiguous(), vol=gv, view0=ctx.view0, nviews=ctx.nviews)
        return gv, None, None, None, None, None


class _TVEps(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vol: torch.Tensor, eps: float, ct: Context):
        ctx.save_for_backward(vol)
        ctx.eps, ctx.ct = eps, ct
        return torch.tensor(ct.tv_eps(vol.contiguous(), eps), device=vol.device, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad: torch.Tensor):
        (vol,) = ctx.saved_tensors
        g = ctx.ct.tv_eps_grad(vol.contiguous(), ctx.eps)
        return g * grad.to(g.dtype), None, None


def project(vol: torch.Tensor, g: Geometry, ct: Context | None = None, view0: int = 0,
            nviews: int | None = None, out_dtype: torch.dtype | None = None) -> torch.Tensor:
    """A x for views [view0, view0+nviews); differentiable (backward = A^T).
    The projections are stored in `out_dtype` (default: the volume's type);
    ray sums accumulate in FP32/FP64 either way."""
    ct = ct or default_context(vol.device.index or 0)
    nv = g.nviews - view0 if nviews is None else nviews
    return _Project.apply(vol, g, ct, view0, nv, out_dtype)


def tv_eps(vol: torch.Tensor, eps: float = 1e-8, ct: Context | None = None) -> torch.Tensor:
    """TV_eps(x) as an FP64 scalar tensor; differentiable."""
    ct = ct or default_context(vol.device.index or 0)
    return _TVEps.apply(vol, eps, ct)
