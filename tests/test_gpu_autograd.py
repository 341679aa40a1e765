"""Differentiation through the C ABI: torch.autograd with the matched adjoint
as the projector's backward and the analytic TV_eps gradient, checked by the
dot-product identity and by FP64 central finite differences (64^3 volume,
90 views of 96^2, x and b drawn N(0,1) from a seed)."""
import numpy as np
import pytest
import torch

from gpu_amd.autograd import project, tv_eps
from gpu_amd.geometry import adjoint_64
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def objective(ctx, g, x, b, lam, eps):
    r = project(x, g, ctx) - b
    return 0.5 * (r * r).sum() + lam * tv_eps(x, eps, ctx)


def test_backward_is_adjoint(ctx):
    g = adjoint_64()
    gen = torch.Generator(device="cuda").manual_seed(11)
    x = torch.randn(g.vol_shape, device="cuda", dtype=torch.float64, generator=gen, requires_grad=True)
    y = torch.randn(g.proj_shape, device="cuda", dtype=torch.float64, generator=gen)
    ax = project(x, g, ctx)
    lhs = float((ax * y).sum())
    (gx,) = torch.autograd.grad(ax, x, grad_outputs=y)
    rhs = float((x.detach() * gx).sum())
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_gradient_matches_central_differences(ctx):
    """Central differences (h = 1e-6) at 200 seeded voxels.  The point is taken
    near the data (b = A x0, x = x0 + 0.01 n, x0, n ~ N(0,1)) so the objective is
    ~1e4: the FP64 rounding of f then contributes ~1e-10/h to each quotient,
    and the bound 1e-4 checks the gradient, not the rounding of f (at x, b ~
    N(0,1) the objective is ~3e7 and rounding alone gives ~1e-3)."""
    g = adjoint_64()
    lam, eps, h = 1e-2, 1e-8, 1e-6
    gen = torch.Generator(device="cuda").manual_seed(5)
    x0 = torch.randn(g.vol_shape, device="cuda", dtype=torch.float64, generator=gen)
    b = project(x0, g, ctx).detach()
    x = x0 + 0.01 * torch.randn(g.vol_shape, device="cuda", dtype=torch.float64, generator=gen)
    xr = x.clone().requires_grad_(True)
    f = objective(ctx, g, xr, b, lam, eps)
    (grad,) = torch.autograd.grad(f, xr)
    rng = np.random.default_rng(0)
    idx = rng.choice(g.nvox, size=200, replace=False)
    flat = x.view(-1)
    errs = []
    with torch.no_grad():
        for i in idx.tolist():
            old = float(flat[i])
            flat[i] = old + h
            fp = float(objective(ctx, g, x, b, lam, eps))
            flat[i] = old - h
            fm = float(objective(ctx, g, x, b, lam, eps))
            flat[i] = old
            fd = (fp - fm) / (2 * h)
            an = float(grad.view(-1)[i])
            errs.append(abs(fd - an) / max(abs(fd), 1e-3))
    print(f"max rel err {max(errs):.3e}, median {np.median(errs):.3e}")
    assert max(errs) < 1e-4
    assert float(np.median(errs)) < 1e-5


def test_gradient_matches_fp64_cpu_oracle(ctx):
    """The autograd gradient equals A^T(Ax-b) + lam*grad TV_eps from the CPU
    FP64 oracle."""
    g = adjoint_64().with_views(18)
    lam, eps = 1e-2, 1e-8
    rng = np.random.default_rng(3)
    x = rng.standard_normal(g.vol_shape)
    b = rng.standard_normal(g.proj_shape)
    xr = torch.as_tensor(x, device="cuda").requires_grad_(True)
    f = objective(ctx, g, xr, torch.as_tensor(b, device="cuda"), lam, eps)
    (grad,) = torch.autograd.grad(f, xr)
    ref = orc.adjoint(g, orc.forward(g, x, prec=64) - b, prec=64) + lam * orc.tv_eps_grad(x, eps)
    got = grad.cpu().numpy()
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-12
