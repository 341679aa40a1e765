"""Differentiation through the C ABI: torch.autograd with the matched adjoint
as the projector's backward and the analytic TV_eps gradient, checked by the
dot-product identity and by FP64 central finite differences (64^3 volume,
90 views of 96^2, x and b drawn N(0,1) from a seed)."""
import numpy as np
import pytest
import torch

from gpu_amd.autograd import project, tv_eps
from gpu_amd.geometry import adjoint_64
from gpu_amd.phantom import phantom
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


def _adam_cpu(g, b, lam, eps, lr, iters):
    """The same Adam recurrence (torch defaults: betas 0.9/0.999, eps 1e-8) in
    FP64 on the oracle's forward, matched adjoint and TV_eps gradient."""
    x = np.zeros(g.vol_shape)
    m = np.zeros_like(x)
    v = np.zeros_like(x)
    losses = []
    for k in range(1, iters + 1):
        r = orc.forward(g, x, prec=64) - b
        losses.append(0.5 * float((r * r).sum()) + lam * orc.tv_eps(x, eps, prec=64))
        grad = orc.adjoint(g, r, prec=64) + lam * orc.tv_eps_grad(x, eps, prec=64)
        m = 0.9 * m + 0.1 * grad
        v = 0.999 * v + 0.001 * grad * grad
        mh = m / (1 - 0.9 ** k)
        vh = v / (1 - 0.999 ** k)
        x = x - lr * mh / (np.sqrt(vh) + 1e-8)
    return x, losses


@pytest.mark.parametrize("storage,loss_tol,tol", [(torch.float32, 1e-4, 2e-2), (torch.bfloat16, 2e-2, 6e-2)])
def test_multisource_adam_vs_fp64_cpu(ctx, storage, loss_tol, tol):
    """Gradient reconstruction in the multi-source geometry at reduced scale
    (24^3 at 0.75 mm, 6 sources 3 mm apart, 12 views of 64x16 at 0.5 mm):
    Adam (lr 1e-3) on 0.5||Ax - b||^2 + 1e-3 TV_eps through the autograd
    projector (forward = ray sums, backward = matched adjoint), volume stored
    in `storage` with FP32 master weights, against the same recurrence in FP64
    on the CPU oracle.  The loss trajectories agree to `loss_tol` at every
    iteration.  The volumes agree to `tol`, a looser bar: Adam's normalised
    step turns rounding-level gradients of voxels that no ray constrains
    into lr-sized steps, which dominate the volume difference (FP32 ~1e-2)."""
    from gpu_amd.geometry import Geometry
    g = Geometry(nx=24, ny=24, nz=24, vx=0.75, ncols=64, nrows=16, du=0.5, nviews=12,
                 nsources=6, src_z0=-7.5, src_dz=3.0)
    x_true = phantom(g, seed=31, scale=0.15).numpy().astype(np.float64)
    b = orc.forward(g, x_true, prec=64)
    lam, eps, lr, iters = 1e-3, 1e-8, 1e-3, 30
    x_ref, loss_ref = _adam_cpu(g, b, lam, eps, lr, iters)

    bt = torch.as_tensor(b, dtype=torch.float32, device="cuda")
    x = torch.zeros(g.vol_shape, device="cuda", dtype=torch.float32, requires_grad=True)
    opt = torch.optim.Adam([x], lr=lr)
    losses = []
    for _ in range(iters):
        opt.zero_grad()
        xs = x.to(storage)
        r = project(xs, g, ctx, out_dtype=torch.float32) - bt
        loss = 0.5 * (r.double() * r.double()).sum() + lam * tv_eps(xs, eps, ctx)
        loss.backward()
        losses.append(float(loss))
        opt.step()
    got = x.detach().double().cpu().numpy()
    err = np.linalg.norm(got - x_ref) / np.linalg.norm(x_ref)
    print(f"{storage}: rel L2 vs FP64 CPU Adam {err:.3e}; loss {losses[0]:.4e} -> {losses[-1]:.4e} "
          f"(CPU {loss_ref[0]:.4e} -> {loss_ref[-1]:.4e})")
    lerr = max(abs(a - b) / abs(b) for a, b in zip(losses, loss_ref))
    print(f"max relative loss difference {lerr:.3e}")
    assert losses[-1] < 0.5 * losses[0]
    assert lerr <= loss_tol
    assert err <= tol
