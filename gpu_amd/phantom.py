"""Seeded synthetic inputs: the ellipsoid phantom and Poisson noise.

The reference loads a phantom file (CBCT/CBCT_1024V_1024P_multiGPU_v6.3_20170823/
main_recon.cu:231-243) that it does not ship; these inputs are this build's own
definition:

* 10 ellipsoids, centres U(-40, 40) mm, semi-axes U(5, 40) mm, both times
  ``scale``; orientation R = Rz(a) Ry(b) Rz(c) with (a, b, c) ~ U(0, 2 pi)
  (z-y-z Euler angles); attenuation U(0.002, 0.02) / mm, added where
  sum_k ((R^T (p - c))_k / a_k)^2 <= 1 at the voxel centre p.
* Parameters come from numpy's PCG64 generator seeded with ``seed``, drawn in
  the order centres, semi-axes, angles, attenuations.
* Noise: counts ~ Poisson(I0 * exp(-p)); p_noisy = -ln(max(counts, 1) / I0).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .geometry import Geometry


def ellipsoid_params(seed: int = 0, n: int = 10, scale: float = 1.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    centres = rng.uniform(-40.0, 40.0, (n, 3)) * scale
    axes = rng.uniform(5.0, 40.0, (n, 3)) * scale
    angles = rng.uniform(0.0, 2.0 * math.pi, (n, 3))
    mu = rng.uniform(0.002, 0.02, n)
    return centres, axes, angles, mu


def rotation_zyz(a: float, b: float, c: float) -> np.ndarray:
    def rz(t):
        return np.array([[math.cos(t), -math.sin(t), 0.0], [math.sin(t), math.cos(t), 0.0], [0.0, 0.0, 1.0]])

    def ry(t):
        return np.array([[math.cos(t), 0.0, math.sin(t)], [0.0, 1.0, 0.0], [-math.sin(t), 0.0, math.cos(t)]])

    return rz(a) @ ry(b) @ rz(c)


def voxel_centres(g: Geometry):
    vy = g.vy if g.vy is not None else g.vx
    vz = g.vz if g.vz is not None else g.vx
    bx = -g.vx * (g.nx / 2.0 + g.x_offset)
    by = -vy * (g.ny / 2.0)
    bz = -vz * (g.nz / 2.0)
    x = bx + (np.arange(g.nx) + 0.5) * g.vx
    y = by + (np.arange(g.ny) + 0.5) * vy
    z = bz + (np.arange(g.nz) + 0.5) * vz
    return x, y, z


def phantom(g: Geometry, seed: int = 0, scale: float = 1.0, device: str | torch.device = "cpu",
            dtype: torch.dtype = torch.float32, n: int = 10) -> torch.Tensor:
    """Volume [nz, ny, nx] of the seeded ellipsoid phantom (evaluated in FP64)."""
    centres, axes, angles, mu = ellipsoid_params(seed, n, scale)
    xs, ys, zs = voxel_centres(g)
    dev = torch.device(device)
    X = torch.as_tensor(xs, device=dev, dtype=torch.float64)
    Y = torch.as_tensor(ys, device=dev, dtype=torch.float64)
    out = torch.zeros(g.vol_shape, device=dev, dtype=dtype)
    YY, XX = torch.meshgrid(Y, X, indexing="ij")
    for e in range(n):
        Rt = torch.as_tensor(rotation_zyz(*angles[e]).T, device=dev, dtype=torch.float64)
        c = centres[e]
        inv_a2 = torch.as_tensor(1.0 / axes[e] ** 2, device=dev, dtype=torch.float64)
        dx, dy = XX - c[0], YY - c[1]
        for k, zv in enumerate(zs):
            dz = float(zv - c[2])
            if abs(dz) > axes[e].max():
                continue
            q = torch.zeros_like(dx)
            for r in range(3):
                comp = Rt[r, 0] * dx + Rt[r, 1] * dy + Rt[r, 2] * dz
                q += comp * comp * inv_a2[r]
            out[k] += (q <= 1.0).to(dtype) * float(mu[e])
    return out


def poisson_log(proj: torch.Tensor, i0: float = 1e5, seed: int = 0) -> torch.Tensor:
    """Poisson counts of I0*exp(-p) then the log transform, in FP32."""
    gen = torch.Generator(device=proj.device)
    gen.manual_seed(seed)
    rate = i0 * torch.exp(-proj.to(torch.float32))
    counts = torch.poisson(rate, generator=gen)
    return -torch.log(torch.clamp(counts, min=1.0) / i0)
