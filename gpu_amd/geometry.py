"""Runtime geometry (the ct_geometry struct of include/ct_core.h) and the
named synthetic workloads.

The reference fixes its geometry as compile-time constants
(CBCT/CBCT_1024V_1024P_multiGPU_v6.3_20170823/main_recon.cu:54-111); here it is
a value passed with every call.  Lengths are millimetres for the workloads
below; SID 1000 mm / SDD 1500 mm map to dso = 1000, dod = -500 (detector
across the isocentre).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field, replace


class CtGeometry(ctypes.Structure):
    """Field-for-field mirror of ct_geometry (include/ct_core.h)."""

    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("x_offset", ctypes.c_int32),
        ("vx", ctypes.c_double), ("vy", ctypes.c_double), ("vz", ctypes.c_double),
        ("ncols", ctypes.c_int32), ("nrows", ctypes.c_int32),
        ("du", ctypes.c_double), ("dv", ctypes.c_double),
        ("det_offset", ctypes.c_double),
        ("dso", ctypes.c_double), ("dod", ctypes.c_double),
        ("nviews", ctypes.c_int32), ("nsources", ctypes.c_int32),
        ("angle_step_deg", ctypes.c_double), ("angle0_deg", ctypes.c_double),
        ("angle_shift_deg", ctypes.c_double),
        ("src_z0", ctypes.c_double), ("src_dz", ctypes.c_double),
        ("style", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


@dataclass(frozen=True)
class Geometry:
    nx: int
    ny: int
    nz: int
    vx: float
    ncols: int
    nrows: int
    du: float
    nviews: int
    dso: float = 1000.0
    dod: float = -500.0
    vy: float | None = None
    vz: float | None = None
    dv: float | None = None
    x_offset: int = 0
    det_offset: float = 0.0
    nsources: int = 1
    angle_step_deg: float | None = None   # default: nviews evenly over 360 degrees
    angle0_deg: float = 0.0
    angle_shift_deg: float = 0.0
    src_z0: float = 0.0
    src_dz: float = 0.0
    style: int = 0                        # 0 cone beam, 1 multi-slice fan beam, 2 parallel beam

    def to_c(self) -> CtGeometry:
        step = self.angle_step_deg if self.angle_step_deg is not None else 360.0 / max(self.nviews, 1)
        return CtGeometry(
            self.nx, self.ny, self.nz, self.x_offset,
            self.vx, self.vy if self.vy is not None else self.vx, self.vz if self.vz is not None else self.vx,
            self.ncols, self.nrows,
            self.du, self.dv if self.dv is not None else self.du,
            self.det_offset, self.dso, self.dod,
            self.nviews, self.nsources, step, self.angle0_deg, self.angle_shift_deg,
            self.src_z0, self.src_dz, self.style, 0,
        )

    @property
    def vol_shape(self) -> tuple[int, int, int]:
        return (self.nz, self.ny, self.nx)

    @property
    def proj_shape(self) -> tuple[int, ...]:
        if self.nsources > 1:
            return (self.nviews, self.nsources, self.nrows, self.ncols)
        return (self.nviews, self.nrows, self.ncols)

    @property
    def nvox(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def rays(self) -> int:
        return self.nviews * self.nsources * self.nrows * self.ncols

    def with_views(self, nviews: int, keep_step: bool = True) -> "Geometry":
        step = self.angle_step_deg if self.angle_step_deg is not None else 360.0 / self.nviews
        return replace(self, nviews=nviews, angle_step_deg=step if keep_step else None)


# Workloads (sizes as the synthetic benchmark configurations state them).
def plain_sart_128() -> Geometry:
    """128^3 at 1 mm, 180 views over 360 deg onto 256x256 at 1 mm."""
    return Geometry(nx=128, ny=128, nz=128, vx=1.0, ncols=256, nrows=256, du=1.0, nviews=180)


def repo_scale_512() -> Geometry:
    """512^3 at 0.5 mm, 1024 views over 360 deg onto 1024x1024 at 0.4 mm."""
    return Geometry(nx=512, ny=512, nz=512, vx=0.5, ncols=1024, nrows=1024, du=0.4, nviews=1024)


def adjoint_64() -> Geometry:
    """64^3 at 1 mm, 90 views onto 96x96 at 1 mm."""
    return Geometry(nx=64, ny=64, nz=64, vx=1.0, ncols=96, nrows=96, du=1.0, nviews=90)


def multisource_256() -> Geometry:
    """256^3 at 0.75 mm; 48 sources 3 mm apart on the rotation axis, 720 views,
    detector 1024 columns x 128 rows at 0.5 mm."""
    return Geometry(nx=256, ny=256, nz=256, vx=0.75, ncols=1024, nrows=128, du=0.5, nviews=720,
                    nsources=48, src_z0=-3.0 * 23.5, src_dz=3.0)


WORKLOADS = {
    "plain_sart_128": plain_sart_128,
    "repo_scale_512": repo_scale_512,
    "adjoint_64": adjoint_64,
    "multisource_256": multisource_256,
}
