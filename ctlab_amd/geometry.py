"""Scan geometry in the argument vocabulary of CTLAB's ``DD3Proj_gpu`` /
``DD3Back_gpu`` / ``DD2_multiGPU``.

``ConeGeometry`` carries exactly the geometry arguments of
``DD3Proj_gpu`` (src/DD3_GPU_Proj.h:16-44): the view-0 source position
``(x0, y0, z0)``, view-0 detector cell centres ``xds, yds`` (DNU) and ``zds``
(DNV), the image centre, ``hangs`` / ``hzPos`` per view, volume sizes, voxel
pitches ``dx`` (in-plane, square) and ``dz``, and the XY ``mask``.  Arrays are
float32, as the reference ABI passes them.

``arc_detector`` / ``flat_detector`` build the cell centres the way the
reference's OS_SART driver is parameterised (sid, sdd, col/row size and
offsets; src/DD3_GPU_recon.h:41-65), with the arc and flat detector frames of
src/utilities.hpp:3294-3326 (arc cell at angle g: (SDD sin g, SID - SDD cos g);
flat detector at y = -(SDD - SID)).
"""
from dataclasses import dataclass, field
from typing import Optional

import numpy as np


@dataclass
class ConeGeometry:
    x0: float
    y0: float
    z0: float
    DNU: int
    DNV: int
    xds: np.ndarray
    yds: np.ndarray
    zds: np.ndarray
    imgXCenter: float
    imgYCenter: float
    imgZCenter: float
    angs: np.ndarray
    zpos: np.ndarray
    XN: int
    YN: int
    ZN: int
    dx: float
    dz: float
    mask: Optional[np.ndarray] = None
    parallel: int = 0          # 1: multi-slice fan beam (DD2_multiGPU)

    def __post_init__(self):
        f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
        self.xds, self.yds, self.zds = f32(self.xds), f32(self.yds), f32(self.zds)
        self.angs, self.zpos = f32(self.angs), f32(self.zpos)
        if self.mask is not None:
            self.mask = np.ascontiguousarray(self.mask, dtype=np.uint8).reshape(-1)
            assert self.mask.size == self.XN * self.YN

    @property
    def PN(self):
        return int(self.angs.size)

    @property
    def vol_shape(self):
        """(YN, XN, ZN): the (Z,X,Y) order of the reference, z fastest."""
        return (self.YN, self.XN, self.ZN)

    @property
    def prj_shape(self):
        return (self.PN, self.DNU, self.DNV)

    def nvox(self):
        return self.XN * self.YN * self.ZN

    def nprj(self):
        return self.PN * self.DNU * self.DNV

    def subset(self, views):
        """Geometry restricted to the given view indices (OS-SART subsets)."""
        views = np.asarray(views)
        g = ConeGeometry(**{k: getattr(self, k) for k in self.__dataclass_fields__})
        g.angs = self.angs[views].copy()
        g.zpos = self.zpos[views].copy()
        return g


def arc_detector(sid, sdd, DNU, DNV, col_size, row_size, col_offset=0.0, row_offset=0.0):
    """Equiangular arc (centred on the source) x flat rows; returns xds, yds, zds."""
    dgam = col_size / sdd
    u = np.arange(DNU, dtype=np.float64) - ((DNU - 1) / 2.0 + col_offset)
    gam = u * dgam
    xds = sdd * np.sin(gam)
    yds = sid - sdd * np.cos(gam)
    zds = (np.arange(DNV, dtype=np.float64) - ((DNV - 1) / 2.0 + row_offset)) * row_size
    return xds.astype(np.float32), yds.astype(np.float32), zds.astype(np.float32)


def flat_detector(sid, sdd, DNU, DNV, col_size, row_size, col_offset=0.0, row_offset=0.0):
    u = np.arange(DNU, dtype=np.float64) - ((DNU - 1) / 2.0 + col_offset)
    xds = u * col_size
    yds = np.full(DNU, sid - sdd)
    zds = (np.arange(DNV, dtype=np.float64) - ((DNV - 1) / 2.0 + row_offset)) * row_size
    return xds.astype(np.float32), yds.astype(np.float32), zds.astype(np.float32)


def circular_cone(XN, YN, ZN, dx, dz, PN, DNU, DNV, col_size, row_size, sid=541.0, sdd=949.0,
                  arc=True, mask=None, ang0=0.0, col_offset=0.0, row_offset=0.0):
    """Circular cone-beam scan, source at (0, sid, 0) at view 0, views over 2*pi."""
    det = arc_detector if arc else flat_detector
    xds, yds, zds = det(sid, sdd, DNU, DNV, col_size, row_size, col_offset, row_offset)
    angs = ang0 + np.arange(PN, dtype=np.float64) * (2.0 * np.pi / PN)
    return ConeGeometry(x0=0.0, y0=sid, z0=0.0, DNU=DNU, DNV=DNV, xds=xds, yds=yds, zds=zds,
                        imgXCenter=0.0, imgYCenter=0.0, imgZCenter=0.0, angs=angs,
                        zpos=np.zeros(PN), XN=XN, YN=YN, ZN=ZN, dx=dx, dz=dz, mask=mask)


def helical_cone(XN, YN, ZN, dx, dz, PN, DNU, DNV, col_size, row_size, pitch, turns=1.0, sid=541.0, sdd=949.0,
                 arc=True, mask=None, ang0=0.0, col_offset=0.0, row_offset=0.0):
    """Helical cone-beam scan: PN views over `turns` rotations while the source
    and detector advance `pitch` mm along z per rotation (hzPos, the per-view
    z offset of DD3Proj_gpu), centred on the volume."""
    g = circular_cone(XN, YN, ZN, dx, dz, PN, DNU, DNV, col_size, row_size, sid, sdd, arc, mask, ang0,
                      col_offset, row_offset)
    t = np.arange(PN, dtype=np.float64) * (turns / PN)        # rotations elapsed at view p
    g.angs = (ang0 + 2.0 * np.pi * t).astype(np.float32)
    g.zpos = (pitch * (t - 0.5 * turns)).astype(np.float32)
    return g


def fan_multislice(XN, YN, SLN, dx, PN, DNU, col_size, sid=541.0, sdd=949.0, arc=True,
                   mask=None, ang0=0.0, col_offset=0.0):
    """Multi-slice fan beam (DD2_multiGPU): SLN independent 2-D slices."""
    det = arc_detector if arc else flat_detector
    xds, yds, _ = det(sid, sdd, DNU, 1, col_size, 1.0, col_offset)
    angs = ang0 + np.arange(PN, dtype=np.float64) * (2.0 * np.pi / PN)
    return ConeGeometry(x0=0.0, y0=sid, z0=0.0, DNU=DNU, DNV=SLN, xds=xds, yds=yds,
                        zds=np.arange(SLN, dtype=np.float32), imgXCenter=0.0, imgYCenter=0.0,
                        imgZCenter=0.0, angs=angs, zpos=np.zeros(PN), XN=XN, YN=YN, ZN=SLN,
                        dx=dx, dz=1.0, mask=mask, parallel=1)
