"""Seeded ellipsoid-in-cylinder phantoms and the standard scan geometries.

The random stream is SplitMix64 (Steele, Lea & Flood 2014) so the same seed
gives the same phantom on every machine.  Voxelisation evaluates the
indicator functions at 2x2x2 sub-samples per voxel and averages them.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def gauss(self) -> float:
        # Box-Muller
        u1 = self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(max(u1, 1e-300))) * math.cos(2.0 * math.pi * u2)


@dataclasses.dataclass
class Ellipsoid:
    center: tuple      # mm
    semi_axes: tuple   # mm
    rotation: np.ndarray  # 3x3, columns = ellipsoid axes in world frame
    mu: float


@dataclasses.dataclass
class Phantom:
    cyl_radius: float
    cyl_half_height: float
    cyl_mu: float
    ellipsoids: list


def _quat_to_rot(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def make_phantom(vol_mm, seed, n_ellipsoids=10, axes_range=(5.0, 50.0), mu_range=(0.01, 0.03),
                 cyl_mu=0.02) -> Phantom:
    """Cylinder of radius 0.45*L_xy and height 0.9*L_z (mu 0.02/mm) plus
    additive ellipsoids with uniform semi-axes, uniformly random rotations
    (normalised Gaussian quaternion) and centres uniform in the inner 60% of
    the cylinder."""
    lx, ly, lz = vol_mm
    rng = SplitMix64(seed)
    R = 0.45 * min(lx, ly)
    H = 0.45 * lz
    ells = []
    for _ in range(n_ellipsoids):
        axes = tuple(rng.uniform(*axes_range) for _ in range(3))
        q = np.array([rng.gauss() for _ in range(4)])
        q /= np.linalg.norm(q)
        r = 0.6 * R * math.sqrt(rng.uniform())
        phi = rng.uniform(0.0, 2.0 * math.pi)
        zc = rng.uniform(-0.6 * H, 0.6 * H)
        c = (r * math.cos(phi), r * math.sin(phi), zc)
        mu = rng.uniform(*mu_range)
        ells.append(Ellipsoid(c, axes, _quat_to_rot(q), mu))
    return Phantom(R, H, cyl_mu, ells)


def voxelize(ph: Phantom, vol_shape_xyz, vol_mm, device=None, dtype=None, supersample=2):
    """(nz, ny, nx) float32 torch tensor of the phantom, averaged over
    supersample^3 points per voxel.  Runs on ``device`` (CPU or GPU)."""
    import torch
    nx, ny, nz = vol_shape_xyz
    lx, ly, lz = vol_mm
    dev = torch.device(device) if device is not None else torch.device("cpu")
    out = torch.zeros((nz, ny, nx), dtype=torch.float32, device=dev)
    s = supersample
    offs = (torch.arange(s, dtype=torch.float64, device=dev) + 0.5) / s - 0.5
    xs = (((torch.arange(nx, dtype=torch.float64, device=dev)[:, None] + 0.5 + offs[None, :]) / nx)
          - 0.5).reshape(-1) * lx
    ys = (((torch.arange(ny, dtype=torch.float64, device=dev)[:, None] + 0.5 + offs[None, :]) / ny)
          - 0.5).reshape(-1) * ly
    zsall = (((torch.arange(nz, dtype=torch.float64, device=dev)[:, None] + 0.5 + offs[None, :]) / nz)
             - 0.5).reshape(-1) * lz
    Y, X = torch.meshgrid(ys.float(), xs.float(), indexing="ij")
    rr = X * X + Y * Y
    cyl_xy = (rr <= ph.cyl_radius ** 2)
    chunk = max(1, int(2 ** 24 // (X.numel())))  # z sub-slices per pass
    for z0 in range(0, nz, max(1, chunk // s)):
        z1 = min(nz, z0 + max(1, chunk // s))
        zs = zsall[z0 * s:z1 * s].float()
        acc = torch.zeros((zs.numel(),) + X.shape, dtype=torch.float32, device=dev)
        inz = (zs.abs() <= ph.cyl_half_height)[:, None, None]
        acc += ph.cyl_mu * (inz & cyl_xy[None]).float()
        for e in ph.ellipsoids:
            Rm = torch.tensor(e.rotation, dtype=torch.float32, device=dev)
            dx = X - e.center[0]
            dy = Y - e.center[1]
            val = 0.0
            for a in range(3):
                # coordinate along ellipsoid axis a: R[:, a] . (p - c)
                ca = Rm[0, a] * dx[None] + Rm[1, a] * dy[None] + Rm[2, a] * (zs[:, None, None] - e.center[2])
                val = val + (ca / e.semi_axes[a]) ** 2
            acc += e.mu * (val <= 1.0).float()
        blk = acc.reshape(z1 - z0, s, ny, s, nx, s).mean(dim=(1, 3, 5))
        out[z0:z1] = blk
    if dtype is not None:
        out = out.to(dtype)
    return out


@dataclasses.dataclass
class ScanConfig:
    name: str
    vol_shape: tuple   # (nx, ny, nz)
    vol_mm: tuple
    det_shape: tuple   # (nu, nv)
    det_mm: tuple
    nviews: int
    sad: float = 1000.0
    sdd: float = 1500.0
    arc_deg: float = 360.0
    seed: int = 1000

    def angles(self):
        # CBCT_forwards_projection convention: start + i*spacing, degrees
        # (xray/CBCT_forwards_projection.cpp:167-172)
        step = self.arc_deg / self.nviews
        return (np.arange(self.nviews, dtype=np.float32) * np.float32(step)).astype(np.float32)

    def offsets(self):
        return np.zeros((self.nviews, 2), np.float32)

    def geom(self):
        from .capi import Geom
        return Geom.make(self.vol_shape, self.vol_mm, self.det_shape, self.det_mm, self.sad,
                         self.sdd)

    def phantom(self):
        return make_phantom(self.vol_mm, self.seed)


# The standard workloads: FDK at 128^3 / 180 views; OS reconstruction at
# 256^3 / 360 views; the 512^3 / 720-view precision case.  Detector pitch
# keeps the 199.68 mm / 128 mm detector-to-volume width ratio of the first.
FDK128 = ScanConfig("fdk128", (128, 128, 128), (128.0, 128.0, 128.0), (256, 192),
                    (199.68, 149.76), 180, seed=1000)
OS256 = ScanConfig("os256", (256, 256, 256), (204.8, 204.8, 204.8), (512, 384),
                   (512 * 0.624, 384 * 0.624), 360, seed=1001)
PREC512 = ScanConfig("prec512", (512, 512, 512), (256.0, 256.0, 256.0), (1024, 768),
                     (1024 * 0.39, 768 * 0.39), 720, seed=1002)
CONFIGS = {c.name: c for c in (FDK128, OS256, PREC512)}


def tiny_config(nx=16, nu=24, nv=20, nviews=12, seed=7, arc=360.0, vol_mm=None):
    """A small scan for parity tests (seconds for the FP64 oracle)."""
    vol_mm = vol_mm or (float(nx), float(nx), float(nx))
    ratio = 199.68 / 128.0
    return ScanConfig("tiny%d" % nx, (nx, nx, nx), vol_mm, (nu, nv),
                      (vol_mm[0] * ratio * nu / (2 * nx), vol_mm[0] * ratio * nv / (2 * nx) * 1.0),
                      nviews, arc_deg=arc, seed=seed)
