"""Shared test helpers: geometry conversion between the product (gtt.capi.Geom)
and the oracle (oracle.Geom), seeded inputs, error metrics."""
import numpy as np

import oracle as O
from gtt import capi
from gtt import phantom as PH


def to_oracle(g: capi.Geom) -> O.Geom:
    return O.Geom(g.nx, g.ny, g.nz, g.lx, g.ly, g.lz, g.nu, g.nv, g.fov_u, g.fov_v, g.sad, g.sdd)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def geom(nx, ny, nz, nu, nv, vol_mm=None, ratio=1.56, sad=1000.0, sdd=1500.0, fov=None):
    vol_mm = vol_mm or (float(nx), float(ny), float(nz))
    if fov is None:
        pitch = vol_mm[0] * ratio / nu
        fov = (pitch * nu, pitch * nv)
    return capi.Geom.make((nx, ny, nz), vol_mm, (nu, nv), fov, sad, sdd)


def phantom_volume(g: capi.Geom, seed=11):
    ph = PH.make_phantom((g.lx, g.ly, g.lz), seed)
    return PH.voxelize(ph, (g.nx, g.ny, g.nz), (g.lx, g.ly, g.lz)).numpy().astype(np.float64)


def angles_full(n, arc=360.0, start=0.0):
    return (np.float32(start) + np.arange(n, dtype=np.float32) * np.float32(arc / n)).astype(np.float32)


def bf16_round(x):
    """Round-to-nearest-even float32 -> bfloat16 -> float32 (numpy)."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)
