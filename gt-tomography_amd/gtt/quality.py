"""Image-quality metric of the reference: multi-scale SSIM (xray/mssim.cpp:77-121)
and the per-phase average of the compare_images program
(xray/compare_images.cpp:47-70), on the gfx950 kernel cbct_mssim."""
from __future__ import annotations

import numpy as np

from . import capi

_ctx = None


def _context():
    global _ctx
    if _ctx is None:
        _ctx = capi.Context(0)
    return _ctx


def _device(a):
    import torch
    if isinstance(a, torch.Tensor):
        return a.to(device="cuda", dtype=torch.float32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def mssim(image, reference, sigma=1.5, scales=3, ctx: capi.Context | None = None):
    """MS-SSIM of a volume (nz, ny, nx) against a reference volume; numpy
    arrays or torch tensors.  sigma: voxels, one value or (sx, sy, sz)."""
    c = ctx or _context()
    return c.mssim(_device(image), _device(reference), sigma, scales)


def compare_images(image, gold, sigma=1.5, ctx: capi.Context | None = None):
    """compare_images: the MS-SSIM of every phase (volumes stacked as
    (phases, nz, ny, nx)) against the gold standard's, averaged in float32
    as the program does.  Returns (average, per-phase values)."""
    image = np.asarray(image)
    gold = np.asarray(gold)
    if image.ndim != 4 or image.shape != gold.shape:
        raise ValueError("compare_images: image and gold must be 4-D arrays of one shape")
    res = [mssim(image[p], gold[p], (sigma,) * 3, 3, ctx) for p in range(image.shape[0])]
    avg = np.float32(0.0)
    for r in res:
        avg = np.float32(avg + np.float32(r))
    return float(np.float32(avg / np.float32(len(res)))), res
