"""Writes the golden fixtures of this directory from the FP64 oracle.

The reference ships no golden vectors for this path, so these fixtures are
the oracle's outputs on small seeded cases.  They pin the oracle against
regressions (tests/test_golden.py recomputes and compares) and give the GPU
tests fixed expected outputs.  Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "gt-tomography_amd"))

import oracle as O  # noqa: E402
from gtt import phantom as PH  # noqa: E402


def cases():
    """(name, oracle geometry, angles, offsets, seed) of the fixtures."""
    out = []
    g = O.Geom(16, 16, 16, 16.0, 16.0, 16.0, 24, 20, 24.96, 20.8, 1000.0, 1500.0)
    out.append(("cube16", g, (np.arange(12) * 30.0).astype(np.float32), None, 21))
    g = O.Geom(13, 11, 9, 13.0, 11.0, 13.5, 21, 17, 30.0, 24.0, 500.0, 800.0)
    a = (5.0 + np.arange(9) * 40.0).astype(np.float32)
    r = np.random.default_rng(4)
    o = np.stack([r.uniform(-2, 2, 9), r.uniform(-1, 1, 9)], -1).astype(np.float32)
    out.append(("ragged_offsets", g, a, o, 22))
    return out


def build(name, g, angles, offsets, seed):
    ph = PH.make_phantom((g.lx, g.ly, g.lz), seed, axes_range=(1.0, 6.0))
    vol = PH.voxelize(ph, (g.nx, g.ny, g.nz), (g.lx, g.ly, g.lz)).numpy().astype(np.float64)
    proj = O.fp(g, angles, offsets, vol)
    bp = O.bp(g, angles, offsets, proj)
    F = O.ramlak_filter(g.nu).astype(np.float32)
    filt = O.fdk_prefilter(g, angles, offsets, proj, F)
    fbp = O.bp(g, angles, offsets, filt, fbp=True)
    geom = np.array([g.nx, g.ny, g.nz, g.lx, g.ly, g.lz, g.nu, g.nv, g.fu, g.fv, g.sad, g.sdd])
    return dict(geom=geom, angles=angles,
                offsets=np.zeros((len(angles), 2), np.float32) if offsets is None else offsets,
                volume=vol.astype(np.float32), projections=proj, backprojection=bp,
                filtered=filt, fdk_backprojection=fbp, filter=F)


def main():
    for name, g, a, o, seed in cases():
        d = build(name, g, a, o, seed)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
        print("wrote", name, {k: v.shape for k, v in d.items()})


if __name__ == "__main__":
    main()
