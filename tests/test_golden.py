"""Golden fixtures (tests/golden/*.npz, written by make_golden.py from the FP64
oracle on small seeded cases).

CPU: the oracle still reproduces them (regression pin of the oracle).
GPU: the gfx950 kernels reproduce them through the C ABI (rel-L2 <= 1e-5).
"""
import glob
import os

import numpy as np
import pytest

import oracle as O
from gtt import capi
from helpers import rel_l2

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(glob.glob(os.path.join(HERE, "*.npz")))


def load(path):
    d = dict(np.load(path, allow_pickle=False))
    v = d["geom"]
    g = O.Geom(int(v[0]), int(v[1]), int(v[2]), v[3], v[4], v[5], int(v[6]), int(v[7]), v[8], v[9],
               v[10], v[11])
    return g, d


def test_fixtures_present():
    assert len(FIXTURES) >= 2


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_reproduces_golden(path):
    g, d = load(path)
    vol = d["volume"].astype(np.float64)
    p = O.fp(g, d["angles"], d["offsets"], vol)
    assert np.allclose(p, d["projections"], rtol=1e-12, atol=1e-14)
    b = O.bp(g, d["angles"], d["offsets"], d["projections"])
    assert np.allclose(b, d["backprojection"], rtol=1e-12, atol=1e-14)
    f = O.fdk_prefilter(g, d["angles"], d["offsets"], d["projections"], d["filter"])
    assert np.allclose(f, d["filtered"], rtol=1e-12, atol=1e-14)
    fb = O.bp(g, d["angles"], d["offsets"], d["filtered"], fbp=True)
    assert np.allclose(fb, d["fdk_backprojection"], rtol=1e-12, atol=1e-14)


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_gpu_reproduces_golden(path):
    torch = pytest.importorskip("torch")
    g, d = load(path)
    print("CASE golden", os.path.basename(path), flush=True)
    cg = capi.Geom.make((g.nx, g.ny, g.nz), (g.lx, g.ly, g.lz), (g.nu, g.nv), (g.fu, g.fv), g.sad, g.sdd)
    ctx = capi.Context(0)
    ctx.set_views(d["angles"], d["offsets"])
    ctx.set_filter(d["filter"])
    vol = torch.from_numpy(d["volume"]).cuda()
    proj = torch.zeros(cg.proj_shape(len(d["angles"])), device="cuda")
    ctx.fp(cg, vol, proj)
    pin = torch.from_numpy(d["projections"].astype(np.float32)).cuda()
    bpv = torch.zeros_like(vol)
    ctx.bp(cg, pin, bpv)
    filt = torch.empty_like(pin)
    ctx.fdk_prefilter(cg, pin, filt)
    fbv = torch.zeros_like(vol)
    ctx.bp(cg, filt, fbv, fbp=True)
    torch.cuda.synchronize()
    assert rel_l2(proj.cpu().numpy(), d["projections"]) <= 1e-5
    assert rel_l2(bpv.cpu().numpy(), d["backprojection"]) <= 1e-5
    assert rel_l2(filt.cpu().numpy(), d["filtered"]) <= 1e-5
    assert rel_l2(fbv.cpu().numpy(), d["fdk_backprojection"]) <= 1e-5
    ctx.close()
