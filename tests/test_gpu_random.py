"""Randomised parity of the projectors on the GPU: seeded small geometries
(odd sizes, anisotropic voxels, magnifications from 1.2 to 2, detectors
narrower and wider than the volume's shadow, detector offsets, short arcs).
For every case:
  * every forward-projector engine (per ray; the column variants; several
    launch groups of adjacent views) gives the same bits;
  * FP, and BP plain and FDK-weighted, match the FP64 oracle within
    rel-L2 1e-5;
  * the backprojector engines give the same bits;
  * with BF16 working copies, the engines again agree bit for bit.
"""
import numpy as np
import pytest

import oracle as O
from gtt import capi
from helpers import angles_full, phantom_volume, rel_l2, to_oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TOL = 1e-5
N_CASES = 16


def random_case(seed):
    r = np.random.default_rng(1000 + seed)
    nx, ny, nz = (int(v) for v in r.integers(5, 33, 3))
    vol_mm = tuple(float(n * r.uniform(0.5, 1.5)) for n in (nx, ny, nz))
    nu, nv = int(r.integers(8, 80)), int(r.integers(6, 72))
    pitch = vol_mm[0] * r.uniform(0.8, 2.2) / nu
    fov = (pitch * nu, pitch * nv * r.uniform(0.6, 1.6))
    sad = float(r.uniform(400.0, 1500.0))
    sdd = sad * float(r.uniform(1.2, 2.0))
    g = capi.Geom.make((nx, ny, nz), vol_mm, (nu, nv), fov, sad, sdd)
    nviews = int(r.integers(3, 23))
    arc = 360.0 if r.random() < 0.6 else float(r.uniform(90.0, 300.0))
    angles = angles_full(nviews, arc, start=float(r.uniform(0.0, 360.0)))
    offsets = None
    if r.random() < 0.5:
        offsets = np.stack([r.uniform(-0.2, 0.2, nviews) * fov[0],
                            r.uniform(-0.1, 0.1, nviews) * fov[1]], -1).astype(np.float32)
    return g, angles, offsets


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.float().cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("seed", range(N_CASES))
def test_random_geometry_parity(ctx, seed):
    g, angles, offsets = random_case(seed)
    print("CASE random", seed, g, len(angles), "offsets" if offsets is not None else "", flush=True)
    vol = phantom_volume(g, seed=seed) + 0.01
    ctx.set_views(angles, offsets)

    def fp_all(v):
        outs = []
        try:
            for e, vg in ((0, 6), (1, 1), (1, 6), (2, 6), (3, 5), (4, 3)):
                ctx.fp_engine(e)
                ctx.fp_view_group(vg)
                out = torch.zeros(g.proj_shape(len(angles)), device="cuda")
                ctx.fp(g, v, out)
                outs.append(out)
        finally:
            ctx.fp_engine(1)
            ctx.fp_view_group(6)
        for i in range(1, len(outs)):
            assert torch.equal(outs[0], outs[i]), i
        return outs[2]

    fp_all(dev(vol).to(torch.bfloat16))
    out = fp_all(dev(vol))
    err = rel_l2(host(out), O.fp(to_oracle(g), angles, offsets, vol))
    print("fp rel-L2", err)
    assert err <= TOL

    proj = np.random.default_rng(seed).random(g.proj_shape(len(angles)))
    p = dev(proj)
    try:   # BF16 working copies: the engines agree
        ctx.storage("bf16")
        b = []
        for e in (0, 1, 2):
            ctx.bp_engine(e)
            out = torch.empty(g.vol_shape, device="cuda")
            ctx.bp(g, p, out, fbp=True)
            b.append(out)
        assert torch.equal(b[0], b[1]) and torch.equal(b[0], b[2])
    finally:
        ctx.storage("f32")
        ctx.bp_engine(2)
    for fbp in (False, True):
        bouts = []
        try:
            for e in (0, 1, 2):
                ctx.bp_engine(e)
                out = torch.empty(g.vol_shape, device="cuda")
                ctx.bp(g, p, out, fbp=fbp)
                bouts.append(out)
        finally:
            ctx.bp_engine(2)
        assert torch.equal(bouts[0], bouts[1]) and torch.equal(bouts[0], bouts[2])
        err = rel_l2(host(bouts[2]), O.bp(to_oracle(g), angles, offsets, proj, fbp=fbp))
        print("bp rel-L2", fbp, err)
        assert err <= TOL


@pytest.mark.parametrize("seed", range(4))
def test_random_os_sart_parity(seed):
    """OS-SART on seeded random geometries (random subset counts, iterations
    and constraint; iterations replayed as HIP graphs) against the oracle
    within rel-L2 1e-4."""
    from gtt import operators as G
    r = np.random.default_rng(2000 + seed)
    # the solvers' geometry family (square slices, a detector 1.56 times the
    # volume's width, a full arc): random sizes, views, subsets and
    # iterations.  Short scans and narrow detectors leave voxels and rays
    # that OS-SART's normalisations amplify, testing conditioning, not parity
    from helpers import geom
    n = int(r.integers(10, 25))
    g = geom(n, n, int(r.integers(6, 25)), int(r.integers(20, 48)), int(r.integers(14, 40)))
    nviews = int(r.integers(8, 28))
    angles = angles_full(nviews, 360.0, start=float(r.uniform(0.0, 360.0)))
    offsets = None
    S = int(r.integers(1, 5))
    iters = int(r.integers(2, 4))
    nonneg = bool(r.random() < 0.5)
    print("CASE random os-sart", seed, g, len(angles), S, iters, nonneg, flush=True)
    vol = phantom_volume(g, seed=seed)
    p = O.fp(to_oracle(g), angles, offsets, vol).astype(np.float32)
    G.set_graphs(True)
    try:
        x = G.os_sart(g, angles, offsets, p, S, iters, beta=0.9, nonneg=nonneg)
    finally:
        G.set_graphs(False)
    op = O.CBSubsetOperator(S)
    o = np.zeros((len(angles), 2), np.float32) if offsets is None else offsets
    pb = op.setup(to_oracle(g), angles, o, p, O.Binning.default_3d(len(angles)))
    ref = O.os_sart(op, pb, iters, beta=0.9, nonneg=nonneg)
    err = rel_l2(x, ref)
    print("os-sart rel-L2", err)
    assert err <= 1e-4
