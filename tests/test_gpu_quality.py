"""GPU parity of the multi-scale SSIM kernel (cbct_mssim) against the FP64
oracle's restatement of xray/mssim.cpp:14-121.

Bound: |MS-SSIM_gpu - MS-SSIM_oracle| <= 2e-6 (FP32 moments of
midpoint-shifted data; the result is a mean of values in [-1, 1])."""
import numpy as np
import pytest
from scipy import ndimage

import oracle as O
from gtt import capi
from gtt import quality as Q

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TOL = 2e-6


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def pair(shape, noise=0.05, offset=0.0, seed=3):
    r = np.random.default_rng(seed)
    Y = ndimage.gaussian_filter(r.random(shape), 1.2) + offset
    X = Y + noise * r.standard_normal(shape)
    return X.astype(np.float32), Y.astype(np.float32)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


@pytest.mark.parametrize("shape,sigma,scales", [
    ((16, 20, 24), (1.5, 1.5, 1.5), 3),
    ((9, 30, 17), (0.8, 1.3, 2.1), 2),
    ((5, 12, 40), (1.5, 1.5, 1.5), 3),       # z: radius 21 at the top scale, 5 voxels
    ((24, 24, 24), (1.0, 1.0, 0.0), 1),      # no filtering along z
])
def test_mssim_matches_oracle(ctx, shape, sigma, scales):
    X, Y = pair(shape)
    print("CASE mssim", shape, sigma, scales, flush=True)
    got = ctx.mssim(dev(X), dev(Y), sigma, scales)
    ref = O.mssim(X.astype(np.float64), Y.astype(np.float64), sigma, scales)
    print("mssim", got, ref, got - ref)
    assert abs(got - ref) <= TOL
    assert 0.0 < ref < 1.0


def test_mssim_offset_data_and_identity(ctx):
    """Data far from zero (values near 1000, range about 1): the midpoint
    shift keeps the FP32 moments accurate.  An image against itself is 1."""
    X, Y = pair((12, 16, 20), noise=0.02, offset=1000.0, seed=4)
    got = ctx.mssim(dev(X), dev(Y))
    ref = O.mssim(X.astype(np.float64), Y.astype(np.float64))
    print("mssim offset", got, ref)
    assert abs(got - ref) <= 1e-5
    assert ctx.mssim(dev(Y), dev(Y)) == pytest.approx(1.0, abs=1e-6)


def test_mssim_larger_volume(ctx):
    X, Y = pair((48, 64, 80), seed=5)
    print("CASE mssim 48x64x80", flush=True)
    got = Q.mssim(X, Y, ctx=ctx)
    ref = O.mssim(X.astype(np.float64), Y.astype(np.float64))
    assert abs(got - ref) <= TOL


def test_compare_images_averages_phases(ctx):
    X1, Y1 = pair((10, 14, 18), seed=6)
    X2, Y2 = pair((10, 14, 18), noise=0.1, seed=7)
    avg, per = Q.compare_images(np.stack([X1, X2]), np.stack([Y1, Y2]), ctx=ctx)
    refs = [O.mssim(X1.astype(np.float64), Y1.astype(np.float64)),
            O.mssim(X2.astype(np.float64), Y2.astype(np.float64))]
    assert np.allclose(per, refs, atol=TOL)
    assert avg == pytest.approx(float(np.float32((np.float32(per[0]) + np.float32(per[1])) / 2)), abs=1e-7)


def test_mssim_rejects_bad_arguments(ctx):
    X, Y = pair((8, 8, 8))
    with pytest.raises(ValueError):
        ctx.mssim(dev(X), dev(Y[:-1]))
    with pytest.raises(capi.CbctError, match="sigma"):
        ctx.mssim(dev(X), dev(Y), (-1.0, 1.0, 1.0))
    with pytest.raises(capi.CbctError, match="scales"):
        ctx.mssim(dev(X), dev(Y), 1.5, 0)
