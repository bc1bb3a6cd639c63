"""GPU parity of the reference's operator classes through their C entry
points, against the FP64 oracle's restatements:

* cuConebeamProjectionOperator (xray/cuConebeamProjectionOperator.cpp):
  mult_M / mult_MH over bins, the offset-corrected mult_MH (full half-fan
  weights on a copy of the projections, :82-116), calculate_mask (the
  space-carving kernel, xray/conebeam_projection.cu:822-921) and the masked
  mult_MH (apply_mask, :1651-1657);
* hoCuConebeamProjectionOperator (xray/hoCuConebeamProjectionOperator.cpp):
  mult_M with the square-root half-fan weights on its output (:168-169),
  mult_MH in FBP mode (cosine, Parker, ramp, half-fan weights, FDK-weighted
  BP, image *= 2/N, :179-266) and in plain mode (square-root weights).

Bounds: rel-L2 <= 1e-5 (single passes, FP32 vs FP64).  Space-carving masks
agree exactly wherever the oracle's nearest-texel choice is unambiguous
(more than 1e-3 texel from a texel edge in every view).
"""
import numpy as np
import pytest

import oracle as O
from gtt import capi
from gtt import operators as ops
from helpers import angles_full, geom, phantom_volume, rel_l2, to_oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TOL = 1e-5


def print_case(*a):
    print("CASE", *a, flush=True)


def half_fan(nviews=40, arc=360.0, start=10.0, frac=0.2, nx=24, nu=48, nv=32):
    g = geom(nx, nx, nx, nu, nv)
    angles = angles_full(nviews, arc, start=start)
    offs = np.zeros((nviews, 2), np.float32)
    offs[:, 0] = np.float32(frac * g.fov_u)
    offs[:, 1] = np.linspace(-0.4, 0.4, nviews).astype(np.float32)
    return g, angles, offs


def shifted(angles):
    a = np.asarray(angles, np.float32)
    return (a - np.min(a)).astype(np.float32)


def dev(a, dtype=torch.float32):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda().to(dtype)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("bins", [None, "two"])
def test_cu_op_offset_corrected_mult_M_and_mult_MH(bins):
    g, angles, offs = half_fan()
    b = None if bins is None else [list(range(0, 40, 2)), list(range(1, 40, 2))]
    print_case("cu op half-fan", bins, g)
    op = ops.CuConebeamProjectionOperator().setup(g, angles, offs, bins=b)
    assert op.flags() == (False, True)         # 20% offset: correction on
    ob = O.Binning(b) if b else O.Binning.default_3d(len(angles))
    oop = O.CuConebeamOperator(to_oracle(g), shifted(angles), offs, ob)
    assert oop.use_offset_correction
    nb = len(ob)
    vol = np.stack([phantom_volume(g, seed=3 + k) for k in range(nb)]) if nb > 1 else phantom_volume(g)
    nviews = sum(len(x) for x in ob.bins)
    proj = torch.zeros((nviews, g.nv, g.nu), device="cuda")
    op.mult_M(dev(vol), proj)
    ref = np.zeros((nviews, g.nv, g.nu))
    oop.mult_M(vol, ref)
    e_fp = rel_l2(host(proj), ref)
    y = np.random.default_rng(5).random((nviews, g.nv, g.nu))
    out = torch.zeros(vol.shape, device="cuda")
    op.mult_MH(dev(y), out)
    refb = np.zeros(vol.shape)
    oop.mult_MH(y, refb)
    e_bp = rel_l2(host(out), refb)
    # the weights do act: without them the result differs
    plain = np.zeros(vol.shape)
    O.CuConebeamOperator(to_oracle(g), shifted(angles), offs, ob,
                         use_offset_correction=False).mult_MH(y, plain)
    print("cu op half-fan fp", e_fp, "bp", e_bp, "vs unweighted", rel_l2(plain, refb))
    assert e_fp <= TOL and e_bp <= TOL
    assert rel_l2(plain, refb) > 1e-2


def test_cu_op_calculate_mask_and_masked_mult_MH():
    g = geom(20, 20, 20, 40, 32)
    angles = angles_full(24, 360.0)
    print_case("cu op mask", g)
    vol = phantom_volume(g)
    p = O.fp(to_oracle(g), angles, None, vol)
    p[p < 0.3 * p.max()] = 0.0                # a silhouette with holes
    limit = 0.6 * p.max()
    op = ops.CuConebeamProjectionOperator().setup(g, angles, transform_angles=False)
    mask = torch.zeros(g.vol_shape, dtype=torch.uint8, device="cuda")
    pd = dev(p)
    op.calculate_mask(pd, limit, mask)
    ref, amb = O.spacecarver(to_oracle(g), angles, None, p, limit)
    got = host(mask)
    ok = amb == 0
    print("mask: voxels", got.size, "unambiguous", int(ok.sum()), "ones", int(got.sum()),
          "mismatches", int((got != ref)[ok].sum()))
    assert ok.mean() > 0.8 and 0 < ref.sum() < ref.size
    assert np.array_equal(got[ok], ref[ok])
    # the kernel entry point directly
    ctx = capi.Context(0)
    ctx.set_views(angles)
    m2 = torch.zeros_like(mask)
    ctx.spacecarver(g, pd, m2, limit)
    assert torch.equal(m2, mask)
    # masked backprojection = BP * mask
    op.set_mask(mask)
    y = np.random.default_rng(1).random(p.shape)
    out = torch.zeros(g.vol_shape, device="cuda")
    op.mult_MH(dev(y), out)
    refb = O.apply_mask(O.bp(to_oracle(g), angles, None, y), got)
    assert rel_l2(host(out), refb) <= TOL
    op.set_mask(None)
    op.mult_MH(dev(y), out)
    assert rel_l2(host(out), O.bp(to_oracle(g), angles, None, y)) <= TOL
    # cbct_apply_mask on its own
    x = torch.from_numpy(np.random.default_rng(2).random(g.vol_shape).astype(np.float32)).cuda()
    want = host(x) * (got != 0)
    ctx.apply_mask(x, mask)
    assert np.array_equal(host(x), want.astype(np.float32))
    ctx.close()


def test_ho_op_mult_M_sqrt_offset_weights_and_accumulate():
    g, angles, offs = half_fan()
    print_case("ho op mult_M half-fan", g)
    op = ops.HoCuConebeamProjectionOperator().setup(g, angles, offs)
    assert op.flags() == (False, True)
    vol = phantom_volume(g).astype(np.float32)
    got = np.zeros((len(angles), g.nv, g.nu), np.float32)
    op.mult_M(vol, got)
    b = O.Binning.default_3d(len(angles))
    ref = O.ho_mult_M(to_oracle(g), shifted(angles), offs, vol, b, use_offset_correction=True)
    unweighted = O.ho_mult_M(to_oracle(g), shifted(angles), offs, vol, b)
    e = rel_l2(got, ref)
    print("ho mult_M rel-L2", e, "vs unweighted", rel_l2(unweighted, ref))
    assert e <= TOL and rel_l2(unweighted, ref) > 1e-2
    acc = np.full_like(got, 0.25)
    op.mult_M(vol, acc, accumulate=True)
    assert rel_l2(acc, ref + 0.25) <= TOL


@pytest.mark.parametrize("case", ["full", "shortscan_offset", "two_bins"])
def test_ho_op_mult_MH_fbp(case):
    arc = 180.0 if case == "shortscan_offset" else 360.0   # 175.5 deg span: short scan
    frac = 0.2 if case == "shortscan_offset" else 0.0
    g, angles, offs = half_fan(nviews=40, arc=arc, frac=frac)
    bins = [list(range(0, 40, 2)), list(range(1, 40, 3))] if case == "two_bins" else None
    print_case("ho op FBP", case, g)
    op = ops.HoCuConebeamProjectionOperator().setup(g, angles, offs, bins=bins, use_fbp=True)
    ss, oc = op.flags()
    assert ss == (case == "shortscan_offset") and oc == (frac > 0)
    p = np.random.default_rng(9).random((40, g.nv, g.nu)).astype(np.float32)
    nb = len(bins) if bins else 1
    out = np.zeros(((nb,) if nb > 1 else ()) + g.vol_shape, np.float32)
    op.mult_MH(p, out)
    ob = O.Binning(bins) if bins else O.Binning.default_3d(40)
    ref = O.fdk_mult_MH(to_oracle(g), shifted(angles), offs, p, ob, O.guo_filter(g.nu).astype(np.float32))
    e = rel_l2(out, ref)
    print("ho FBP rel-L2", case, e)
    assert e <= TOL
    # the 2/N scale: twice the views' FDK sum over N (.cpp:253)
    if case == "full":
        raw = O.bp(to_oracle(g), shifted(angles), offs,
                   O.fdk_prefilter(to_oracle(g), shifted(angles), offs, p.astype(np.float64),
                                   O.guo_filter(g.nu).astype(np.float32)), fbp=True)
        assert np.allclose(ref, raw * np.float32(2.0 / 40), rtol=1e-6)
    # FBP-mode mult_M leaves the half-fan weights off its output
    proj = np.zeros_like(p)
    op.mult_M(np.ascontiguousarray(out), proj)
    refp = O.ho_mult_M(to_oracle(g), shifted(angles), offs, out, ob, use_offset_correction=oc,
                       use_fbp=True)
    assert rel_l2(proj, refp) <= TOL


def test_ho_op_mult_MH_plain_sqrt_weights():
    g, angles, offs = half_fan()
    print_case("ho op plain mult_MH half-fan", g)
    op = ops.HoCuConebeamProjectionOperator().setup(g, angles, offs)
    p = np.random.default_rng(4).random((len(angles), g.nv, g.nu)).astype(np.float32)
    out = np.zeros(g.vol_shape, np.float32)
    op.mult_MH(p, out)
    ref = O.ho_mult_MH(to_oracle(g), shifted(angles), offs, p, O.Binning.default_3d(len(angles)))
    e = rel_l2(out, ref)
    print("ho plain mult_MH rel-L2", e)
    assert e <= TOL
    acc = np.full_like(out, 1.5)
    op.mult_MH(p, acc, accumulate=True)
    assert rel_l2(acc, ref + 1.5) <= TOL


@pytest.mark.parametrize("case", ["plain", "half_fan_two_bins"])
def test_cbct_forwards_projection_program(case):
    """CBCT_forwards_projection (xray/CBCT_forwards_projection.cpp:60-254):
    angles start + i*spacing in float arithmetic, zero-filled projections,
    hoCu mult_M with the automatic offset correction (square-root weights on
    every view for a half-fan detector); views in no bin stay zero."""
    g = geom(20, 20, 20, 40, 30)
    nviews, start, step = 30, np.float32(7.5), np.float32(11.25)
    offs, bins = None, None
    if case != "plain":
        offs = np.zeros((nviews, 2), np.float32)
        offs[:, 0] = np.float32(0.25 * g.fov_u)
        bins = [list(range(0, 24, 2)), list(range(1, 24, 2))]     # views 24.. in no bin
    print_case("CBCT_forwards_projection", case)
    nb = len(bins) if bins else 1
    vol = np.stack([phantom_volume(g, seed=k + 2) for k in range(nb)]) if nb > 1 else phantom_volume(g)
    proj, angles = ops.cbct_forwards_projection(g, vol, start, step, nviews, offs, bins)
    want_angles = np.array([start + np.float32(i) * step for i in range(nviews)], np.float32)
    assert np.array_equal(angles, want_angles)
    ob = O.Binning(bins) if bins else O.Binning.default_3d(nviews)
    o = np.zeros((nviews, 2), np.float32) if offs is None else offs
    ref = O.ho_mult_M(to_oracle(g), shifted(angles), o, vol, ob,
                      use_offset_correction=O._auto_offset(to_oracle(g), o))
    e = rel_l2(proj, ref)
    print("CBCT_forwards_projection rel-L2", case, e)
    assert e <= TOL
    if bins:
        assert not np.any(proj[24:])
