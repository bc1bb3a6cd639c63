"""GPU parity of the gfx950 kernels against the FP64 oracle, through the C ABI.

Tolerances (FP32 arithmetic vs FP64 oracle): projectors and filter rel-L2
<= 1e-5; element-wise passes <= 1e-6; FP64 reductions <= 1e-12.
BF16-storage paths are compared with the oracle run on the bf16-rounded
input (same tolerance: the arithmetic is FP32), and their end-to-end error
against the FP32-input oracle is bounded separately (<= 1e-3, reported).
"""
import sys

import numpy as np
import pytest

import oracle as O
from gtt import capi
from helpers import angles_full, bf16_round, geom, phantom_volume, rel_l2, to_oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = 1e-5


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()
    return t if dtype is None else t.to(dtype)


def host(t):
    torch.cuda.synchronize()
    return t.float().cpu().numpy().astype(np.float64)


CASES = {
    # name: (geometry, nviews, offsets?)
    "cube32": (dict(nx=32, ny=32, nz=32, nu=48, nv=40), 24, False),
    "ragged": (dict(nx=23, ny=19, nz=17, nu=37, nv=29, vol_mm=(23.0, 19.0, 25.5)), 13, True),
    "flat": (dict(nx=40, ny=36, nz=6, nu=64, nv=12), 18, False),
    # ~6 detector texels per voxel: half of the 8-deep blocks' detector
    # rectangles (and none of the 16-deep ones) fit the pair kernel's LDS
    # tile, so its staged and direct paths alternate between views
    "zoom": (dict(nx=20, ny=20, nz=16, nu=100, nv=80), 24, True),
}


def make_case(name):
    gk, nviews, offs = CASES[name]
    g = geom(**gk)
    angles = angles_full(nviews, 360.0, start=3.0)
    offsets = None
    if offs:
        rng = np.random.default_rng(5)
        offsets = np.stack([rng.uniform(-2, 2, nviews), rng.uniform(-1, 1, nviews)], -1).astype(np.float32)
    return g, angles, offsets


def print_case(*a):
    print("CASE", *a, flush=True)


@pytest.mark.parametrize("name", list(CASES))
def test_fp_parity(ctx, name):
    g, angles, offsets = make_case(name)
    print_case("fp", name, g)
    vol = phantom_volume(g) + 0.001 * np.random.default_rng(1).random((g.nz, g.ny, g.nx))
    ref = O.fp(to_oracle(g), angles, offsets, vol)
    ctx.set_views(angles, offsets)
    out = torch.zeros(g.proj_shape(len(angles)), device="cuda")
    ctx.fp(g, dev(vol), out)
    err = rel_l2(host(out), ref)
    print("fp rel-L2", name, err)
    assert err <= TOL


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
@pytest.mark.parametrize("name", list(CASES) + ["tall", "wide", "deep"])
def test_fp_engines_bitwise_identical(ctx, name, dtype):
    """The detector-column engines (1-4) and the per-ray engine give
    identical bits.  'tall' has most rays leaving through the top/bottom
    faces (per-ray fallback); 'wide' has detector columns that miss the box
    entirely; 'deep' has a 600-voxel z range: with 12 rows per lane one
    row group's z window spans 5 chunks of 128 z, more than the cached
    engines hold (they fall back to per-step loads), with the default tile
    it spans 3."""
    tiles = (8,)
    if name == "tall":
        g, angles, offsets = geom(20, 20, 5, 30, 150, fov=(40.0, 60.0)), angles_full(7), None
    elif name == "wide":
        g, angles, offsets = geom(12, 12, 12, 40, 70, fov=(400.0, 30.0)), angles_full(5), None
    elif name == "deep":
        g, angles, offsets = geom(8, 8, 600, 6, 720, fov=(20.0, 900.0)), angles_full(3), None
        tiles = (8, 12)
    else:
        g, angles, offsets = make_case(name)
    print_case("engines", name, dtype, g)
    vol = phantom_volume(g) + 0.01
    tdt = torch.float32 if dtype == "f32" else torch.bfloat16
    ctx.set_views(angles, offsets)
    outs = []
    for tile in tiles:
        ctx.fp_tile(tile)
        for e in (0, 1, 2, 3, 4):   # per ray; the column-sharing variants
            ctx.fp_engine(e)
            out = torch.zeros(g.proj_shape(len(angles)), device="cuda")
            ctx.fp(g, dev(vol, tdt), out)
            torch.cuda.synchronize()
            outs.append(out)
    ctx.fp_engine(1)
    ctx.fp_tile(8)
    for e in range(1, len(outs)):
        assert torch.equal(outs[0], outs[e]), e
    if dtype == "f32":
        assert rel_l2(host(outs[1]), O.fp(to_oracle(g), angles, offsets, vol)) <= TOL


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_fp_view_groups_bitwise_identical(ctx, dtype):
    """The column engine's launch order (adjacent views of one column
    launched together, cbct_fp_view_group) changes no bit: 13 views (a
    ragged last group for every group size but 1 and 13, and a group larger
    than the view count), two row groups per column, a view offset."""
    g = geom(16, 16, 12, 24, 150)
    angles = angles_full(13)
    print_case("view groups", dtype, g)
    vol = phantom_volume(g) + 0.01
    tdt = torch.float32 if dtype == "f32" else torch.bfloat16
    ctx.set_views(angles, None)
    ctx.fp_tile(2)   # 150 rows at 2 rows per lane: two row groups
    ref = {}
    try:
        for e in (1, 2, 3):
            for vg in (1, 2, 5, 6, 13, 64):
                ctx.fp_engine(e)
                ctx.fp_view_group(vg)
                for view0, nv in ((0, 13), (2, 11)):
                    out = torch.zeros(g.proj_shape(nv), device="cuda")
                    ctx.fp(g, dev(vol, tdt), out, view0=view0, nviews=nv)
                    torch.cuda.synchronize()
                    key = (view0, nv)
                    if key not in ref:
                        ref[key] = out
                    assert torch.equal(out, ref[key]), (e, vg, view0)
    finally:
        ctx.fp_engine(1)
        ctx.fp_tile(8)
        ctx.fp_view_group(6)
    assert torch.equal(ref[(2, 11)], ref[(0, 13)][2:])
    for bad in (0, 65):
        with pytest.raises(capi.CbctError, match="views outside"):
            ctx.fp_view_group(bad)


def test_fp_split_launches(ctx):
    """A call whose column-engine grid exceeds one launch's 2^32 - 1
    work-items (an 8192 x 1 detector over 8200 views: 6.7e7 one-wave blocks)
    goes out as launches over consecutive views, with the per-ray engine's
    bits."""
    g = geom(8, 8, 8, 8192, 1)
    angles = angles_full(8200)
    print_case("split launches", g, len(angles))
    vol = phantom_volume(g) + 0.01
    ctx.set_views(angles, None)
    outs = []
    try:
        for e in (0, 1):
            ctx.fp_engine(e)
            out = torch.zeros(g.proj_shape(len(angles)), device="cuda")
            ctx.fp(g, dev(vol), out)
            outs.append(out)
    finally:
        ctx.fp_engine(1)
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1])
    assert float(outs[1].abs().sum()) > 0.0


@pytest.mark.parametrize("kz", [0, 16])
@pytest.mark.parametrize("fbp", [False, True])
@pytest.mark.parametrize("name", list(CASES))
def test_bp_parity(ctx, name, fbp, kz):
    """kz 0: z-voxels per thread chosen by size (8 at these sizes); 16: the
    instantiation large volumes use, forced."""
    g, angles, offsets = make_case(name)
    print_case("bp", name, fbp, kz, g)
    proj = np.random.default_rng(2).random(g.proj_shape(len(angles)))
    ref = O.bp(to_oracle(g), angles, offsets, proj, fbp=fbp)
    ctx.set_views(angles, offsets)
    out = torch.empty(g.vol_shape, device="cuda")
    ctx.bp_kz(kz)
    try:
        ctx.bp(g, dev(proj), out, fbp=fbp)
    finally:
        ctx.bp_kz(0)
    err = rel_l2(host(out), ref)
    print("bp rel-L2", name, fbp, kz, err)
    assert err <= TOL


def test_kernel_names(ctx):
    """The kernel-name queries follow the engine settings (all views by
    default): the staged backprojector where its tile holds the rectangles
    (cube32), quads where it does not (zoom at 16 z-voxels per thread) or
    when engine 1 is set, the staged kernel whenever engine 3 is set."""
    g, angles, offsets = make_case("cube32")
    ctx.set_views(angles, offsets)
    assert ctx.fp_kernel(g) in ("k_fp_ray", "k_fp_col", "k_fp_colc")
    assert ctx.bp_kernel(g) == "k_bpp"
    gz, az, oz = make_case("zoom")
    ctx.set_views(az, oz)
    try:
        ctx.bp_kz(16)
        assert ctx.bp_kernel(gz) == "k_bp"
        ctx.bp_engine(3)
        assert ctx.bp_kernel(gz) == "k_bpp"
        ctx.bp_engine(1)
        assert ctx.bp_kernel(gz, 0, 4) == "k_bp"
    finally:
        ctx.bp_engine(2)
        ctx.bp_kz(0)


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
@pytest.mark.parametrize("name", list(CASES))
def test_bp_engines_bitwise_identical(ctx, name, dtype):
    """Quad-texel and plain padded projection copies, and the LDS-staged
    kernels (engine 3 forces the BF16 pair kernel where engine 2 would pick
    quads), give identical bits at 8 and at 16 z-voxels per thread (FBP and
    plain, FP32 and BF16 storage, accumulating into a volume).  The two
    z-depths differ in rounding: each thread steps v from its first voxel."""
    g, angles, offsets = make_case(name)
    print_case("bp engines", name, dtype, g)
    proj = np.random.default_rng(4).random(g.proj_shape(len(angles)))
    tdt = torch.float32 if dtype == "f32" else torch.bfloat16
    ctx.set_views(angles, offsets)
    start = torch.from_numpy(np.random.default_rng(6).random(g.vol_shape).astype(np.float32)).cuda()
    try:
        for fbp in (False, True):
            for kz in (8, 16):
                ctx.bp_kz(kz)
                outs = []
                for e in (0, 1, 2, 3):
                    ctx.bp_engine(e)
                    out = start.clone()
                    ctx.bp(g, dev(proj, tdt), out, fbp=fbp, accumulate=True)
                    torch.cuda.synchronize()
                    outs.append(out)
                for i in range(1, len(outs)):
                    assert torch.equal(outs[0], outs[i]), (fbp, kz, i)
    finally:
        ctx.bp_engine(2)
        ctx.bp_kz(0)


@pytest.mark.parametrize("name", list(CASES))
def test_bf16_working_copies(ctx, name):
    """FP32 arrays with BF16 working copies (cbct_storage) give the same bits
    as BF16 arrays rounded by the caller, for FP and (F)BP, and stay within the
    BF16 bound of the FP32 result."""
    g, angles, offsets = make_case(name)
    print_case("bf16 working copies", name, g)
    vol = phantom_volume(g) + 0.01
    proj = np.random.default_rng(12).random(g.proj_shape(len(angles)))
    ctx.set_views(angles, offsets)
    outs = {}
    for mode in ("explicit", "storage", "f32"):
        ctx.storage("bf16" if mode == "storage" else "f32")
        v = dev(vol, torch.bfloat16) if mode == "explicit" else dev(vol)
        p = dev(proj, torch.bfloat16) if mode == "explicit" else dev(proj)
        fp = torch.zeros(g.proj_shape(len(angles)), device="cuda")
        ctx.fp(g, v, fp)
        bps = []
        for fbp in (False, True):
            b = torch.empty(g.vol_shape, device="cuda")
            ctx.bp(g, p, b, fbp=fbp)
            bps.append(b)
        torch.cuda.synchronize()
        outs[mode] = (fp, bps)
    ctx.storage("f32")
    assert torch.equal(outs["explicit"][0], outs["storage"][0])
    for a, b in zip(outs["explicit"][1], outs["storage"][1]):
        assert torch.equal(a, b)
    e_fp = rel_l2(host(outs["storage"][0]), host(outs["f32"][0]))
    e_bp = rel_l2(host(outs["storage"][1][0]), host(outs["f32"][1][0]))
    print("bf16 working copies rel-L2 vs f32: fp", e_fp, "bp", e_bp)
    assert e_fp <= 3e-3 and e_bp <= 3e-3


def test_fp_bp_accumulate_and_view_ranges(ctx):
    g, angles, offsets = make_case("cube32")
    print_case("accumulate", g)
    vol = phantom_volume(g)
    og = to_oracle(g)
    ctx.set_views(angles, offsets)
    # FP of views [5, 15) accumulated onto a non-zero buffer
    base = np.random.default_rng(3).random(g.proj_shape(10))
    out = dev(base)
    ctx.fp(g, dev(vol), out, view0=5, nviews=10, accumulate=True)
    ref = base + O.fp(og, angles[5:15], None, vol)
    assert rel_l2(host(out), ref) <= TOL
    # BP of views [3, 9) accumulated
    proj = np.random.default_rng(4).random(g.proj_shape(6))
    vb = np.random.default_rng(5).random(g.vol_shape)
    v = dev(vb)
    ctx.bp(g, dev(proj), v, view0=3, nviews=6, accumulate=True)
    ref = vb + O.bp(og, angles[3:9], None, proj)
    assert rel_l2(host(v), ref) <= TOL


def test_fp_bf16_storage(ctx):
    g, angles, offsets = make_case("cube32")
    print_case("fp-bf16", g)
    vol = phantom_volume(g)
    volb = bf16_round(vol).astype(np.float64)
    ctx.set_views(angles, offsets)
    out = torch.zeros(g.proj_shape(len(angles)), device="cuda")
    ctx.fp(g, dev(vol, torch.bfloat16), out)
    got = host(out)
    assert rel_l2(got, O.fp(to_oracle(g), angles, None, volb)) <= TOL
    e = rel_l2(got, O.fp(to_oracle(g), angles, None, vol))
    print("fp bf16-storage rel-L2 vs fp32-input oracle", e)
    assert e <= 1e-3


@pytest.mark.parametrize("fbp", [False, True])
def test_bp_bf16_storage(ctx, fbp):
    g, angles, offsets = make_case("ragged")
    print_case("bp-bf16", fbp, g)
    proj = np.random.default_rng(6).random(g.proj_shape(len(angles)))
    projb = bf16_round(proj).astype(np.float64)
    ctx.set_views(angles, offsets)
    out = torch.empty(g.vol_shape, device="cuda")
    ctx.bp(g, dev(proj, torch.bfloat16), out, fbp=fbp)
    got = host(out)
    assert rel_l2(got, O.bp(to_oracle(g), angles, offsets, projb, fbp=fbp)) <= TOL
    e = rel_l2(got, O.bp(to_oracle(g), angles, offsets, proj, fbp=fbp))
    print("bp bf16-storage rel-L2 vs fp32-input oracle", e)
    assert e <= 1e-3


def test_fp_transpose_parity_and_adjoint(ctx):
    g, angles, offsets = make_case("ragged")
    print_case("fpT", g)
    rng = np.random.default_rng(7)
    x = rng.random(g.vol_shape)
    y = rng.random(g.proj_shape(len(angles)))
    ctx.set_views(angles, offsets)
    xd, yd = dev(x), dev(y)
    fx = torch.zeros_like(yd)
    ctx.fp(g, xd, fx)
    fty = torch.zeros_like(xd)
    ctx.fp_transpose(g, yd, fty)
    assert rel_l2(host(fty), O.fp_transpose(to_oracle(g), angles, offsets, y)) <= TOL
    lhs = ctx.dot(fx, yd)
    rhs = ctx.dot(xd, fty)
    print("adjoint", lhs, rhs, abs(lhs - rhs) / abs(lhs))
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)


FILTER_ENGINES = [0, 1, 2]   # 0: FP32 VALU Toeplitz, 1: FP16 hi/lo split MFMA, 2: FFT


@pytest.mark.parametrize("engine", FILTER_ENGINES)
@pytest.mark.parametrize("nu", [48, 37, 256, 2052])
@pytest.mark.parametrize("which", ["ramlak", "guo"])
def test_ramp_filter_parity(ctx, nu, which, engine):
    print_case("filter", nu, which, "engine", engine)
    F = O.ramlak_filter(nu) if which == "ramlak" else O.guo_filter(nu)
    rows = np.random.default_rng(nu).standard_normal((70, nu))
    ref = O.filter_rows(rows, F.astype(np.float32))
    ctx.set_filter(F)
    ctx.filter_engine(engine)
    out = torch.empty((70, nu), device="cuda")
    ctx.ramp_filter(nu, dev(rows), out)
    err = rel_l2(host(out), ref)
    print("filter rel-L2", nu, which, "engine", engine, err)
    assert err <= TOL


@pytest.mark.parametrize("nu", [256, 1024])
def test_mfma_filter_row_paths_bitwise(ctx, nu):
    """The MFMA engine's row split keeps rows of up to 2048 elements in
    registers when the input is 16-byte aligned and reads them twice
    otherwise: a 4-byte-offset copy of the same rows must give the same
    bits."""
    print_case("filter row paths", nu)
    rows = np.random.default_rng(nu + 1).standard_normal((70, nu)).astype(np.float32)
    ctx.set_filter(O.ramlak_filter(nu))
    ctx.filter_engine(1)
    base = torch.empty(70 * nu + 4, device="cuda")
    aligned = base[4:].view(70, nu)
    shifted = base[1:70 * nu + 1].view(70, nu)
    assert aligned.data_ptr() % 16 == 0 and shifted.data_ptr() % 16 == 4
    aligned.copy_(torch.from_numpy(rows))
    out_a = torch.empty((70, nu), device="cuda")
    ctx.ramp_filter(nu, aligned, out_a)
    shifted.copy_(torch.from_numpy(rows))
    out_s = torch.empty((70, nu), device="cuda")
    ctx.ramp_filter(nu, shifted, out_s)
    torch.cuda.synchronize()
    assert torch.equal(out_a, out_s)


@pytest.mark.parametrize("engine", FILTER_ENGINES)
def test_ramp_filter_dynamic_range(ctx, engine):
    """Rows spanning 1e-6..1e8 (beyond the fp16 range), an all-zero row and a
    single spike: each row must meet the tolerance on its own, which the
    per-row power-of-two scaling of the split engine has to provide."""
    nu, nr = 1024, 130          # 130 rows: one full 128-row tile plus a ragged one
    print_case("filter dynamic range", nu, nr, "engine", engine)
    rng = np.random.default_rng(11)
    rows = rng.standard_normal((nr, nu)) * (10.0 ** rng.uniform(-6, 8, (nr, 1)))
    rows[5] = 0.0
    rows[7] = 0.0
    rows[7, 300] = 3e7
    F = O.guo_filter(nu)
    ref = O.filter_rows(rows, F.astype(np.float32))
    ctx.set_filter(F)
    ctx.filter_engine(engine)
    out = torch.empty((nr, nu), device="cuda")
    ctx.ramp_filter(nu, dev(rows), out)
    got = host(out)
    assert np.all(got[5] == 0.0)
    worst = max(rel_l2(got[r], ref[r]) for r in range(nr) if r != 5)
    print("filter dynamic-range worst row rel-L2", engine, worst)
    assert worst <= TOL


@pytest.mark.parametrize("nu,nrows", [(1, 3), (2, 1), (3, 5), (33, 65), (1000, 7), (4096, 3),
                                      (4097, 3)])
def test_fft_filter_edges(ctx, nu, nrows):
    """The FFT engine at the edges of its transform lengths: the smallest
    (N = 64 for nu <= 32, 32 row pairs per workgroup), odd row counts (a
    pair whose second row is missing), a length that is no power of two,
    the largest (nu = 4096, N = 8192) and the first row length it hands to
    the MFMA engine (4097); out of place and in place."""
    print_case("fft filter edges", nu, nrows)
    F = O.ramlak_filter(nu)
    rows = np.random.default_rng(7 * nu + nrows).standard_normal((nrows, nu))
    ref = O.filter_rows(rows, F.astype(np.float32))
    ctx.set_filter(F)
    ctx.filter_engine(2)
    x = dev(rows)
    out = torch.empty_like(x)
    ctx.ramp_filter(nu, x, out)
    err = rel_l2(host(out), ref)
    ctx.ramp_filter(nu, x, x)
    err_in_place = rel_l2(host(x), ref)
    print("fft filter edges rel-L2", nu, nrows, err, err_in_place)
    assert err <= TOL and err_in_place <= TOL


def test_fft_filter_extreme_rows(ctx):
    """Rows whose sums of squares leave the FP32 range (1e20: squares
    overflow; 1e-25: squares underflow) take the FFT engine's FP64 norm
    path, here at nu = 1024 where a row pair spans two waves; each row must
    still meet the tolerance and a zero row stay zero."""
    nu = 1024
    print_case("fft filter extreme rows", nu)
    rng = np.random.default_rng(5)
    mags = [1e20, 1e-25, 1.0, 0.0, 3e-25, 1e20, 1e-3]
    rows = rng.standard_normal((len(mags), nu)) * np.array(mags)[:, None]
    F = O.guo_filter(nu)
    ref = O.filter_rows(rows, F.astype(np.float32))
    ctx.set_filter(F)
    ctx.filter_engine(2)
    out = torch.empty((len(mags), nu), device="cuda")
    ctx.ramp_filter(nu, dev(rows), out)
    got = host(out)
    assert np.all(got[3] == 0.0)
    errs = [rel_l2(got[r], ref[r]) for r in range(len(mags)) if mags[r] != 0.0]
    print("fft filter extreme rows rel-L2", errs)
    assert max(errs) <= TOL


@pytest.mark.parametrize("nu", [300, 1024, 2052])
def test_fft_prefilter_short_scan_wide(ctx, nu):
    """The FFT engine's short-scan (Parker) instantiation on the 16-point
    transforms (N = 1024, 2048 and 8192: one, two and eight waves per row
    pair), cosine and Parker weights fused into its loads, against the
    oracle's pre-filter chain; in place as well.  At nu = 1024 all three
    engines measure 3.6e-6: the FP32 Parker weights (shared by the engines)
    against the oracle's, at the last view's steep edge; the same rows
    filtered without them agree to 2e-7."""
    g = geom(8, 8, 8, nu, 6, fov=(0.4 * nu, 2.4))
    n = 9
    angles = angles_full(n, 220.0)
    print_case("fft prefilter short scan", nu)
    proj = np.random.default_rng(nu).random(g.proj_shape(n))
    F = O.guo_filter(nu)
    ref = O.fdk_prefilter(to_oracle(g), angles, np.zeros((n, 2), np.float32), proj,
                          F.astype(np.float32), short_scan=True, offset_correct=False)
    ctx.set_views(angles)
    ctx.set_filter(F)
    ctx.filter_engine(2)
    pin = dev(proj)
    out = torch.empty_like(pin)
    ctx.fdk_prefilter(g, pin, out, short_scan=True)
    err = rel_l2(host(out), ref)
    ctx.fdk_prefilter(g, pin, pin, short_scan=True)
    err_in_place = rel_l2(host(pin), ref)
    print("fft prefilter short scan rel-L2", nu, err, err_in_place)
    assert err <= TOL and err_in_place <= TOL


@pytest.mark.parametrize("engine", FILTER_ENGINES)
@pytest.mark.parametrize("short_scan,offset", [(False, False), (True, False), (False, True)])
def test_fdk_prefilter_parity(ctx, short_scan, offset, engine):
    g = geom(16, 16, 16, 40, 24, fov=(60.0, 36.0))
    n = 20
    arc = 200.0 if short_scan else 360.0
    angles = angles_full(n, arc)
    offsets = np.zeros((n, 2), np.float32)
    if offset:
        offsets[:, 0] = 9.0
    print_case("prefilter", short_scan, offset, "engine", engine)
    proj = np.random.default_rng(8).random(g.proj_shape(n))
    F = O.ramlak_filter(g.nu)
    ref = O.fdk_prefilter(to_oracle(g), angles, offsets, proj, F.astype(np.float32),
                          short_scan=short_scan, offset_correct=offset)
    ctx.set_views(angles, offsets)
    ctx.set_filter(F)
    ctx.filter_engine(engine)
    pin = dev(proj)
    out = torch.empty_like(pin)
    ctx.fdk_prefilter(g, pin, out, short_scan=short_scan, offset_correct=offset)
    err = rel_l2(host(out), ref)
    print("prefilter rel-L2", err)
    assert err <= TOL
    # in place
    ctx.fdk_prefilter(g, pin, pin, short_scan=short_scan, offset_correct=offset)
    assert rel_l2(host(pin), ref) <= TOL


@pytest.mark.parametrize("sqrt_variant", [False, True])
def test_offset_and_parker_weights(ctx, sqrt_variant):
    g = geom(8, 8, 8, 40, 6, fov=(40.0, 6.0))
    angles = angles_full(10, 220.0)
    offsets = np.stack([np.linspace(4, 8, 10), np.zeros(10)], -1).astype(np.float32)
    print_case("weights", sqrt_variant)
    ctx.set_views(angles, offsets)
    p = np.random.default_rng(9).random(g.proj_shape(10))
    d = dev(p)
    ctx.offset_weights(g, d, sqrt_variant=sqrt_variant)
    ref = O.offset_weights(to_oracle(g), offsets, p.copy(), sqrt_variant=sqrt_variant)
    assert rel_l2(host(d), ref) <= 1e-6
    d = dev(p)
    ctx.parker_weights(g, d)
    ref = O.parker_weights(to_oracle(g), angles, p.copy())
    assert rel_l2(host(d), ref) <= 1e-5


@pytest.mark.parametrize("shape", [(6, 7, 9), (2, 5, 6, 8)])
def test_tv_parity_and_fused_update(ctx, shape):
    print_case("tv", shape)
    rng = np.random.default_rng(10)
    x = rng.random(shape)
    out0 = rng.random(shape)
    sigma, alpha, w = 3.0, 0.0, 1.0
    ref = out0.copy()
    O.tv_primal_dual(x, ref, sigma, alpha, w, accumulate=True)
    xd = dev(x)
    od = dev(out0)
    ctx.tv_primal_dual(xd, od, sigma, alpha, w, accumulate=True)
    assert rel_l2(host(od), ref) <= 1e-6
    # fused TV + updateG equals clear + TV + updateG
    s, P = rng.random(shape), rng.random(shape) + 0.5
    tau, lam = 0.05, 1e-3
    g = torch.zeros_like(xd)
    ctx.tv_primal_dual(xd, g, sigma, alpha, w, accumulate=True)
    x_unf = xd.clone()
    ctx.update_g(g, x_unf, dev(s), dev(P), tau, lam)
    x_f = torch.empty_like(xd)
    ctx.tv_update_g(xd, x_f, dev(s), dev(P), sigma, alpha, w, tau, lam)
    assert torch.equal(x_unf, x_f)
    gref = np.zeros(shape)
    O.tv_primal_dual(x, gref, sigma, alpha, w)
    assert rel_l2(host(x_f), O.update_g(gref, x, s, P, tau, lam)) <= 1e-6


def test_elementwise_and_reductions(ctx):
    print_case("elementwise")
    rng = np.random.default_rng(11)
    n = 100003
    a, b, c = rng.random(n) + 0.1, rng.random(n) + 0.1, rng.random(n)
    A, B, C = dev(a), dev(b), dev(c)
    ctx.ew("mul", A, B)
    ctx.ew("axpy", -2.0, C, B)
    ctx.ew("clamp_min", B, 0.05)
    ctx.ew("reciprocal", B)
    exp = 1.0 / np.maximum(b * a - 2.0 * c, 0.05)
    assert rel_l2(host(B), exp) <= 1e-6
    x = rng.standard_normal(n)
    X = dev(x)
    xf = x.astype(np.float32).astype(np.float64)
    assert ctx.asum(X) == pytest.approx(np.abs(xf).sum(), rel=1e-12)
    assert ctx.nrm2(X) == pytest.approx(np.sqrt((xf * xf).sum()), rel=1e-12)
    yf = b.astype(np.float32).astype(np.float64)
    assert ctx.dot(X, dev(b)) == pytest.approx(np.dot(xf, yf), rel=1e-12)
    # fused solver passes
    r, bb, w = rng.random(n), rng.random(n), rng.random(n)
    R = dev(r)
    ctx.sart_residual(R, dev(bb), dev(w))
    assert rel_l2(host(R), -(r - bb) * w) <= 1e-6
    xs, img, wx = rng.standard_normal(n), rng.random(n), rng.random(n)
    XS = dev(xs)
    ctx.sart_update(XS, dev(img), dev(wx), 0.7, True)
    assert rel_l2(host(XS), np.maximum(xs + 0.7 * img * wx, 0.0)) <= 1e-6
    z, xo = rng.random(n), rng.random(n)
    xx = rng.standard_normal(n).astype(np.float32).astype(np.float64)
    XX, Z = dev(xx), dev(z)
    ctx.mom_momentum(XX, dev(xo), Z, 0.4, True)
    xc = np.maximum(xx, 0)
    assert rel_l2(host(Z), xc * 1.4 - 0.4 * xo) <= 1e-6
    assert rel_l2(host(XX), xc) == 0.0


def test_bf16_conversion(ctx):
    x = np.random.default_rng(12).standard_normal(5000).astype(np.float32)
    X = dev(x)
    H = torch.empty(5000, dtype=torch.bfloat16, device="cuda")
    capi.load().cbct_to_bf16(ctx.h, 5000, capi._ptr(X), capi._ptr(H), capi._stream())
    assert np.array_equal(host(H).astype(np.float32), bf16_round(x))
    assert torch.equal(H, X.to(torch.bfloat16))


def test_errors_are_reported(ctx):
    g = geom(8, 8, 8, 10, 10)
    ctx.set_views(angles_full(4))
    out = torch.zeros((4, 10, 10), device="cuda")
    vol = torch.zeros((8, 8, 8), device="cuda")
    with pytest.raises(capi.CbctError, match="inconsistent sizes"):
        ctx.fp(g, vol, out, view0=2, nviews=4)
    bad = geom(8, 8, 8, 10, 10)
    bad.nx = 0
    with pytest.raises(capi.CbctError, match="dimensions"):
        ctx.fp(bad, vol, out)
    c2 = capi.Context(0)
    with pytest.raises(capi.CbctError, match="setup not performed"):
        c2.bp(g, out, vol)
    with pytest.raises(capi.CbctError, match="filter must be provided"):
        c2.set_views(angles_full(4))
        c2.fdk_prefilter(g, out, out)
    c2.close()


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_fp_autotune_keeps_bits(ctx, dtype):
    """The autotuner (cbct_fp_autotune) times column-engine variants and tiles
    and records the fastest for the geometry; projections afterwards are
    bitwise those of the built-in choice."""
    g, angles, offsets = make_case("ragged")
    print_case("autotune", dtype, g)
    capi.fp_autotune_clear()
    vol = phantom_volume(g) + 0.01
    tdt = torch.float32 if dtype == "f32" else torch.bfloat16
    ctx.set_views(angles, offsets)
    before = torch.zeros(g.proj_shape(len(angles)), device="cuda")
    ctx.fp(g, dev(vol, tdt), before)
    ms = ctx.fp_autotune(g, 0, len(angles), dtype)
    assert ms > 0.0
    assert ctx.fp_kernel(g, 0, len(angles), dtype) in ("k_fp_col", "k_fp_colc")
    after = torch.zeros_like(before)
    ctx.fp(g, dev(vol, tdt), after)
    torch.cuda.synchronize()
    assert torch.equal(before, after)
    capi.fp_autotune_clear()
