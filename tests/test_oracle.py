"""Pins the FP64 oracle (CPU only).

The reference ships no tests, fixtures or golden vectors for this path, so
each oracle function is pinned here by something independent of its own
code: an independent numpy/scipy restatement, a closed form, a textbook
special case or a mathematical invariant, chosen so that a plausible slip
(a sign, a transposed axis, a dropped half-pixel) fails.
"""
import math

import numpy as np
import pytest
from scipy import ndimage

import oracle as O


def small_geom(nx=10, ny=None, nz=None, nu=14, nv=12, l=None, fu=None, fv=None,
               sad=200.0, sdd=300.0):
    ny = ny or nx
    nz = nz or nx
    l = l or (float(nx), float(ny), float(nz))
    fu = fu or 1.6 * l[0] * 1.5
    fv = fv or 1.6 * l[2] * 1.5 * nv / nu * 1.2
    return O.Geom(nx, ny, nz, l[0], l[1], l[2], nu, nv, fu, fv, sad, sdd)


def rand_vol(g, seed=0):
    r = np.random.default_rng(seed)
    return r.random(g.vol_shape)


# ---------------------------------------------------------------------------
# Independent numpy/scipy restatement of the forward projector
# ---------------------------------------------------------------------------

def np_fp(g: O.Geom, angles, offsets, vol, spp=1.5):
    """Vectorised restatement: sample points per ray, then
    scipy.ndimage.map_coordinates(order=1, mode='grid-constant') which has
    the texture unit's border semantics (zero outside, linear inside)."""
    gc = g.c()
    L = np.array([gc.lx, gc.ly, gc.lz])
    N = np.array([g.nx, g.ny, g.nz], float)
    n = int(np.float32(spp) * np.float32(g.nx))
    out = np.zeros((len(angles), g.nv, g.nu))
    jj, ii = np.meshgrid(np.arange(g.nv), np.arange(g.nu), indexing="ij")
    for k, ang in enumerate(angles):
        th = float(np.float32(ang)) * math.pi / 180.0
        c, s = math.cos(th), math.sin(th)
        S = np.array([s * gc.sad, -c * gc.sad, 0.0])
        ou, ov = (offsets[k] if offsets is not None else (0.0, 0.0))
        pu = ((ii + 0.5) / g.nu - 0.5) * gc.fu + float(np.float32(ou))
        pv = ((jj + 0.5) / g.nv - 0.5) * gc.fv + float(np.float32(ov))
        add = gc.sdd - gc.sad
        E = np.stack([c * pu - s * add, s * pu + c * add, pv], -1)
        d = E - S
        with np.errstate(divide="ignore", invalid="ignore"):
            tA = (L / 2 - S) / d
            tB = (-L / 2 - S) / d
        lo = np.fmin(tA, tB)
        hi = np.fmax(tA, tB)
        a1 = np.maximum(np.nanmax(lo, -1), 0.0)
        a2 = np.minimum(np.nanmin(hi, -1), 1.0)
        step = np.linalg.norm((a2 - a1)[..., None] * d, axis=-1) / n
        m = np.arange(n)
        tpar = a1[..., None] + (a2 - a1)[..., None] * m / n       # (nv, nu, n)
        P = S + tpar[..., None] * d[..., None, :]                  # (nv, nu, n, 3) in mm
        tex = (P / L + 0.5) * N - 0.5
        coords = np.stack([tex[..., 2].ravel(), tex[..., 1].ravel(), tex[..., 0].ravel()])
        vals = ndimage.map_coordinates(vol, coords, order=1, mode="grid-constant", cval=0.0)
        out[k] = vals.reshape(g.nv, g.nu, n).sum(-1) * step
    return out


def np_bp(g: O.Geom, angles, offsets, proj, fbp=False):
    """Restatement of the voxel-driven backprojector with map_coordinates."""
    gc = g.c()
    z, y, x = np.meshgrid(np.arange(g.nz), np.arange(g.ny), np.arange(g.nx), indexing="ij")
    px = ((x + 0.5) / g.nx - 0.5) * gc.lx
    py = ((y + 0.5) / g.ny - 0.5) * gc.ly
    pz = ((z + 0.5) / g.nz - 0.5) * gc.lz
    out = np.zeros(g.vol_shape)
    for k, ang in enumerate(angles):
        th = float(np.float32(ang)) * math.pi / 180.0
        c, s = math.cos(th), math.sin(th)
        qx = c * px + s * py
        qy = -s * px + c * py
        mag = gc.sdd / (qy + gc.sad)
        ou, ov = (offsets[k] if offsets is not None else (0.0, 0.0))
        eu = qx * mag - float(np.float32(ou))
        ev = pz * mag - float(np.float32(ov))
        tu = (eu / gc.fu + 0.5) * g.nu - 0.5
        tv = (ev / gc.fv + 0.5) * g.nv - 0.5
        vals = ndimage.map_coordinates(proj[k], np.stack([tv.ravel(), tu.ravel()]), order=1,
                                       mode="grid-constant", cval=0.0).reshape(g.vol_shape)
        if fbp:
            U = 1.0 + (px * s - py * c) / gc.sad   # r sin(beta - phi) = x sin(beta) - y cos(beta)
            vals = vals / (U * U)
        out += vals
    return out


ANGLES = np.array([0.0, 17.0, 45.0, 90.0, 133.5, 200.0, 271.0, 333.0], np.float32)


@pytest.mark.parametrize("off", [False, True])
def test_fp_matches_scipy_restatement(off):
    g = small_geom(nx=10, ny=9, nz=8, nu=14, nv=11)
    offsets = None
    if off:
        offsets = np.stack([np.linspace(-3, 2, len(ANGLES)), np.linspace(1, -1, len(ANGLES))], -1)
    vol = rand_vol(g, 1)
    ref = np_fp(g, ANGLES, offsets, vol)
    got = O.fp(g, ANGLES, offsets, vol)
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("fbp", [False, True])
def test_bp_matches_scipy_restatement(fbp):
    g = small_geom(nx=9, ny=10, nz=7, nu=13, nv=12)
    offsets = np.stack([np.linspace(-2, 2, len(ANGLES)), np.linspace(0.5, -0.5, len(ANGLES))], -1)
    proj = np.random.default_rng(2).random(g.proj_shape(len(ANGLES)))
    ref = np_bp(g, ANGLES, offsets, proj, fbp)
    got = O.bp(g, ANGLES, offsets, proj, fbp=fbp)
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


def test_num_samples_truncates_float_product():
    # int(float(1.5) * float(nx)) as the reference passes it to an int parameter
    assert O.num_samples(1.5, 128) == 192
    assert O.num_samples(1.5, 7) == 10
    assert O.num_samples(0.7, 10) == 7   # 0.7f*10 = 7.0000005 -> 7


def test_fp_rot90_invariance():
    """Projecting at 90 degrees equals projecting the volume rotated by 90
    degrees about z at 0 degrees (square volume): pins the rotation sense."""
    g = small_geom(nx=8, nu=12, nv=10)
    vol = rand_vol(g, 3)
    a = O.fp(g, [90.0], None, vol)
    # FP_theta(V) = FP_0(V o R(theta)); R(90): (x, y) -> (-y, x), so the volume
    # seen at 0 deg is V'(x, y) = V(-y, x), and -y_j = x_{N-1-j}:
    # V'[z, j, i] = V[z, i, N-1-j]
    volr = np.transpose(vol, (0, 2, 1))[:, ::-1, :]
    b = O.fp(g, [0.0], None, np.ascontiguousarray(volr))
    assert np.abs(a - b).max() <= 1e-10 * np.abs(a).max()
    # and the opposite rotation sense does not match
    volw = np.transpose(vol, (0, 2, 1))[:, :, ::-1]
    c = O.fp(g, [0.0], None, np.ascontiguousarray(volw))
    assert np.abs(a - c).max() > 1e-3 * np.abs(a).max()


def test_fp_zflip_is_detector_vflip():
    g = small_geom(nx=8, nu=12, nv=10)
    vol = rand_vol(g, 4)
    a = O.fp(g, ANGLES[:4], None, vol)
    b = O.fp(g, ANGLES[:4], None, np.ascontiguousarray(vol[::-1]))
    assert np.abs(a[:, ::-1, :] - b).max() <= 1e-11 * np.abs(a).max()


def test_fp_constant_volume_is_chord_length():
    """A constant-one volume projects to the chord length through the box, up
    to the half-voxel border ramp of the zero-bordered interpolation."""
    g = small_geom(nx=12, nu=16, nv=12)
    vol = np.ones(g.vol_shape)
    got = O.fp(g, [0.0, 30.0], None, vol)
    voxel = g.lx / g.nx
    # chord by the same slab clipping, computed here independently in mm
    gc = g.c()
    for k, ang in enumerate([0.0, 30.0]):
        th = ang * math.pi / 180
        c, s = math.cos(th), math.sin(th)
        S = np.array([s * gc.sad, -c * gc.sad, 0.0])
        for j in range(g.nv):
            for i in range(g.nu):
                pu = ((i + 0.5) / g.nu - 0.5) * gc.fu
                pv = ((j + 0.5) / g.nv - 0.5) * gc.fv
                E = np.array([c * pu - s * (gc.sdd - gc.sad), s * pu + c * (gc.sdd - gc.sad), pv])
                d = E - S
                L = np.array([gc.lx, gc.ly, gc.lz])
                with np.errstate(divide="ignore"):
                    t1, t2 = (L / 2 - S) / d, (-L / 2 - S) / d
                a1 = max(np.max(np.minimum(t1, t2)), 0.0)
                a2 = min(np.min(np.maximum(t1, t2)), 1.0)
                chord = max(a2 - a1, 0.0) * np.linalg.norm(d)
                if chord > 3 * voxel:
                    assert abs(got[k, j, i] - chord) <= 1.5 * voxel + 0.05 * chord


def test_fp_linear_and_accumulate():
    g = small_geom(nx=7, nu=9, nv=8)
    a, b = rand_vol(g, 5), rand_vol(g, 6)
    pa = O.fp(g, ANGLES[:3], None, a)
    pb = O.fp(g, ANGLES[:3], None, b)
    pab = O.fp(g, ANGLES[:3], None, 2.0 * a - 3.0 * b)
    assert np.allclose(pab, 2 * pa - 3 * pb, rtol=1e-12, atol=1e-12)
    acc = pa.copy()
    O.fp(g, ANGLES[:3], None, b, out=acc, accumulate=True)
    assert np.allclose(acc, pa + pb, rtol=1e-14, atol=1e-14)


def test_fp_ray_missing_box_is_not_guarded():
    """A ray that misses the box keeps a positive step and samples outside the
    volume: the result is zero for a volume with a zero border."""
    g = small_geom(nx=6, nu=9, nv=5, fu=450.0, fv=10.0)
    vol = np.zeros(g.vol_shape)
    vol[1:-1, 1:-1, 1:-1] = 1.0
    p = O.fp(g, [0.0], None, vol)
    assert np.all(p[0, :, 0] == 0.0)  # outermost column (pu = -200 mm) misses the box
    assert p[0, 2, 4] > 0             # central ray hits


def test_fp_rays_and_bp_voxels_are_pointwise_fp_and_bp():
    """The subsample forms used at full-size geometries give exactly the
    full projector's values at their rays / voxels (offsets and FBP too)."""
    g = small_geom(nx=9, ny=7, nz=8, nu=13, nv=11)
    angles = np.array([3.0, 47.0, 181.5, 300.0], np.float32)
    offs = np.array([[0.5, -0.3], [0.0, 0.0], [-1.1, 0.7], [2.0, 0.1]], np.float32)
    vol = rand_vol(g, 8)
    full = O.fp(g, angles, offs, vol)
    rng = np.random.default_rng(2)
    rays = np.stack([rng.integers(0, 4, 50), rng.integers(0, g.nu, 50), rng.integers(0, g.nv, 50)], -1)
    got = O.fp_rays(g, angles, offs, vol, rays)
    assert np.array_equal(got, full[rays[:, 0], rays[:, 2], rays[:, 1]])
    for fbp in (False, True):
        bfull = O.bp(g, angles, offs, full, fbp=fbp)
        vox = rng.choice(g.nx * g.ny * g.nz, 40, replace=False)
        assert np.array_equal(O.bp_voxels(g, angles, offs, full, vox, fbp=fbp), bfull.ravel()[vox])
    with pytest.raises(ValueError):
        O.fp_rays(g, angles, offs, vol, [[4, 0, 0]])
    with pytest.raises(ValueError):
        O.bp_voxels(g, angles, offs, full, [g.nx * g.ny * g.nz])


def test_adjoint_fp_transpose():
    g = small_geom(nx=7, ny=6, nz=5, nu=9, nv=8)
    rng = np.random.default_rng(7)
    x = rng.random(g.vol_shape)
    y = rng.random(g.proj_shape(4))
    offs = np.stack([np.linspace(-1, 1, 4), np.linspace(1, -1, 4)], -1)
    lhs = np.vdot(O.fp(g, ANGLES[:4], offs, x), y)
    rhs = np.vdot(x, O.fp_transpose(g, ANGLES[:4], offs, y))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize("fbp", [False, True])
def test_adjoint_bp_transpose(fbp):
    g = small_geom(nx=6, ny=7, nz=5, nu=9, nv=8)
    rng = np.random.default_rng(8)
    x = rng.random(g.vol_shape)
    y = rng.random(g.proj_shape(4))
    lhs = np.vdot(O.bp(g, ANGLES[:4], None, y, fbp=fbp), x)
    rhs = np.vdot(y, O.bp_transpose(g, ANGLES[:4], None, x, fbp=fbp))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_reference_pair_is_not_adjoint():
    """The reference (mult_M, mult_MH) pair is unmatched (different
    interpolation and a step-length scale only in FP)."""
    g = small_geom(nx=6, nu=9, nv=8)
    rng = np.random.default_rng(9)
    x = rng.random(g.vol_shape)
    y = rng.random(g.proj_shape(4))
    lhs = np.vdot(O.fp(g, ANGLES[:4], None, x), y)
    rhs = np.vdot(x, O.bp(g, ANGLES[:4], None, y))
    assert abs(lhs - rhs) > 1e-3 * abs(lhs)


def test_bp_constant_sinogram_known_answer():
    """A constant sinogram c backprojects to c*nviews wherever every view's
    projection of the voxel centre lies at least a pixel inside the detector."""
    g = small_geom(nx=8, nu=40, nv=40, fu=60.0, fv=60.0)
    angles = np.arange(0, 360, 15, dtype=np.float32)
    proj = np.full(g.proj_shape(len(angles)), 2.5)
    got = O.bp(g, angles, None, proj)
    assert np.allclose(got, 2.5 * len(angles), rtol=1e-13)


def test_bp_single_pixel_bilinear_weights():
    """One non-zero detector pixel: the voxel response equals the hand-computed
    bilinear weight of that pixel at the voxel's projected position."""
    g = small_geom(nx=4, nu=10, nv=10, fu=20.0, fv=20.0)
    proj = np.zeros(g.proj_shape(1))
    proj[0, 5, 4] = 1.0
    got = O.bp(g, [0.0], None, proj)
    gc = g.c()
    for z in range(4):
        for y in range(4):
            for x in range(4):
                px = ((x + 0.5) / 4 - 0.5) * gc.lx
                py = ((y + 0.5) / 4 - 0.5) * gc.ly
                pz = ((z + 0.5) / 4 - 0.5) * gc.lz
                mag = gc.sdd / (py + gc.sad)   # view 0: q = pos
                tu = (px * mag / gc.fu + 0.5) * 10 - 0.5
                tv = (pz * mag / gc.fv + 0.5) * 10 - 0.5
                wu = max(0.0, 1 - abs(tu - 4))
                wv = max(0.0, 1 - abs(tv - 5))
                assert got[z, y, x] == pytest.approx(wu * wv, abs=1e-13)


def test_fbp_weight_closed_form():
    """FBP of a constant-one sinogram with full detector coverage is
    sum_k 1/U_k^2 with U_k = 1 + (x sin b_k - y cos b_k)/SAD, while the
    reference's projection magnification uses SAD + q_y (mirrored)."""
    g = small_geom(nx=6, nu=40, nv=40, fu=60.0, fv=60.0)
    angles = np.array([0, 40, 95, 180, 250], np.float32)
    proj = np.ones(g.proj_shape(len(angles)))
    got = O.bp(g, angles, None, proj, fbp=True)
    gc = g.c()
    z, y, x = np.meshgrid(np.arange(6), np.arange(6), np.arange(6), indexing="ij")
    px = ((x + 0.5) / 6 - 0.5) * gc.lx
    py = ((y + 0.5) / 6 - 0.5) * gc.ly
    exp = np.zeros(g.vol_shape)
    for a in angles:
        b = float(a) * math.pi / 180
        U = 1.0 + (px * math.sin(b) - py * math.cos(b)) / gc.sad
        exp += 1.0 / (U * U)
    assert np.allclose(got, exp, rtol=1e-13)


# ---------------------------------------------------------------------------
# Filters and weights
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("nu", [8, 13, 64])
def test_filter_toeplitz_equals_fft_chain(nu):
    """The exact Toeplitz form equals the reference's pad -> R2C -> *F -> C2R
    -> /2Nu -> crop chain, evaluated with numpy.fft."""
    rng = np.random.default_rng(nu)
    rows = rng.standard_normal((5, nu))
    for F in (O.ramlak_filter(nu), O.guo_filter(nu), rng.random(nu + 1)):
        pad = np.zeros((5, 2 * nu))
        off = nu // 2
        pad[:, off:off + nu] = rows
        spec = np.fft.rfft(pad, axis=-1) * F
        back = np.fft.irfft(spec, n=2 * nu, axis=-1)  # irfft already divides by 2nu
        ref = back[:, off:off + nu]
        got = O.filter_rows(rows, F)
        assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def test_ramlak_kernel_closed_form():
    """For F[k] = k the spatial kernel has the closed form
    h[0] = Nu/4 + ..., checked against the direct cosine sum evaluated with
    numpy (independent of the C loop) and its known zero-mean property:
    sum_m h over the full period equals F[0]/... = 0."""
    nu = 32
    F = O.ramlak_filter(nu)
    h = O.filter_kernel(F)
    full = np.concatenate([h, [np.sum(F * np.cos(np.pi * np.arange(nu + 1) * nu / nu) * np.r_[1, 2 * np.ones(nu - 1), 1]) / (2 * nu)], h[1:][::-1]])
    # inverse DFT of the symmetric extension has mean F[0]/(2nu)... = 0 for Ram-Lak
    assert abs(full.sum()) < 1e-12
    # h[m] for odd m is negative, even m > 0 is ~0 (ramp kernel structure)
    assert all(h[m] < 0 for m in range(1, nu, 2))
    assert h[0] > 0


def test_guo_filter_normalisation():
    for nu in (16, 256):
        F = O.guo_filter(nu)
        assert F[0] == 0.0
        assert np.sum(np.abs(F)) == pytest.approx(nu + 1, rel=1e-12)
        k = np.arange(nu + 1)
        A2 = 0.2 * (nu + 1) ** 2
        raw = np.where(k * k < A2, k * A2 / (A2 - k * k) * np.exp(-A2 / np.maximum(A2 - k * k, 1e-300)), 0.0)
        assert np.allclose(F, raw * (nu + 1) / raw.sum(), rtol=1e-12)


def test_cosine_weights_closed_form():
    g = small_geom(nx=4, nu=12, nv=10, fu=40.0, fv=30.0)
    w = O.cosine_weights(g)
    gc = g.c()
    u = np.arange(12)
    v = np.arange(10)
    V, Uu = np.meshgrid(v, u, indexing="ij")
    s = gc.sad * ((Uu / 12) - 0.5) * gc.fu / gc.sdd     # sampled at u/Nu, not centres
    t = gc.sad * ((V / 10) - 0.5) * gc.fv / gc.sdd
    assert np.allclose(w, gc.sad / np.sqrt(gc.sad ** 2 + s ** 2 + t ** 2), rtol=1e-15)
    assert w[5, 6] == pytest.approx(1.0)   # u = Nu/2, v = Nv/2 is the centre


def test_parker_q1_is_conventional_parker():
    """With q = 1 the reference's formula reduces to Parker's weights:
    sin^2(pi/4 b/(d-a)) ramp-up, 1, sin^2(pi/4 (pi+2d-b)/(d+a)) ramp-down."""
    delta = 0.15
    for alpha in (-0.1, -0.03, 0.0, 0.05, 0.12):
        for beta in np.linspace(0.0, math.pi + 2 * delta, 41):
            got = O.parker_weight(beta, alpha, delta, q=1.0, eps=0.0)
            if beta <= 2 * (delta - alpha):
                exp = math.sin(math.pi / 4 * beta / (delta - alpha)) ** 2
            elif beta <= math.pi - 2 * alpha:
                exp = 1.0
            elif beta <= math.pi + 2 * delta:
                exp = math.sin(math.pi / 4 * (math.pi + 2 * delta - beta) / (delta + alpha)) ** 2
            else:
                exp = 0.0
            assert got == pytest.approx(exp, abs=1e-12), (alpha, beta)


def test_parker_redundancy_identity():
    """Conjugate rays (beta, alpha) and (beta + pi + 2 alpha, -alpha) carry
    weights summing to one (q = 1, eps = 0)."""
    delta = 0.2
    rng = np.random.default_rng(1)
    for _ in range(200):
        alpha = rng.uniform(-0.9 * delta, 0.9 * delta)
        beta = rng.uniform(0.0, math.pi - 2 * alpha)
        w1 = O.parker_weight(beta, alpha, delta, 1.0, 0.0)
        w2 = O.parker_weight(beta + math.pi + 2 * alpha, -alpha, delta, 1.0, 0.0)
        assert w1 + w2 == pytest.approx(1.0, abs=1e-12)


def test_offset_weights_redundancy():
    """Half-fan weights satisfy w(t) + w(-t) = 1 (full form) and
    w_sqrt^2 = w on the same sample positions; zero offset is a no-op."""
    g = small_geom(nx=4, nu=40, nv=2, fu=40.0, fv=4.0)
    off_u = 6.0    # 6 pixels of 1 mm
    offs = np.array([[off_u, 0.0]], np.float32)
    w = O.offset_weights(g, offs, np.ones(g.proj_shape(1)))[0, 0]
    u = np.arange(40)
    t = (40.0 * (u / 40 - 0.5) + off_u)
    omega = 20.0 - off_u
    for i in range(40):
        j = 40 - i - 2 * int(off_u)  # t_j = -t_i
        if 0 <= j < 40 and abs(t[i]) <= omega:
            assert w[i] + w[j] == pytest.approx(1.0, abs=1e-12)
    assert np.all(w[t < -omega] == 0.0)
    assert np.all(w[t > omega] == 1.0)
    # zero offset: sign is NaN and every weight test is false
    w0 = O.offset_weights(g, np.zeros((1, 2), np.float32), np.ones(g.proj_shape(1)))
    assert np.all(w0 == 1.0)


def test_offset_sqrt_squares_to_full_at_shifted_samples():
    g = small_geom(nx=4, nu=40, nv=1, fu=40.0, fv=4.0)
    offs = np.array([[-5.0, 0.0]], np.float32)
    ws = O.offset_weights(g, offs, np.ones(g.proj_shape(1)), sqrt_variant=True)[0, 0]
    u = np.arange(40)
    sign = -1.0
    t = (40.0 * ((u + 0.5) / 40 - 0.5) - 5.0) * sign
    omega = 20.0 - 5.0
    inside = np.abs(t) <= omega
    full = np.sin(np.pi * (t + omega) / (4 * omega)) ** 2
    assert np.allclose(ws[inside] ** 2, full[inside], rtol=1e-12)


# ---------------------------------------------------------------------------
# TV and update steps
# ---------------------------------------------------------------------------

def np_tv(x, sigma, alpha, w):
    """Independent numpy restatement with np.roll (periodic differences)."""
    g = np.stack([np.roll(x, -1, axis=2) - x, np.roll(x, -1, axis=1) - x,
                  np.roll(x, -1, axis=0) - x])       # (dx, dy, dz)
    g *= w * sigma / (1 + sigma * alpha)
    g /= np.maximum(1.0, np.sqrt((g ** 2).sum(0)))
    div = (np.roll(g[0], 1, axis=2) - g[0]) + (np.roll(g[1], 1, axis=1) - g[1]) + \
          (np.roll(g[2], 1, axis=0) - g[2])
    return w * div


@pytest.mark.parametrize("sigma", [0.3, 50.0])
def test_tv_matches_numpy_roll(sigma):
    x = np.random.default_rng(3).random((5, 6, 7))
    out = np.zeros_like(x)
    O.tv_primal_dual(x, out, sigma, 0.1, 0.7, accumulate=True)
    assert np.allclose(out, np_tv(x, sigma, 0.1, 0.7), rtol=1e-13, atol=1e-14)


def test_tv_hand_case_3cubed():
    """A single bright voxel in a 3^3 zero cube, small sigma (no projection):
    out = w^2 sigma * D^T D x = w^2 sigma * (-Laplacian) = w^2 sigma * (+6 at
    the voxel, -1 at its six neighbours)."""
    x = np.zeros((3, 3, 3))
    x[1, 1, 1] = 1.0
    out = np.zeros_like(x)
    sigma, w = 0.01, 2.0
    O.tv_primal_dual(x, out, sigma, 0.0, w, accumulate=False)
    exp = np.zeros_like(x)
    exp[1, 1, 1] = 6
    for d in [(0, 1, 1), (2, 1, 1), (1, 0, 1), (1, 2, 1), (1, 1, 0), (1, 1, 2)]:
        exp[d] = -1
    assert np.allclose(out, w * w * sigma * exp, atol=1e-15)


def test_update_g_special_cases():
    rng = np.random.default_rng(4)
    x, s, P, g = rng.random(10), rng.random(10), rng.random(10) + 0.5, rng.random(10)
    assert np.allclose(O.update_g(g, x, s, P, 0.0, 1.0), x)
    # g = 0, P = 1: x = (x + (tau/l) s) / (1 + tau/l)
    r = O.update_g(np.zeros(10), x, s, np.ones(10), 0.3, 0.5)
    assert np.allclose(r, (x + 0.6 * s) / 1.6)
    # stationary point: x = s with g = 0 is a fixed point for any P
    assert np.allclose(O.update_g(np.zeros(10), s, s, P, 0.7, 2.0), s)


# ---------------------------------------------------------------------------
# Subsets and solvers
# ---------------------------------------------------------------------------

def test_subset_split_round_robin():
    subsets, perm, _ = O.subset_split(36, O.Binning.default_3d(36), 10)
    for s in range(10):
        assert subsets[s] == [i for i in range(36) if i % 10 == s]
    assert perm == [i for s in range(10) for i in range(36) if i % 10 == s]
    # views missing from every bin are skipped, and the counter k skips them too
    b = O.Binning([[0, 2, 3, 5, 6]])
    subsets, _, _ = O.subset_split(7, b, 2)
    assert subsets == [[0, 3, 6], [2, 5]]


def test_subset_split_sorts_by_bin():
    b = O.Binning([[1, 3, 5], [0, 2, 4]])
    subsets, perm, _ = O.subset_split(6, b, 2)
    # round robin over present views 0..5: s0 = [0,2,4], s1 = [1,3,5]; then by bin
    assert subsets == [[0, 2, 4], [1, 3, 5]]
    b = O.Binning([[2, 3], [0, 1]])
    subsets, _, _ = O.subset_split(4, b, 1)
    assert subsets == [[2, 3, 0, 1]]


def tiny_problem(nx=6, nviews=6, nu=9, nv=8):
    g = small_geom(nx=nx, nu=nu, nv=nv)
    angles = np.arange(nviews, dtype=np.float32) * np.float32(360.0 / nviews)
    vol = np.zeros(g.vol_shape)
    vol[1:-1, 1:-1, 1:-1] = 0.02
    vol[2:4, 2:4, 2:4] += 0.01
    return g, angles, vol


def test_os_sart_one_subset_one_iteration_closed_form():
    g, angles, vol = tiny_problem()
    b = O.fp(g, angles, None, vol)
    op = O.CBSubsetOperator(1)
    bp_ = op.setup(g, angles, np.zeros((len(angles), 2)), b, O.Binning.default_3d(len(angles)))
    x = O.os_sart(op, bp_, 1, beta=0.8)
    ones_v = np.ones(g.vol_shape)
    Wp = 1.0 / np.maximum(O.fp(g, angles, None, ones_v), 1e-6)
    Wx = 1.0 / np.maximum(O.bp(g, angles, None, np.ones(b.shape)), 1e-6)
    exp = 0.8 * O.bp(g, angles, None, b * Wp) * Wx
    assert np.allclose(x, exp, rtol=1e-12, atol=1e-15)


def test_os_sart_decreases_residual():
    g, angles, vol = tiny_problem()
    b = O.fp(g, angles, None, vol)
    op = O.CBSubsetOperator(3)
    bp_ = op.setup(g, angles, np.zeros((len(angles), 2)), b, O.Binning.default_3d(len(angles)))
    obj = []
    O.os_sart(op, bp_, 4, nonneg=True, objective=obj)
    assert obj[-1] < obj[0] < np.vdot(b, b)


def test_os_mom_without_regulariser_is_pure_momentum():
    """No regulariser: denoise is x = z, so two subset steps from zero are
    z1 = -S P A1^T(A1 x0 - b1) (t1 = (1+sqrt5)/2, momentum factor 0),
    then x2 = z1 - S P A2^T(A2 z1 - b2) and z2 = x2 (1+c) - c z1, c = (t1-1)/t2."""
    g, angles, vol = tiny_problem(nviews=4)
    b = O.fp(g, angles, None, vol)
    op = O.CBSubsetOperator(2)
    perm_b = op.setup(g, angles, np.zeros((4, 2)), b, O.Binning.default_3d(4))
    z = O.os_mom(op, perm_b, 1)
    # direct evaluation
    ones = np.ones(g.vol_shape)
    Pm = 1.0 / np.maximum(O.bp(g, angles, None, O.fp(g, angles, None, ones)), 1e-6)
    a1, a2 = angles[[0, 2]], angles[[1, 3]]
    b1, b2 = b[[0, 2]], b[[1, 3]]
    t1 = 0.5 * (1 + math.sqrt(5.0))
    x1 = -2 * Pm * O.bp(g, a1, None, O.fp(g, a1, None, np.zeros(g.vol_shape)) - b1)
    z1 = x1 * (1 + 0.0 / t1) - 0.0 * np.zeros(g.vol_shape)
    t2 = 0.5 * (1 + math.sqrt(1 + 4 * t1 * t1))
    x2 = z1 - 2 * Pm * O.bp(g, a2, None, O.fp(g, a2, None, z1) - b2)
    c = (t1 - 1) / t2
    z2 = x2 * (1 + c) - c * x1
    assert np.allclose(z, z2, rtol=1e-12, atol=1e-16)


def test_os_mom_tv_reduces_noise_energy():
    g, angles, vol = tiny_problem(nviews=6)
    b = O.fp(g, angles, None, vol)
    b += np.random.default_rng(0).normal(0, 0.02 * b.max(), b.shape)
    op = O.CBSubsetOperator(2)
    pb = op.setup(g, angles, np.zeros((6, 2)), b, O.Binning.default_3d(6))
    z0 = O.os_mom(op, pb, 3, nonneg=True)
    z1 = O.os_mom(op, pb, 3, tv_weight=1e-3, tau0=1e-3, reg_steps=2, nonneg=True)
    tv = lambda v: np.abs(np.diff(v, axis=0)).sum() + np.abs(np.diff(v, axis=1)).sum() + np.abs(np.diff(v, axis=2)).sum()
    assert tv(z1) < tv(z0)


class _IdentityOp:
    """One-subset identity 'projector' (codomain = domain), so that the
    preconditioner is P = 1 and a solver step can be followed by hand."""
    nsubsets = 1

    def __init__(self, shape):
        self.shape = shape

    def domain_shape(self):
        return self.shape

    def projection_subsets(self, arr):
        return [arr]

    def mult_M(self, x, out, accumulate=False):
        out[...] = (out if accumulate else 0.0) + x

    mult_MH = mult_M

    def mult_M_subset(self, x, out, subset, accumulate=False):
        self.mult_M(x, out, accumulate)

    mult_MH_subset = mult_M_subset


def test_os_mom_tv_schedule_by_hand():
    """osMOMSolverD::denoise schedule (solvers/osMOMSolverD.h:205-227) with
    every constant written out: one subset, identity projector (P = 1),
    TV weight 0.5 (so avg_lambda = 0.5 and the normalised TV weight is 1),
    tau0 = 1e-3, reg_steps = 2, start x0.
      gamma~ = 0.35 / 0.5 / (N / asum(P)) = 0.7
      sigma0 = 1 / (tau0 * L^2), L = 4         = 62.5
      theta1 = 1 / sqrt(1 + 2 * 0.7 * 1e-3)    = 0.999300734143549
      tau1   = tau0 * theta1                   = 0.000999300734143549
      sigma1 = sigma0 / theta1                 = 62.543734698209384
    The gradient step takes z to b (S = beta = 1, A = I, P = 1); each TV
    step is x <- (x - tau g + (tau/0.5) b) / (1 + tau/0.5) with
    g = D^T proj(sigma D x) (updateG, :233-247); the momentum factor of the
    first subset is 0, so the solver returns x2.  Dual fields stay inside
    the unit ball, so sigma enters linearly: sigma0 = 1/(4 tau0), a theta of
    1, or gamma~ without the 0.35 all change the result by far more than
    the tolerance."""
    rng = np.random.default_rng(21)
    shape = (4, 4, 4)
    x0 = 0.004 * rng.random(shape)
    b = 0.004 * rng.random(shape)
    tau0, tau1 = 1e-3, 0.000999300734143549
    sigma0, sigma1 = 62.5, 62.543734698209384
    lam = 0.5

    def dual_norm_max(x, sigma):
        d = np.stack([np.roll(x, -1, axis=2) - x, np.roll(x, -1, axis=1) - x,
                      np.roll(x, -1, axis=0) - x]) * sigma
        return np.sqrt((d ** 2).sum(0)).max()

    assert dual_norm_max(x0, sigma0) < 1.0
    g1 = np_tv(x0, sigma0, 0.0, 1.0)
    x1 = (x0 - tau0 * g1 + (tau0 / lam) * b) / (1.0 + tau0 / lam)
    assert dual_norm_max(x1, sigma1) < 1.0
    g2 = np_tv(x1, sigma1, 0.0, 1.0)
    x2 = (x1 - tau1 * g2 + (tau1 / lam) * b) / (1.0 + tau1 / lam)

    z = O.os_mom(_IdentityOp(shape), b, 1, tv_weight=lam, tau0=tau0, reg_steps=2, x0=x0)
    assert np.allclose(z, x2, rtol=1e-13, atol=1e-17)
    # the schedule matters at this tolerance: a wrong sigma0 (1/(4 tau0))
    # moves the result by orders of magnitude more than it
    g1w = np_tv(x0, 1.0 / (4 * tau0), 0.0, 1.0)
    x1w = (x0 - tau0 * g1w + (tau0 / lam) * b) / (1.0 + tau0 / lam)
    assert np.abs(x1w - x1).max() > 1e3 * 1e-13 * np.abs(x1).max()


def test_os_mom_second_subset_momentum_with_tv():
    """Two iterations of one subset on the identity projector: the second
    subset step starts from z1 = x2 of the first (factor 0), then
    t2 = (1 + sqrt(1 + 4 t1^2)) / 2 and z = x' (1 + c) - c x_prev with
    c = (t1 - 1) / t2 (solvers/osMOMSolverD.h:139-181), with the TV schedule
    restarting at tau0 in every denoise call."""
    rng = np.random.default_rng(5)
    shape = (4, 4, 4)
    b = 0.004 * rng.random(shape)
    tau0, tau1 = 1e-3, 0.000999300734143549
    sigma0, sigma1 = 62.5, 62.543734698209384
    lam = 0.5

    def denoise(x, s):
        x = (x - tau0 * np_tv(x, sigma0, 0.0, 1.0) + (tau0 / lam) * s) / (1.0 + tau0 / lam)
        return (x - tau1 * np_tv(x, sigma1, 0.0, 1.0) + (tau1 / lam) * s) / (1.0 + tau1 / lam)

    t1 = 0.5 * (1 + math.sqrt(5.0))
    xa = denoise(np.zeros(shape), b)          # subset step 1 from x = z = 0: z <- b
    z1 = xa                                   # factor (1 - 1)/t1 = 0
    t2 = 0.5 * (1 + math.sqrt(1 + 4 * t1 * t1))
    s2 = z1 - (z1 - b)                        # gradient step on z = z1 (A = I): b
    xb = denoise(z1, s2)                      # x = z1 at the second step
    c = (t1 - 1) / t2
    z2 = xb * (1 + c) - c * xa
    z = O.os_mom(_IdentityOp(shape), b, 2, tv_weight=lam, tau0=tau0, reg_steps=2)
    assert np.allclose(z, z2, rtol=1e-13, atol=1e-17)


def test_fdk_scale_cancels_in_driver():
    """FDK_reconstruct_3d rescales by <Af,p>/<Af,Af>, so the absolute filter
    scale cancels."""
    g, angles, vol = tiny_problem(nviews=8, nu=12, nv=10)
    p = O.fp(g, angles, None, vol)
    F = O.ramlak_filter(g.nu)
    f1, s1 = O.fdk_reconstruct_3d(g, angles, None, p, F)
    f2, s2 = O.fdk_reconstruct_3d(g, angles, None, p, 7.0 * F)
    assert np.allclose(f1, f2, rtol=1e-12, atol=1e-15)
    assert s1 == pytest.approx(7.0 * s2, rel=1e-12)


# ---------------------------------------------------------------------------
# McKinnon-Bates 4-D FDK (xray/FDK_reconstruct_4d.cpp:100-170)
# ---------------------------------------------------------------------------

def mb_problem(nviews=12, arc=360.0):
    g = small_geom(nx=6, nu=10, nv=8)
    angles = (np.arange(nviews, dtype=np.float32) * np.float32(arc / nviews)).astype(np.float32)
    vol = rand_vol(g, 3)
    p = O.fp(g, angles, None, vol)
    # a view-dependent change, so that the bins see different data
    p *= (1.0 + 0.2 * np.sin(np.arange(nviews)))[:, None, None]
    return g, angles, p


@pytest.mark.parametrize("arc", [360.0, 200.0])
def test_fdk4d_scale_and_3d_image_are_the_3d_drivers(arc):
    """When the bins cover every view, A(expand f) is the 3-D projection of f:
    the 4-D scale and the union FBP are those of FDK_reconstruct_3d."""
    g, angles, p = mb_problem(arc=arc)
    bins = O.Binning([[0, 3, 6, 9], [1, 4, 7, 10, 2], [5, 8, 11]])
    F = O.guo_filter(g.nu)
    v4, f3, s4 = O.fdk_reconstruct_4d(g, angles, None, p, F, bins)
    v3, s3 = O.fdk_reconstruct_3d(g, angles, None, p, F)
    assert s4 == pytest.approx(s3, rel=1e-12)
    assert np.allclose(f3 * s3, v3, rtol=1e-12, atol=1e-14)
    assert v4.shape == (3,) + g.vol_shape


@pytest.mark.parametrize("arc", [360.0, 200.0])
def test_fdk4d_bin_weighted_sum_identity(arc):
    """With bins partitioning the views, sum_b n_b FBP_b(q) = N FBP_all(q)
    (the pre-filter acts on each view alone, FBP_b carries 2/n_b), so
    sum_b n_b x_b / N = s (2 f - s FBP_all(A f)).  Uneven bin sizes make a
    dropped +f, a wrong sign, a 2/N weight or a swapped bin fail.  The
    tolerance covers the float32 2/n factors the reference uses."""
    g, angles, p = mb_problem(arc=arc)
    bins = [[0, 3, 6, 9, 11], [1, 4, 7], [2, 5, 8, 10]]
    F = O.guo_filter(g.nu)
    v4, f3, s = O.fdk_reconstruct_4d(g, angles, None, p, F, O.Binning(bins))
    a0 = (angles - angles.min()).astype(np.float32)
    off = np.zeros((len(angles), 2), np.float32)
    Af = O.fp(g, a0, off, f3)
    fbp_Af = O.fdk_mult_MH(g, a0, off, Af, O.Binning.default_3d(len(angles)), F)
    lhs = sum(len(b) * v4[i] for i, b in enumerate(bins)) / len(angles)
    rhs = s * (2.0 * f3 - s * fbp_Af)
    assert np.abs(lhs - rhs).max() <= 1e-7 * np.abs(rhs).max()


def test_fdk4d_views_outside_the_bins_take_no_part():
    """A view in no bin changes neither the union FBP, the scale nor any bin."""
    g, angles, p = mb_problem()
    bins = O.Binning([[0, 2, 4, 6, 8, 10], [1, 3, 5, 9]])      # 7 and 11 in no bin
    F = O.guo_filter(g.nu)
    v1, f1, s1 = O.fdk_reconstruct_4d(g, angles, None, p, F, bins)
    q = p.copy()
    q[[7, 11]] = 5.0 * np.random.default_rng(1).random(q[[7, 11]].shape)
    v2, f2, s2 = O.fdk_reconstruct_4d(g, angles, None, q, F, bins)
    assert np.array_equal(v1, v2) and np.array_equal(f1, f2) and s1 == s2


def test_fdk4d_one_bin_of_every_view():
    """A single bin of all views: x = s (2 f - s FBP(A f)) exactly, and the
    result differs from the 3-D driver's s f by the second FBP pass."""
    g, angles, p = mb_problem()
    F = O.guo_filter(g.nu)
    v4, f3, s = O.fdk_reconstruct_4d(g, angles, None, p, F, O.Binning.default_3d(len(angles)))
    a0 = (angles - angles.min()).astype(np.float32)
    off = np.zeros((len(angles), 2), np.float32)
    fbp_Af = O.fdk_mult_MH(g, a0, off, O.fp(g, a0, off, f3), O.Binning.default_3d(len(angles)), F)
    assert np.allclose(v4[0], s * (2.0 * f3 - s * fbp_Af), rtol=1e-10, atol=1e-14)
    assert not np.allclose(v4[0], s * f3, rtol=1e-3)


# ---------------------------------------------------------------------------
# Multi-scale SSIM (xray/mssim.cpp:14-121)
# ---------------------------------------------------------------------------

def scipy_gauss(v, sigma_xyz):
    return ndimage.gaussian_filter(np.asarray(v, np.float64), sigma=tuple(sigma_xyz)[::-1],
                                   mode="reflect", truncate=3.5)


@pytest.mark.parametrize("sigma", [(1.5, 1.5, 1.5), (0.7, 2.0, 3.3), (0.0, 1.0, 6.0)])
def test_gauss_filter_is_scipys(sigma):
    """The oracle's separable filter is scipy.ndimage.gaussian_filter with
    truncate=3.5 and mode='reflect', also where the radius exceeds the axis
    (z: 5 voxels against 21 taps a side at sigma 6)."""
    v = np.random.default_rng(4).random((5, 9, 12))
    assert np.allclose(O.gauss_filter(v, sigma), scipy_gauss(v, sigma), rtol=0, atol=1e-13)
    # the 11-tap window of SSIM at sigma 1.5
    assert O.gauss_taps(1.5).size == 11 and abs(O.gauss_taps(1.5).sum() - 1.0) < 1e-15


def ssim_parts_scipy(X, Y, sigma, c1, c2):
    ux, uy = scipy_gauss(X, sigma), scipy_gauss(Y, sigma)
    vx = scipy_gauss(X * X, sigma) - ux ** 2
    vy = scipy_gauss(Y * Y, sigma) - uy ** 2
    cxy = scipy_gauss(X * Y, sigma) - ux * uy
    lum = (2 * ux * uy + c1) / (ux ** 2 + uy ** 2 + c1)
    cs = (2 * cxy + c2) / (vx + vy + c2)
    return lum, cs


def mssim_pair(seed=5, shape=(10, 12, 14)):
    r = np.random.default_rng(seed)
    Y = ndimage.gaussian_filter(r.random(shape), 1.0)
    X = Y + 0.05 * r.standard_normal(shape)
    return X, Y


def test_mssim_one_scale_is_textbook_ssim():
    """scales = 1 is the mean SSIM of Wang et al. (2004) with a Gaussian
    window, C1 = (0.01 R)^2, C2 = (0.03 R)^2, R the reference's range."""
    X, Y = mssim_pair()
    R = Y.max() - Y.min()
    lum, cs = ssim_parts_scipy(X, Y, (1.5,) * 3, (0.01 * R) ** 2, (0.03 * R) ** 2)
    assert O.mssim(X, Y, (1.5,) * 3, scales=1) == pytest.approx(float((lum * cs).mean()), rel=1e-6)


def test_mssim_scales_compose_cs_and_top_luminance():
    """scales = 3: CS at sigma, 2 sigma, 4 sigma and the luminance at 4 sigma."""
    X, Y = mssim_pair(shape=(12, 12, 16))
    R = Y.max() - Y.min()
    c1, c2 = (0.01 * R) ** 2, (0.03 * R) ** 2
    S = np.ones_like(X)
    for f in (1.0, 2.0, 4.0):
        lum, cs = ssim_parts_scipy(X, Y, (0.8 * f, 1.0 * f, 1.2 * f), c1, c2)
        S *= cs
    S *= lum
    assert O.mssim(X, Y, (0.8, 1.0, 1.2), scales=3) == pytest.approx(float(S.mean()), rel=1e-6)


def test_mssim_invariants():
    X, Y = mssim_pair()
    assert O.mssim(Y, Y) == pytest.approx(1.0, abs=1e-12)
    # C1, C2 follow the reference's range: a common positive scale cancels
    assert O.mssim(3.0 * X, 3.0 * Y) == pytest.approx(O.mssim(X, Y), rel=1e-6)
    # a worse image scores lower
    r = np.random.default_rng(9)
    assert O.mssim(Y + 0.2 * r.standard_normal(Y.shape), Y) < O.mssim(X, Y) < 1.0


# ---------------------------------------------------------------------------
# CBSubsetWeightOperator (xray/CBSubsetWeightOperator.h:11-63)
# ---------------------------------------------------------------------------

def weighted_pair(S=3, nviews=10, weights=None):
    g, angles, vol = tiny_problem(nviews=nviews)
    b = O.fp(g, angles, None, vol)
    off = np.zeros((nviews, 2))
    op = O.CBSubsetOperator(S)
    pb = op.setup(g, angles, off, b, O.Binning.default_3d(nviews))
    ow = O.CBSubsetWeightOperator(S)
    w = np.ones_like(b) if weights is None else weights(angles, b)
    pbw = ow.setup(g, angles, off, b, O.Binning.default_3d(nviews), w)
    return g, angles, vol, b, op, pb, ow, pbw


def test_weighted_subset_op_unit_weights_is_unweighted():
    g, angles, vol, b, op, pb, ow, pbw = weighted_pair()
    assert np.array_equal(pb, pbw)
    y0, y1 = np.zeros(pb.shape), np.zeros(pb.shape)
    op.mult_M(vol, y0)
    ow.mult_M(vol, y1)
    assert np.array_equal(y0, y1)
    x0, x1 = np.zeros(g.vol_shape), np.zeros(g.vol_shape)
    op.mult_MH(pb, x0)
    ow.mult_MH(pb, x1)
    assert np.array_equal(x0, x1)


def test_weighted_subset_op_weights_follow_their_views():
    """Weights 1 + angle/360, a function of each view's own angle: every
    weighted subset projection is the unweighted one times 1 + a/360 for the
    angle the subset's projector holds for that view, so the weights are
    permuted with the views (a transposed permutation fails)."""
    wfun = lambda a, b: np.broadcast_to((1.0 + a.astype(np.float64) / 360.0)[:, None, None], b.shape).copy()
    g, angles, vol, b, op, pb, ow, pbw = weighted_pair(weights=wfun)
    assert np.allclose(pbw, pb * (1.0 + op.angles.astype(np.float64) / 360.0)[:, None, None], rtol=1e-12)
    for s in range(3):
        nv = op.codomain[s][0]
        y0, y1 = np.zeros((nv, g.nv, g.nu)), np.zeros((nv, g.nv, g.nu))
        op.mult_M_subset(vol, y0, s)
        ow.mult_M_subset(vol, y1, s)
        sub_angles = op.ops[s].all_angles.astype(np.float64)
        assert np.allclose(y1, y0 * (1.0 + sub_angles / 360.0)[:, None, None], rtol=1e-12, atol=0)
        # accumulate adds the weighted projection
        y2 = y1.copy()
        ow.mult_M_subset(vol, y2, s, accumulate=True)
        assert np.allclose(y2, 2.0 * y1, rtol=1e-12)


def test_weighted_subset_op_zero_weights_drop_views():
    """Views weighted zero take no part in the backprojection."""
    drop = {2, 7}
    wfun = lambda a, b: np.stack([np.zeros(b.shape[1:]) if k in drop else np.ones(b.shape[1:])
                                  for k in range(b.shape[0])])
    g, angles, vol, b, op, pb, ow, pbw = weighted_pair(weights=wfun)
    y = np.random.default_rng(2).random(pb.shape)
    yz = y.copy()
    for i, p in enumerate(op.perm):
        if p in drop:
            yz[i] = 0.0
    x0, x1 = np.zeros(g.vol_shape), np.zeros(g.vol_shape)
    op.mult_MH(yz, x0)
    ow.mult_MH(y, x1)
    assert np.allclose(x0, x1, rtol=1e-12, atol=1e-15)


def test_weighted_solvers_uniform_weights_cancel():
    """A uniform weight c scales data, M and M^H alike; OS-SART's row/column
    normalisers and osMOMSolverD's preconditioner cancel it (no TV)."""
    wfun = lambda a, b: np.full(b.shape, 2.5)
    g, angles, vol, b, op, pb, ow, pbw = weighted_pair(weights=wfun)
    assert np.allclose(O.os_sart(ow, pbw, 2, beta=0.8), O.os_sart(op, pb, 2, beta=0.8), rtol=1e-9, atol=1e-12)
    assert np.allclose(O.os_mom(ow, pbw, 2), O.os_mom(op, pb, 2), rtol=1e-9, atol=1e-12)


# ---------------------------------------------------------------------------
# Space carving, masks and the host operator's backprojection
# ---------------------------------------------------------------------------

def np_spacecarver(g: O.Geom, angles, offsets, proj, limit):
    """Independent restatement of conebeam_spacecarver_kernel
    (xray/conebeam_projection.cu:822-921): z negated, nearest texel of the
    normalised coordinate, zero outside, count of views with 0 != value <
    limit, mask = count < 3."""
    gc = g.c()
    z, y, x = np.meshgrid(np.arange(g.nz), np.arange(g.ny), np.arange(g.nx), indexing="ij")
    px = ((x + 0.5) / g.nx - 0.5) * gc.lx
    py = ((y + 0.5) / g.ny - 0.5) * gc.ly
    pz = -((z + 0.5) / g.nz - 0.5) * gc.lz
    count = np.zeros(px.shape, int)
    for k, ang in enumerate(angles):
        th = math.radians(float(np.float32(ang)))
        qx = math.cos(th) * px + math.sin(th) * py
        qy = -math.sin(th) * px + math.cos(th) * py
        t = gc.sdd / (qy + gc.sad)
        ou, ov = (float(np.float32(offsets[k][0])), float(np.float32(offsets[k][1]))) \
            if offsets is not None else (0.0, 0.0)
        iu = np.floor(((qx * t - ou) / gc.fu + 0.5) * g.nu).astype(int)
        iv = np.floor(((pz * t - ov) / gc.fv + 0.5) * g.nv).astype(int)
        inside = (iu >= 0) & (iu < g.nu) & (iv >= 0) & (iv < g.nv)
        val = np.where(inside, proj[k][np.clip(iv, 0, g.nv - 1), np.clip(iu, 0, g.nu - 1)], 0.0)
        count += (val < limit) & (val != 0)
    return (count < 3).astype(np.uint8)


def test_spacecarver_matches_numpy_restatement():
    g = small_geom(nx=11, ny=9, nz=10, nu=17, nv=15)
    rng = np.random.default_rng(31)
    angles = np.sort(rng.uniform(0, 360, 7)).astype(np.float32)
    offs = rng.uniform(-1, 1, (7, 2)).astype(np.float32)
    proj = rng.random((7, g.nv, g.nu))
    proj[proj < 0.2] = 0.0
    mask, amb = O.spacecarver(g, angles, offs, proj, 0.6)
    ref = np_spacecarver(g, angles, offs, proj, 0.6)
    ok = amb == 0
    assert ok.mean() > 0.9
    assert np.array_equal(mask[ok], ref[ok])
    assert 0 < mask.sum() < mask.size


def test_spacecarver_known_answers():
    """A constant sinogram c on a detector that covers the volume: every view
    counts when c < limit (mask 0), none when c >= limit or c = 0 (mask 1).
    Rows of the lower detector half only: the z negation sends them to the
    upper half of the volume."""
    g = small_geom(nx=8, nu=40, nv=40)
    angles = np.arange(6, dtype=np.float32) * np.float32(60)
    ones = np.ones((6, g.nv, g.nu))
    assert O.spacecarver(g, angles, None, 0.5 * ones, 1.0)[0].sum() == 0
    assert O.spacecarver(g, angles, None, 0.5 * ones, 0.5)[0].min() == 1
    assert O.spacecarver(g, angles, None, 0.0 * ones, 1.0)[0].min() == 1
    low = 0.5 * ones
    low[:, g.nv // 2:, :] = 0.0          # only rows below the centre are nonzero
    m, amb = O.spacecarver(g, angles, None, low, 1.0)
    assert np.all(m[g.nz // 2:] == 0) and np.all(m[:g.nz // 2] == 1)


def test_apply_mask_multiplies():
    img = np.array([1.5, -2.0, np.inf, 3.0])
    m = np.array([1, 0, 1, 0], np.uint8)
    assert np.array_equal(O.apply_mask(img, m), [1.5, -0.0, np.inf, 0.0])


def test_ho_mult_MH_plain_bins_and_sqrt_offset_weights():
    """Non-FBP host backprojection: per bin the plain BP of the bin's views
    into the bin's volume, after the square-root half-fan weights when
    offset correction is on (conebeam_projection.cu:1522-1523)."""
    g = small_geom(nx=8, nu=14, nv=10)
    angles = np.arange(6, dtype=np.float32) * np.float32(60)
    offs = np.zeros((6, 2), np.float32)
    offs[:, 0] = 0.3 * g.fu                     # > 0.1 FOV: correction on by default
    proj = np.random.default_rng(3).random((6, g.nv, g.nu))
    b = O.Binning([[0, 2, 4], [1, 3, 5]])
    out = O.ho_mult_MH(g, angles, offs, proj, b)
    w = O.offset_weights(g, offs, np.ones_like(proj), sqrt_variant=True)
    assert not np.allclose(w, 1.0)
    for k, idx in enumerate(b.bins):
        ref = O.bp(g, angles[idx], offs[idx], proj[idx] * w[idx])
        assert np.allclose(out[k], ref, rtol=1e-13, atol=1e-15)
    plain = O.ho_mult_MH(g, angles, offs, proj, b, use_offset_correction=False)
    assert np.allclose(plain[1], O.bp(g, angles[[1, 3, 5]], offs[[1, 3, 5]], proj[[1, 3, 5]]))
