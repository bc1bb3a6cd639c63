"""FP64 CPU oracle for the cone-beam CT path (test infrastructure only).

Only the test suite, the smoke check and the CPU-baseline leg of the
benchmark may import this module: it is the checker, never the product.

The numerical kernels (forward projection, backprojection, filters, weights,
TV) live in ``cbct_oracle.c`` and run in double precision.  This module
restates the host-side logic of the reference in numpy float64:

* ``CBSubsetOperator`` setup (xray/CBSubsetOperator.cpp:52-166)
* ``cuConebeamProjectionOperator`` dispatch (xray/cuConebeamProjectionOperator.cpp:38-215)
* ``hoCuConebeamProjectionOperator`` FBP mode (xray/hoCuConebeamProjectionOperator.cpp:105-266)
* ``osSARTSolver::solve`` (solvers/osSARTSolver.h:33-126)
* ``osMOMSolverD::solve`` (solvers/osMOMSolverD.h:82-247, base ``updateG``)
* the FDK driver (xray/FDK_reconstruct_3d.cpp:128-170)
* the McKinnon-Bates 4-D FDK driver (xray/FDK_reconstruct_4d.cpp:100-170)
* ``CBSubsetWeightOperator`` (xray/CBSubsetWeightOperator.h:11-63)
* multi-scale SSIM (xray/mssim.cpp:14-121) over a Gaussian filter restated
  from its definition (the reference's filter, Gadgetron's
  cuGaussianFilterOperator, is not in the reference sources and no version
  of it is pinned there: its tap truncation and border rule are parity
  unpinned; the convention here is fixed by the tests against scipy)

Array layout: volumes are numpy arrays of shape (nz, ny, nx) (x fastest) or
(nbins, nz, ny, nx); projections are (nviews, nv, nu).
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


class _OrcGeom(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
                ("nu", ctypes.c_int), ("nv", ctypes.c_int),
                ("fu", ctypes.c_double), ("fv", ctypes.c_double),
                ("sad", ctypes.c_double), ("sdd", ctypes.c_double)]


@dataclasses.dataclass
class Geom:
    """Volume + detector geometry (CBCT_geometry plus is_dims_in_mm)."""
    nx: int
    ny: int
    nz: int
    lx: float
    ly: float
    lz: float
    nu: int
    nv: int
    fu: float
    fv: float
    sad: float
    sdd: float

    def c(self) -> _OrcGeom:
        # Geometry scalars are FP32 in the reference (CBCT_geometry stores
        # float); round them to float before widening.
        f = lambda v: float(np.float32(v))
        return _OrcGeom(self.nx, self.ny, self.nz, f(self.lx), f(self.ly), f(self.lz),
                        self.nu, self.nv, f(self.fu), f(self.fv), f(self.sad), f(self.sdd))

    @property
    def vol_shape(self):
        return (self.nz, self.ny, self.nx)

    def proj_shape(self, nviews):
        return (nviews, self.nv, self.nu)


_lib = None
_cpu = None


def _load():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liborcbct.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", _HERE])
        _lib = ctypes.CDLL(path)
        P = ctypes.POINTER
        d = P(ctypes.c_double)
        f = P(ctypes.c_float)
        g = P(_OrcGeom)
        i = ctypes.c_int
        _lib.orc_fp.argtypes = [g, f, f, i, d, d, ctypes.c_float, i]
        _lib.orc_fp_transpose.argtypes = [g, f, f, i, d, d, ctypes.c_float, i]
        _lib.orc_bp.argtypes = [g, f, f, i, d, d, i, i]
        _lib.orc_bp_transpose.argtypes = [g, f, f, i, d, d, i, i]
        _lib.orc_fp_rays.argtypes = [g, f, f, d, ctypes.c_float, ctypes.c_size_t, P(i), d]
        _lib.orc_bp_voxels.argtypes = [g, f, f, i, d, i, ctypes.c_size_t, P(ctypes.c_longlong), d]
        _lib.orc_cosine_weights.argtypes = [g, d]
        u8 = P(ctypes.c_ubyte)
        _lib.orc_spacecarver.argtypes = [g, f, f, i, d, ctypes.c_double, u8, u8]
        _lib.orc_guo_filter.argtypes = [i, d]
        _lib.orc_filter_kernel.argtypes = [i, d, d]
        _lib.orc_filter_rows.argtypes = [i, ctypes.c_size_t, d, d]
        _lib.orc_parker_weights.argtypes = [g, f, i, d]
        _lib.orc_offset_weights.argtypes = [g, f, i, i, d]
        _lib.orc_tv_primal_dual.argtypes = [i, i, i, d, d, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, i]
        _lib.orc_parker_weight.argtypes = [ctypes.c_double] * 5
        _lib.orc_parker_weight.restype = ctypes.c_double
        _lib.orc_num_samples.argtypes = [ctypes.c_float, i]
        _lib.orc_num_samples.restype = i
    return _lib


def _cpu_lib():
    global _cpu
    if _cpu is None:
        path = os.path.join(_HERE, "libcpucbct.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", _HERE])
        _cpu = ctypes.CDLL(path)
        P = ctypes.POINTER
        f = P(ctypes.c_float)
        g = P(_OrcGeom)
        i = ctypes.c_int
        _cpu.cpu_fp_f32.argtypes = [g, f, f, i, f, f, ctypes.c_float]
        _cpu.cpu_bp_f32.argtypes = [g, f, f, i, f, f, i]
    return _cpu


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fp_(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _views(angles, offsets, nviews=None):
    a = np.ascontiguousarray(angles, dtype=np.float32).ravel()
    n = a.size if nviews is None else nviews
    if offsets is None:
        o = np.zeros(2 * n, np.float32)
    else:
        o = np.ascontiguousarray(offsets, dtype=np.float32).reshape(-1)
    if o.size != 2 * a.size:
        raise ValueError("offsets must hold (u, v) per view")
    return a, o


# --------------------------------------------------------------------------
# Projector kernels
# --------------------------------------------------------------------------

def num_samples(samples_per_pixel: float, nx: int) -> int:
    return _load().orc_num_samples(ctypes.c_float(samples_per_pixel), nx)


def fp(geom: Geom, angles, offsets, vol, samples_per_pixel=1.5, out=None, accumulate=False):
    """Forward projection (xray/conebeam_projection.cu:412-528)."""
    lib = _load()
    a, o = _views(angles, offsets)
    v = np.ascontiguousarray(vol, dtype=np.float64).reshape(geom.vol_shape)
    if out is None:
        out = np.zeros(geom.proj_shape(a.size), np.float64)
    assert out.dtype == np.float64 and out.flags.c_contiguous
    g = geom.c()
    lib.orc_fp(ctypes.byref(g), _fp_(a), _fp_(o), a.size, _dp(v), _dp(out),
               ctypes.c_float(samples_per_pixel), int(accumulate))
    return out


def fp_rays(geom: Geom, angles, offsets, vol, rays, samples_per_pixel=1.5):
    """:func:`fp` at a list of rays only: ``rays`` is (nrays, 3) of (view, u, v)
    indices; returns the nrays projection values fp() gives there."""
    lib = _load()
    a, o = _views(angles, offsets)
    v = np.ascontiguousarray(vol, dtype=np.float64).reshape(geom.vol_shape)
    r = np.ascontiguousarray(rays, dtype=np.int32).reshape(-1, 3)
    if r.size and (r[:, 0].min() < 0 or r[:, 0].max() >= a.size or r[:, 1].min() < 0
                   or r[:, 1].max() >= geom.nu or r[:, 2].min() < 0 or r[:, 2].max() >= geom.nv):
        raise ValueError("ray index out of range")
    out = np.zeros(len(r), np.float64)
    g = geom.c()
    lib.orc_fp_rays(ctypes.byref(g), _fp_(a), _fp_(o), _dp(v), ctypes.c_float(samples_per_pixel),
                    len(r), r.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(out))
    return out


def fp_transpose(geom: Geom, angles, offsets, proj, samples_per_pixel=1.5, out=None,
                 accumulate=False):
    """Exact transpose of :func:`fp` (scatter)."""
    lib = _load()
    a, o = _views(angles, offsets)
    p = np.ascontiguousarray(proj, dtype=np.float64).reshape(geom.proj_shape(a.size))
    if out is None:
        out = np.zeros(geom.vol_shape, np.float64)
    g = geom.c()
    lib.orc_fp_transpose(ctypes.byref(g), _fp_(a), _fp_(o), a.size, _dp(p), _dp(out),
                         ctypes.c_float(samples_per_pixel), int(accumulate))
    return out


def bp(geom: Geom, angles, offsets, proj, fbp=False, out=None, accumulate=False):
    """Voxel-driven backprojection (xray/conebeam_projection.cu:922-1037)."""
    lib = _load()
    a, o = _views(angles, offsets)
    p = np.ascontiguousarray(proj, dtype=np.float64).reshape(geom.proj_shape(a.size))
    if out is None:
        out = np.zeros(geom.vol_shape, np.float64)
    assert out.dtype == np.float64 and out.flags.c_contiguous
    g = geom.c()
    lib.orc_bp(ctypes.byref(g), _fp_(a), _fp_(o), a.size, _dp(p), _dp(out), int(fbp),
               int(accumulate))
    return out


def bp_voxels(geom: Geom, angles, offsets, proj, voxels, fbp=False):
    """:func:`bp` at a list of voxels only (linear indices x + nx*(y + ny*z));
    returns the values bp() gives there."""
    lib = _load()
    a, o = _views(angles, offsets)
    p = np.ascontiguousarray(proj, dtype=np.float64).reshape(geom.proj_shape(a.size))
    vx = np.ascontiguousarray(voxels, dtype=np.int64).ravel()
    if vx.size and (vx.min() < 0 or vx.max() >= geom.nx * geom.ny * geom.nz):
        raise ValueError("voxel index out of range")
    out = np.zeros(vx.size, np.float64)
    g = geom.c()
    lib.orc_bp_voxels(ctypes.byref(g), _fp_(a), _fp_(o), a.size, _dp(p), int(fbp), vx.size,
                      vx.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), _dp(out))
    return out


def bp_transpose(geom: Geom, angles, offsets, vol, fbp=False, out=None, accumulate=False):
    """Exact transpose of :func:`bp` (splat)."""
    lib = _load()
    a, o = _views(angles, offsets)
    v = np.ascontiguousarray(vol, dtype=np.float64).reshape(geom.vol_shape)
    if out is None:
        out = np.zeros(geom.proj_shape(a.size), np.float64)
    g = geom.c()
    lib.orc_bp_transpose(ctypes.byref(g), _fp_(a), _fp_(o), a.size, _dp(v), _dp(out),
                         int(fbp), int(accumulate))
    return out


def spacecarver(geom: Geom, angles, offsets, proj, limit):
    """Space-carving mask (xray/conebeam_projection.cu:822-921, 1043-1167):
    returns (mask, ambiguous), uint8 volumes; ambiguous marks voxels whose
    detector coordinate lies within 1e-3 texel of a texel edge in some view."""
    lib = _load()
    a, o = _views(angles, offsets)
    p = np.ascontiguousarray(proj, dtype=np.float64).reshape(geom.proj_shape(a.size))
    mask = np.zeros(geom.vol_shape, np.uint8)
    amb = np.zeros(geom.vol_shape, np.uint8)
    g = geom.c()
    u8 = ctypes.POINTER(ctypes.c_ubyte)
    lib.orc_spacecarver(ctypes.byref(g), _fp_(a), _fp_(o), a.size, _dp(p), float(limit),
                        mask.ctypes.data_as(u8), amb.ctypes.data_as(u8))
    return mask, amb


def apply_mask(image, mask):
    """image * mask (apply_mask, xray/conebeam_projection.cu:1651-1657)."""
    return np.asarray(image, np.float64) * (np.asarray(mask) != 0)


# --------------------------------------------------------------------------
# FDK pre-filter pieces
# --------------------------------------------------------------------------

def cosine_weights(geom: Geom):
    """(nv, nu) cosine weights (xray/hoCuConebeamProjectionOperator.cpp:64-102)."""
    w = np.zeros((geom.nv, geom.nu), np.float64)
    g = geom.c()
    _load().orc_cosine_weights(ctypes.byref(g), _dp(w))
    return w


def guo_filter(nu: int):
    """Default frequency filter, nu+1 taps (hoCuConebeamProjectionOperator.cpp:26-62)."""
    F = np.zeros(nu + 1, np.float64)
    _load().orc_guo_filter(nu, _dp(F))
    return F


def ramlak_filter(nu: int):
    """Ram-Lak |k| response on the nu+1 one-sided bins (a caller-supplied
    filter; the reference accepts any through set_frequency_filter)."""
    return np.arange(nu + 1, dtype=np.float64)


def filter_kernel(F):
    """Spatial kernel h[0..nu-1] of the pad/FFT/crop filter chain."""
    F = np.ascontiguousarray(F, dtype=np.float64)
    nu = F.size - 1
    h = np.zeros(nu, np.float64)
    _load().orc_filter_kernel(nu, _dp(F), _dp(h))
    return h


def filter_rows(proj, F):
    """Apply the ramp filter to every detector row (exact Toeplitz form)."""
    p = np.array(proj, dtype=np.float64, copy=True, order="C")
    nu = p.shape[-1]
    h = filter_kernel(F)
    _load().orc_filter_rows(nu, p.size // nu, _dp(h), _dp(p))
    return p


def parker_weights(geom: Geom, angles, proj):
    """Apply short-scan weights in place (conebeam_projection.cu:141-227)."""
    a, _ = _views(angles, None)
    p = np.ascontiguousarray(proj, dtype=np.float64)
    g = geom.c()
    _load().orc_parker_weights(ctypes.byref(g), _fp_(a), a.size, _dp(p))
    return p


def parker_weight(beta, alpha, delta, q=0.1, eps=0.001):
    """Short-scan weight of one ray (source angle beta, fan angle alpha)."""
    return _load().orc_parker_weight(beta, alpha, delta, q, eps)


def offset_weights(geom: Geom, offsets, proj, sqrt_variant=False):
    """Apply half-fan offset weights in place (conebeam_projection.cu:230-331)."""
    o = np.ascontiguousarray(offsets, dtype=np.float32).reshape(-1)
    p = np.ascontiguousarray(proj, dtype=np.float64)
    g = geom.c()
    _load().orc_offset_weights(ctypes.byref(g), _fp_(o), o.size // 2, int(sqrt_variant), _dp(p))
    return p


def fdk_prefilter(geom: Geom, angles, offsets, proj, F, short_scan=False, offset_correct=False,
                  cos_w=None):
    """Per-view FDK chain (conebeam_projection.cu:1483-1521): cosine weights,
    [Parker], ramp filter, [full offset weights]."""
    p = np.array(proj, dtype=np.float64, copy=True, order="C")
    p *= cosine_weights(geom) if cos_w is None else cos_w
    if short_scan:
        parker_weights(geom, angles, p)
    p = filter_rows(p, F)
    if offset_correct:
        offset_weights(geom, offsets, p, sqrt_variant=False)
    return p


# --------------------------------------------------------------------------
# TV and element-wise solver steps
# --------------------------------------------------------------------------

def tv_primal_dual(x, out, sigma, alpha=0.0, weight=1.0, accumulate=True):
    """cuTVPrimalDualOperator::primalDual (operators/cuTVPrimalDualOperator.cu:69-106);
    loops the 3-D kernel over trailing dimensions."""
    lib = _load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    assert out.dtype == np.float64 and out.flags.c_contiguous and out.shape == x.shape
    nz, ny, nx = x.shape[-3:]
    n3 = nx * ny * nz
    xf = x.reshape(-1)
    of = out.reshape(-1)
    for b in range(x.size // n3):
        xs = np.ascontiguousarray(xf[b * n3:(b + 1) * n3])
        os_ = np.ascontiguousarray(of[b * n3:(b + 1) * n3])
        lib.orc_tv_primal_dual(nx, ny, nz, _dp(xs), _dp(os_), sigma, alpha, weight,
                               int(accumulate))
        of[b * n3:(b + 1) * n3] = os_
    return out


def update_g(g, x, s, precon, tau, avg_lambda):
    """osMOMSolverD::updateG (solvers/osMOMSolverD.h:233-247, base class)."""
    x = x - tau * g
    x = x + (tau / avg_lambda) * (s / precon)
    return x / ((1.0 / precon) * (tau / avg_lambda) + 1.0)


# --------------------------------------------------------------------------
# Operators
# --------------------------------------------------------------------------

class Binning:
    """CBCT_binning (xray/CBCT_binning.h:18-172): a list of bins of view indices."""

    def __init__(self, bins):
        self.bins = [list(int(i) for i in b) for b in bins]

    @staticmethod
    def default_3d(nviews):
        return Binning([list(range(nviews))])

    def get_3d_binning(self):
        u = sorted(set(i for b in self.bins for i in b))
        return Binning([u])

    def __len__(self):
        return len(self.bins)


def _short_scan(geom: Geom, angles):
    """hoCuConebeamProjectionOperator.h:57-65 (float arithmetic)."""
    span = np.float32(np.max(angles))
    delta = np.float32(np.arctan(np.float32(geom.fu) / (np.float32(2.0) * np.float32(geom.sdd))))
    pi = np.float32(np.pi)
    return not (span * pi / np.float32(180.0) > pi + np.float32(3.0) * delta)


def _auto_offset(geom: Geom, offsets):
    """hoCuConebeamProjectionOperator.h:72-76: |mean offset_u| > 0.1 FOV_u."""
    o = np.asarray(offsets, np.float32).reshape(-1, 2)
    mean = np.float32(np.sum(o[:, 0], dtype=np.float32) / np.float32(len(o)))
    return abs(mean) > np.float32(geom.fu) * np.float32(0.1)


class CuConebeamOperator:
    """Device-resident projector semantics (xray/cuConebeamProjectionOperator.cpp).
    Angles/offsets are already the operator's own (setup with
    transform_angles=False, as CBSubsetOperator calls it)."""

    def __init__(self, geom: Geom, angles, offsets, binning: Binning, samples_per_pixel=1.5,
                 use_offset_correction=None):
        self.geom = geom
        self.all_angles = np.asarray(angles, np.float32)
        self.all_offsets = np.asarray(offsets, np.float32).reshape(-1, 2)
        self.binning = binning
        self.spp = samples_per_pixel
        self.short_scan = _short_scan(geom, self.all_angles) if len(self.all_angles) else False
        if use_offset_correction is None:
            use_offset_correction = _auto_offset(geom, self.all_offsets) if len(self.all_angles) else False
        self.use_offset_correction = use_offset_correction
        self.angles = [self.all_angles[b] for b in binning.bins]
        self.offsets = [self.all_offsets[b] for b in binning.bins]

    def mult_M(self, vol, out, accumulate=False):
        g = self.geom
        n3 = g.nx * g.ny * g.nz
        vf = vol.reshape(-1)
        of = out.reshape(-1)
        ip = op = 0
        for b, bin_ in enumerate(self.binning.bins):
            if len(bin_) == 0:
                continue
            nb = len(self.angles[b])
            view = np.ascontiguousarray(of[op:op + nb * g.nu * g.nv]).reshape(g.proj_shape(nb))
            fp(g, self.angles[b], self.offsets[b], vf[ip:ip + n3], self.spp, out=view,
               accumulate=accumulate)
            of[op:op + nb * g.nu * g.nv] = view.reshape(-1)
            ip += n3
            op += nb * g.nu * g.nv

    def mult_MH(self, proj, out, accumulate=False):
        g = self.geom
        n3 = g.nx * g.ny * g.nz
        pf = proj.reshape(-1)
        of = out.reshape(-1)
        ip = op = 0
        for b, bin_ in enumerate(self.binning.bins):
            if len(bin_) == 0:
                continue
            nb = len(self.angles[b])
            pin = pf[ip:ip + nb * g.nu * g.nv].reshape(g.proj_shape(nb)).astype(np.float64)
            if self.use_offset_correction:
                pin = offset_weights(g, self.offsets[b], pin.copy(), sqrt_variant=False)
            view = np.ascontiguousarray(of[op:op + n3]).reshape(g.vol_shape)
            bp(g, self.angles[b], self.offsets[b], pin, fbp=False, out=view, accumulate=accumulate)
            of[op:op + n3] = view.reshape(-1)
            ip += nb * g.nu * g.nv
            op += n3


def subset_split(nviews, binning: Binning, nsubsets):
    """CBSubsetOperator::setup view bookkeeping (xray/CBSubsetOperator.cpp:68-104).
    Returns (subset_projections, permutation, bin_of_view).  Within a subset
    views are ordered by bin with a stable sort (ascending view order)."""
    bins = {}
    for b, bin_ in enumerate(binning.bins):
        for p in bin_:
            bins[p] = b
    subsets = [[] for _ in range(nsubsets)]
    k = 0
    for i in range(nviews):
        if i in bins:
            subsets[k % nsubsets].append(i)
            k += 1
    for s in subsets:
        s.sort(key=lambda v: bins[v])
    perm = [p for s in subsets for p in s]
    return subsets, perm, bins


class CBSubsetOperator:
    """Ordered-subset operator (xray/CBSubsetOperator.cpp:19-212) plus the
    Gadgetron subsetOperator full-operator semantics: full mult_M fills every
    subset's block; full mult_MH clears (unless accumulating) and then adds
    every subset's backprojection."""

    def __init__(self, nsubsets):
        self.nsubsets = nsubsets

    def setup(self, geom: Geom, angles, offsets, projections, binning: Binning, nbins_dom=None,
              samples_per_pixel=1.5):
        """Returns the permuted projections (the reference replaces the
        acquisition's projections in place)."""
        self.geom = geom
        angles = np.asarray(angles, np.float32)
        offsets = np.asarray(offsets, np.float32).reshape(-1, 2)
        angles = (angles - np.min(angles)).astype(np.float32)
        subsets, perm, bins = subset_split(len(angles), binning, self.nsubsets)
        self.subsets, self.perm = subsets, perm
        self.angles = angles[perm]
        self.offsets = offsets[perm]
        self.ops = []
        for s in range(self.nsubsets):
            sub_bins = [[] for _ in range(len(binning))]
            for i, p in enumerate(subsets[s]):
                sub_bins[bins[p]].append(i)
            self.ops.append(CuConebeamOperator(geom, angles[subsets[s]], offsets[subsets[s]],
                                               Binning(sub_bins), samples_per_pixel))
        self.nbins = len(binning)
        self.codomain = [(len(s), geom.nv, geom.nu) for s in subsets]
        if projections is None:
            return None
        return np.ascontiguousarray(np.asarray(projections)[perm])

    def domain_shape(self):
        g = self.geom
        return (self.nbins,) + g.vol_shape if self.nbins > 1 else g.vol_shape

    def codomain_shape(self):
        return (sum(c[0] for c in self.codomain), self.geom.nv, self.geom.nu)

    def projection_subsets(self, proj):
        res, start = [], 0
        for c in self.codomain:
            res.append(proj[start:start + c[0]])
            start += c[0]
        return res

    def mult_M_subset(self, x, out, subset, accumulate=False):
        self.ops[subset].mult_M(x, out, accumulate)

    def mult_MH_subset(self, y, out, subset, accumulate=False):
        self.ops[subset].mult_MH(y, out, accumulate)

    def mult_M(self, x, out, accumulate=False):
        subs = self.projection_subsets(out)
        for i in range(self.nsubsets):
            self.mult_M_subset(x, subs[i], i, accumulate)

    def mult_MH(self, y, out, accumulate=False):
        subs = self.projection_subsets(y)
        if not accumulate:
            out[...] = 0.0
        for i in range(self.nsubsets):
            self.mult_MH_subset(subs[i], out, i, True)


class CBSubsetWeightOperator(CBSubsetOperator):
    """CBSubsetWeightOperator (xray/CBSubsetWeightOperator.h:11-63): setup
    multiplies the projections by the weights (the reference does so in the
    acquisition, in place, :20) before the subset split, and keeps the
    weights in the same permuted order; mult_M weights a subset's projection
    (:33-43), mult_MH weights its input before backprojecting (:44-49)."""

    def setup(self, geom: Geom, angles, offsets, projections, binning: Binning, weights,
              nbins_dom=None, samples_per_pixel=1.5):
        w = np.asarray(weights, np.float64)
        p = np.asarray(projections, np.float64)
        if w.shape != p.shape:
            raise ValueError("Weight dimensions must match that of the projection data")
        pb = super().setup(geom, angles, offsets, p * w, binning, nbins_dom, samples_per_pixel)
        self.weights = np.ascontiguousarray(w[self.perm])
        self.subset_weights = self.projection_subsets(self.weights)
        return pb

    def mult_M_subset(self, x, out, subset, accumulate=False):
        tmp = np.zeros(out.shape)
        super().mult_M_subset(x, tmp, subset, False)
        tmp *= self.subset_weights[subset]
        if accumulate:
            out += tmp
        else:
            out[...] = tmp

    def mult_MH_subset(self, y, out, subset, accumulate=False):
        super().mult_MH_subset(np.asarray(y, np.float64) * self.subset_weights[subset], out, subset,
                               accumulate)


# --------------------------------------------------------------------------
# Solvers
# --------------------------------------------------------------------------

def os_sart(op: CBSubsetOperator, b, iterations, beta=1.0, gamma=0.0, nonneg=False, x0=None,
            objective=None):
    """osSARTSolver::solve (solvers/osSARTSolver.h:33-126).  ``b`` is the
    permuted projection array.  If ``objective`` is a list, ||Ax-b||^2 after
    each iteration is appended (the reference prints it)."""
    dom = op.domain_shape()
    x = np.zeros(dom) if x0 is None else np.array(x0, np.float64)
    b = np.asarray(b, np.float64)
    subsets = op.projection_subsets(b)
    ones_proj = np.zeros(b.shape)
    op.mult_M(np.ones(dom), ones_proj, False)
    ones_proj = 1.0 / np.maximum(ones_proj, 1e-6)
    W_p = op.projection_subsets(ones_proj)
    tmp_projection = np.ones(b.shape)
    tmp = op.projection_subsets(tmp_projection)
    W_x = []
    for i in range(op.nsubsets):
        img = np.zeros(dom)
        op.mult_MH_subset(tmp[i], img, i, False)
        W_x.append(1.0 / np.maximum(img, 1e-6))
    order = list(range(op.nsubsets))
    tmp_image = np.zeros(dom)
    for it in range(iterations):
        for s in order:
            op.mult_M_subset(x, tmp[s], s, False)
            tmp[s] -= subsets[s]
            tmp[s] *= -1.0
            tmp[s] *= W_p[s]
            op.mult_MH_subset(tmp[s], tmp_image, s, False)
            tmp_image *= W_x[s]
            x += (beta / (1.0 + gamma * it)) * tmp_image
            if nonneg:
                np.maximum(x, 0.0, out=x)
        order.reverse()
        if objective is not None:
            r = np.zeros(b.shape)
            op.mult_M(x, r, False)
            r -= b
            objective.append(float(np.dot(r.ravel(), r.ravel())))
    return x


def os_mom(op: CBSubsetOperator, b, iterations, tv_weight=0.0, tau0=1e-3, reg_steps=1,
           beta=1.0, gamma=0.0, nonneg=False, x0=None, precon=None, tv_alpha=0.0):
    """osMOMSolverD::solve with an optional single isotropic TV regulariser
    (solvers/osMOMSolverD.h:82-247; operators/cuTVPrimalDualOperator.cu).
    Returns z (the extrapolated iterate, as the reference does)."""
    dom = op.domain_shape()
    z = np.zeros(dom) if x0 is None else np.array(x0, np.float64)
    x = z.copy()
    xold = z.copy()
    b = np.asarray(b, np.float64)
    subsets = op.projection_subsets(b)
    tmp_projection = np.zeros(b.shape)
    tmp = op.projection_subsets(tmp_projection)
    if precon is None:
        P = np.ones(dom)
        op.mult_M(P, tmp_projection, False)
        op.mult_MH(tmp_projection, P, False)
        P = 1.0 / np.maximum(P, 1e-6)
    else:
        P = np.asarray(precon, np.float64)
    regs = [tv_weight] if tv_weight > 0 else []
    avg_lambda = sum(w / len(regs) for w in regs) if regs else 0.0
    regs = [w / avg_lambda for w in regs]
    t = 1.0
    told = 1.0
    S = op.nsubsets
    for it in range(iterations):
        for s in range(S):
            t = 0.5 * (1.0 + np.sqrt(1.0 + 4.0 * t * t))
            op.mult_M_subset(x, tmp[s], s, False)
            tmp[s] -= subsets[s]
            img = np.zeros(dom)
            op.mult_MH_subset(tmp[s], img, s, False)
            img *= P
            z += -(beta / (1.0 + gamma * it)) * S * img
            # denoise(x, z, P, avg_lambda)
            if not regs:
                x[...] = z
            else:
                gam = 0.35 / avg_lambda / (P.size / np.sum(np.abs(P)))
                tau = tau0
                sigma = 1.0 / (tau * 16.0)
                for _ in range(reg_steps):
                    g = np.zeros(dom)
                    for w in regs:
                        tv_primal_dual(x, g, sigma, tv_alpha, w, accumulate=True)
                    x[...] = update_g(g, x, z, P, tau, avg_lambda)
                    theta = 1.0 / np.sqrt(1.0 + 2.0 * gam * tau)
                    tau *= theta
                    sigma /= theta
            if nonneg:
                np.maximum(x, 0.0, out=x)
            z[...] = x * (1.0 + (told - 1.0) / t) - ((told - 1.0) / t) * xold
            x, xold = xold, x
            x[...] = z
            told = t
    return z


# --------------------------------------------------------------------------
# FDK
# --------------------------------------------------------------------------

def fdk_mult_MH(geom: Geom, angles, offsets, proj, binning: Binning, F, short_scan=None,
                use_offset_correction=None, out=None, accumulate=False):
    """hoCuConebeamProjectionOperator::mult_MH in FBP mode
    (xray/hoCuConebeamProjectionOperator.cpp:179-266).  ``angles`` must
    already be shifted by setup."""
    angles = np.asarray(angles, np.float32)
    offsets = np.asarray(offsets, np.float32).reshape(-1, 2)
    if short_scan is None:
        short_scan = _short_scan(geom, angles)
    if use_offset_correction is None:
        use_offset_correction = _auto_offset(geom, offsets)
    nb = len(binning)
    shape = ((nb,) if nb > 1 else ()) + geom.vol_shape
    if out is None:
        out = np.zeros(shape, np.float64)
    of = out.reshape(nb, -1)
    cw = cosine_weights(geom)
    for b, bin_ in enumerate(binning.bins):
        view = np.ascontiguousarray(of[b]).reshape(geom.vol_shape)
        if len(bin_) == 0:
            if not accumulate:
                view[...] = 0.0
        else:
            pb = np.asarray(proj, np.float64)[bin_]
            filt = fdk_prefilter(geom, angles[bin_], offsets[bin_], pb, F, short_scan,
                                 use_offset_correction, cos_w=cw)
            bp(geom, angles[bin_], offsets[bin_], filt, fbp=True, out=view, accumulate=accumulate)
        # the reference scales by 2/|bin| unconditionally: an empty bin's zero
        # volume becomes 0 * inf = NaN there, and here
        with np.errstate(invalid="ignore"):
            view *= np.float32(2.0) / np.float32(len(bin_)) if len(bin_) else np.inf
        of[b] = view.reshape(-1)
    return out


def ho_mult_MH(geom: Geom, angles, offsets, proj, binning: Binning, use_fbp=False, F=None,
               short_scan=None, use_offset_correction=None, out=None, accumulate=False):
    """hoCuConebeamProjectionOperator::mult_MH
    (xray/hoCuConebeamProjectionOperator.cpp:179-266): the FBP branch is
    :func:`fdk_mult_MH`; otherwise per bin the square-root half-fan weights
    when offset correction is on (conebeam_projection.cu:1522-1523) and the
    plain backprojection.  ``angles`` must already be shifted by setup."""
    angles = np.asarray(angles, np.float32)
    offsets = np.asarray(offsets, np.float32).reshape(-1, 2)
    if use_offset_correction is None:
        use_offset_correction = _auto_offset(geom, offsets)
    if use_fbp:
        return fdk_mult_MH(geom, angles, offsets, proj, binning, F, short_scan,
                           use_offset_correction, out=out, accumulate=accumulate)
    nb = len(binning)
    shape = ((nb,) if nb > 1 else ()) + geom.vol_shape
    if out is None:
        out = np.zeros(shape, np.float64)
    of = out.reshape(nb, -1)
    for b, bin_ in enumerate(binning.bins):
        view = np.ascontiguousarray(of[b]).reshape(geom.vol_shape)
        if len(bin_) == 0:
            if not accumulate:
                view[...] = 0.0
        else:
            pb = np.array(np.asarray(proj, np.float64)[bin_])
            if use_offset_correction:
                pb = offset_weights(geom, offsets[bin_], pb, sqrt_variant=True)
            bp(geom, angles[bin_], offsets[bin_], pb, fbp=False, out=view, accumulate=accumulate)
        of[b] = view.reshape(-1)
    return out


def ho_mult_M(geom: Geom, angles, offsets, vol, binning: Binning, samples_per_pixel=1.5,
              out=None, use_offset_correction=False, use_fbp=False):
    """hoCuConebeamProjectionOperator::mult_M (xray/hoCuConebeamProjectionOperator.cpp:105-176),
    non-accumulating form: each bin's volume is projected into the bin's views."""
    angles = np.asarray(angles, np.float32)
    offsets = np.asarray(offsets, np.float32).reshape(-1, 2)
    if out is None:
        out = np.zeros(geom.proj_shape(len(angles)))
    n3 = geom.nx * geom.ny * geom.nz
    vf = np.asarray(vol, np.float64).reshape(-1)
    for b, bin_ in enumerate(binning.bins):
        if not bin_:
            continue
        p = fp(geom, angles[bin_], offsets[bin_], vf[b * n3:(b + 1) * n3], samples_per_pixel)
        out[bin_] = p
    if use_offset_correction and not use_fbp:
        offset_weights(geom, offsets, out, sqrt_variant=True)
    return out


def fdk_reconstruct_3d(geom: Geom, angles, offsets, proj, F, binning: Binning = None,
                       samples_per_pixel=1.5):
    """FDK_reconstruct_3d driver (xray/FDK_reconstruct_3d.cpp:128-161):
    setup (angles -= min), FBP mult_MH, mult_M, least-squares rescale.
    Returns (volume, scale)."""
    angles = np.asarray(angles, np.float32)
    angles = (angles - np.min(angles)).astype(np.float32)
    if offsets is None:
        offsets = np.zeros((len(angles), 2), np.float32)
    if binning is None:
        binning = Binning.default_3d(len(angles))
    f = fdk_mult_MH(geom, angles, offsets, proj, binning, F)
    Af = ho_mult_M(geom, angles, offsets, f, binning, samples_per_pixel, use_fbp=True)
    p = np.asarray(proj, np.float64)
    scale = float(np.dot(Af.ravel(), p.ravel()) / np.dot(Af.ravel(), Af.ravel()))
    return f * scale, scale


def fdk_reconstruct_4d(geom: Geom, angles, offsets, proj, F, binning: Binning,
                       samples_per_pixel=1.5):
    """FDK_reconstruct_4d driver (xray/FDK_reconstruct_4d.cpp:100-170), in the
    reference's own steps: FBP over the 3-D binning (union of the bins, :112,
    :126-130); the result expanded into every bin and projected bin by bin
    through the 4-D operator (:145-150; views in no bin stay zero); the
    least-squares scale (:151); the scaled projection subtracted (:153-154);
    the 4-D FBP of the residual plus the expanded image, times the scale
    (:158-166).  Returns (volume (nbins, nz, ny, nx), fdk_3d, scale)."""
    angles = np.asarray(angles, np.float32)
    angles = (angles - np.min(angles)).astype(np.float32)
    if offsets is None:
        offsets = np.zeros((len(angles), 2), np.float32)
    b3 = binning.get_3d_binning()
    f3 = fdk_mult_MH(geom, angles, offsets, proj, b3, F)
    nb = len(binning)
    fdk = np.broadcast_to(f3, (nb,) + f3.shape).copy()
    diff = ho_mult_M(geom, angles, offsets, fdk, binning, samples_per_pixel, use_fbp=True)
    p = np.asarray(proj, np.float64)
    scale = float(np.dot(diff.ravel(), p.ravel()) / np.dot(diff.ravel(), diff.ravel()))
    resid = p - scale * diff
    res = fdk_mult_MH(geom, angles, offsets, resid, binning, F).reshape(fdk.shape)
    return (res + fdk) * scale, f3, scale


# --------------------------------------------------------------------------
# Multi-scale SSIM
# --------------------------------------------------------------------------

def gauss_taps(sigma, truncate=3.5):
    """Normalised taps exp(-k^2/(2 sigma^2)), |k| <= int(truncate sigma + 0.5)."""
    sigma = float(sigma)
    if sigma <= 0.0:
        return np.ones(1)
    r = int(truncate * sigma + 0.5)
    k = np.arange(-r, r + 1, dtype=np.float64)
    w = np.exp(-0.5 * k * k / (sigma * sigma))
    return w / w.sum()


def _reflect(i, n):
    """Half-sample symmetric index (-1 -> 0, n -> n-1), any distance."""
    period = 2 * n
    i = np.mod(i, period)
    return np.where(i < n, i, period - 1 - i)


def gauss_filter(vol, sigma_xyz, truncate=3.5):
    """Separable Gaussian filter of a (nz, ny, nx) array; sigma per (x, y, z)."""
    out = np.asarray(vol, np.float64)
    for axis, sigma in ((2, sigma_xyz[0]), (1, sigma_xyz[1]), (0, sigma_xyz[2])):
        w = gauss_taps(sigma, truncate)
        r = (w.size - 1) // 2
        n = out.shape[axis]
        acc = np.zeros_like(out)
        for j, wk in enumerate(w):
            idx = _reflect(np.arange(n) + (j - r), n)
            acc += wk * np.take(out, idx, axis=axis)
        out = acc
    return out


def _cs_map(X, Y, sigma, c2):
    """CS (xray/mssim.cpp:14-74): (2 cov + C2) / (var_x + var_y + C2)."""
    ux, uy = gauss_filter(X, sigma), gauss_filter(Y, sigma)
    xx, yy, xy = gauss_filter(X * X, sigma), gauss_filter(Y * Y, sigma), gauss_filter(X * Y, sigma)
    return (2.0 * (xy - ux * uy) + c2) / ((xx - ux * ux) + (yy - uy * uy) + c2)


def mssim(image, reference, sigma_xyz=(1.5, 1.5, 1.5), scales=3):
    """Gadgetron::mssim (xray/mssim.cpp:77-121) in float64: the CS maps at
    sigma 2^i for i < scales multiplied together, times the luminance term at
    sigma 2^(scales-1), averaged.  C1 = (0.01 R)^2, C2 = (0.03 R)^2 with R
    the reference's range, formed in float32 as the reference does."""
    X = np.asarray(image, np.float64)
    Y = np.asarray(reference, np.float64)
    rng = np.float32(np.float32(Y.max()) - np.float32(Y.min()))
    c1 = float(np.float32(np.float64(np.float32(0.01) * rng) ** 2))
    c2 = float(np.float32(np.float64(np.float32(0.03) * rng) ** 2))
    sig = np.asarray(sigma_xyz, np.float32)
    S = np.ones_like(X)
    for i in range(scales):
        S *= _cs_map(X, Y, sig * np.float32(2.0 ** i), c2)
    top = sig * np.float32(2.0 ** (scales - 1))
    ux, uy = gauss_filter(X, top), gauss_filter(Y, top)
    S *= (2.0 * ux * uy + c1) / (ux * ux + uy * uy + c1)
    return float(S.mean())


# --------------------------------------------------------------------------
# FP32 CPU baseline (timed beside the GPU path)
# --------------------------------------------------------------------------

def cpu_fp_f32(geom: Geom, angles, offsets, vol, samples_per_pixel=1.5):
    a, o = _views(angles, offsets)
    v = np.ascontiguousarray(vol, dtype=np.float32)
    out = np.zeros(geom.proj_shape(a.size), np.float32)
    g = geom.c()
    _cpu_lib().cpu_fp_f32(ctypes.byref(g), _fp_(a), _fp_(o), a.size, _fp_(v), _fp_(out),
                          ctypes.c_float(samples_per_pixel))
    return out


def cpu_bp_f32(geom: Geom, angles, offsets, proj, fbp=False):
    a, o = _views(angles, offsets)
    p = np.ascontiguousarray(proj, dtype=np.float32)
    out = np.zeros(geom.vol_shape, np.float32)
    g = geom.c()
    _cpu_lib().cpu_bp_f32(ctypes.byref(g), _fp_(a), _fp_(o), a.size, _fp_(p), _fp_(out), int(fbp))
    return out
