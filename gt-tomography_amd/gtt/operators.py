"""ctypes binding of libgtt.so (include/gtt_c.h): the C++ operator/solver
layer -- FDK_reconstruct_3d, osSARTSolver and osMOMSolverD over
CBSubsetOperator, and the subset operator itself."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import capi

_lib = None


def lib_path():
    return os.path.join(capi.LIBDIR, "libgtt.so")


def load():
    global _lib
    if _lib is not None:
        return _lib
    capi.load()
    if not os.path.exists(lib_path()):
        raise ImportError("libgtt.so not built (run __graft_entry__.build()); there is no CPU fallback")
    L = ctypes.CDLL(lib_path())
    P = ctypes.c_void_p
    I = ctypes.c_int
    F = ctypes.c_float
    G = ctypes.POINTER(capi.Geom)
    D = ctypes.POINTER(ctypes.c_double)
    L.gtt_last_error.restype = ctypes.c_char_p
    L.gtt_last_error.argtypes = []
    L.gtt_set_default_storage.argtypes = [I]
    L.gtt_set_default_storage.restype = I
    L.gtt_set_autotune.argtypes = [I]
    L.gtt_set_autotune.restype = I
    L.gtt_set_graphs.argtypes = [I]
    L.gtt_set_graphs.restype = I
    L.gtt_graph_launches.argtypes = []
    L.gtt_graph_launches.restype = ctypes.c_longlong
    L.gtt_fdk_reconstruct_3d.argtypes = [G, P, P, I, P, P, I, P, P, F, P, D]
    L.gtt_fdk_reconstruct_4d.argtypes = [G, P, P, I, P, P, I, P, P, F, P, P, D]
    L.gtt_os_sart.argtypes = [G, P, P, I, P, P, I, P, I, I, F, F, I, P, D]
    L.gtt_os_mom.argtypes = [G, P, P, I, P, P, I, P, I, I, F, F, I, F, F, I, I, P]
    L.gtt_os_sart_weighted.argtypes = [G, P, P, I, P, P, I, P, P, I, I, F, F, I, P, D]
    L.gtt_os_mom_weighted.argtypes = [G, P, P, I, P, P, I, P, P, I, I, F, F, I, F, F, I, I, P]
    L.gtt_subset_op_setup_weighted.argtypes = [P, G, P, P, I, P, P, I, P, P, P, P]
    L.gtt_subset_op_create.argtypes = [I, ctypes.POINTER(P)]
    L.gtt_subset_op_destroy.argtypes = [P]
    L.gtt_subset_op_setup.argtypes = [P, G, P, P, I, P, P, I, P, P, P]
    L.gtt_subset_op_views.argtypes = [P, I]
    L.gtt_subset_op_mult_M.argtypes = [P, P, P, I, I]
    L.gtt_subset_op_mult_MH.argtypes = [P, P, P, I, I]
    for n in ("gtt_fdk_reconstruct_3d", "gtt_fdk_reconstruct_4d", "gtt_os_sart_weighted",
              "gtt_os_mom_weighted", "gtt_subset_op_setup_weighted", "gtt_os_sart", "gtt_os_mom", "gtt_subset_op_create",
              "gtt_subset_op_destroy", "gtt_subset_op_setup", "gtt_subset_op_views",
              "gtt_subset_op_mult_M", "gtt_subset_op_mult_MH"):
        getattr(L, n).restype = I
    _lib = L
    return L


def set_default_storage(dtype):
    """Working copies of the projectors ('f32' or 'bf16') for reconstructions
    this thread starts from now on."""
    _check(load().gtt_set_default_storage(capi.CBCT_BF16 if dtype == "bf16" else capi.CBCT_F32))


def set_autotune(on):
    """OS reconstructions this thread starts from now on tune the forward
    projector for their geometry first."""
    _check(load().gtt_set_autotune(int(bool(on))))


def set_graphs(on):
    """OS-SART reconstructions this thread starts from now on replay their
    iterations 2.. as captured HIP graphs (gamma = 0 only; same bits)."""
    _check(load().gtt_set_graphs(int(bool(on))))


def graph_launches():
    """HIP graphs this thread has launched so far."""
    return int(load().gtt_graph_launches())


def _check(rc):
    if rc != 0:
        raise RuntimeError(load().gtt_last_error().decode())


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _views(angles, offsets):
    a = np.ascontiguousarray(angles, np.float32).ravel()
    o = None if offsets is None else np.ascontiguousarray(offsets, np.float32).reshape(-1)
    return a, o


def _bins(bins):
    if bins is None:
        return None, None, 0
    idx = np.ascontiguousarray(np.concatenate([np.asarray(b, np.uint32) for b in bins]) if bins else
                               np.zeros(0, np.uint32), np.uint32)
    sizes = np.ascontiguousarray([len(b) for b in bins], np.int32)
    return idx, sizes, len(bins)


def _vol_shape(g, nbins):
    return ((nbins,) if nbins > 1 else ()) + (g.nz, g.ny, g.nx)


def fdk_reconstruct_3d(g: capi.Geom, angles, offsets, projections, filt=None, bins=None,
                       samples_per_pixel=1.5):
    """FDK_reconstruct_3d: returns (volume, least-squares scale)."""
    L = load()
    a, o = _views(angles, offsets)
    idx, sizes, nb = _bins(bins)
    p = np.ascontiguousarray(projections, np.float32)
    F = None if filt is None else np.ascontiguousarray(filt, np.float32)
    out = np.zeros(_vol_shape(g, nb), np.float32)
    scale = ctypes.c_double()
    _check(L.gtt_fdk_reconstruct_3d(ctypes.byref(g), _p(a), _p(o), a.size, _p(idx), _p(sizes), nb,
                                    _p(p), _p(F), samples_per_pixel, _p(out), ctypes.byref(scale)))
    return out, scale.value


def fdk_reconstruct_4d(g: capi.Geom, angles, offsets, projections, bins, filt=None,
                       samples_per_pixel=1.5):
    """FDK_reconstruct_4d (McKinnon-Bates): returns (volumes (nbins, nz, ny, nx),
    the FBP over the union of the bins (nz, ny, nx), least-squares scale)."""
    L = load()
    a, o = _views(angles, offsets)
    if not bins:
        raise ValueError("fdk_reconstruct_4d needs at least one bin")
    idx, sizes, nb = _bins(bins)
    p = np.ascontiguousarray(projections, np.float32)
    if p.shape != (a.size, g.nv, g.nu):
        raise ValueError(f"projections must be (nviews, nv, nu) = {(a.size, g.nv, g.nu)}, got {p.shape}")
    F = None if filt is None else np.ascontiguousarray(filt, np.float32)
    if F is not None and F.size != g.nu + 1:
        raise ValueError("the frequency filter holds nu+1 taps")
    out = np.zeros((nb, g.nz, g.ny, g.nx), np.float32)
    f3 = np.zeros((g.nz, g.ny, g.nx), np.float32)
    scale = ctypes.c_double()
    _check(L.gtt_fdk_reconstruct_4d(ctypes.byref(g), _p(a), _p(o), a.size, _p(idx), _p(sizes), nb,
                                    _p(p), _p(F), samples_per_pixel, _p(out), _p(f3),
                                    ctypes.byref(scale)))
    return out, f3, scale.value


def _weights(weights, p):
    if weights is None:
        return None
    w = np.ascontiguousarray(weights, np.float32)
    if w.shape != p.shape:
        raise ValueError("Weight dimensions must match that of the projection data")
    return w


def os_sart(g: capi.Geom, angles, offsets, projections, nsubsets, iterations, beta=1.0, gamma=0.0,
            nonneg=False, bins=None, objective=False, weights=None):
    """osSARTSolver over CBSubsetOperator (CBSubsetWeightOperator when
    per-pixel weights of the projections' shape are given); projections in
    acquisition order."""
    L = load()
    a, o = _views(angles, offsets)
    idx, sizes, nb = _bins(bins)
    p = np.ascontiguousarray(projections, np.float32)
    w = _weights(weights, p)
    out = np.zeros(_vol_shape(g, nb), np.float32)
    obj = np.zeros(iterations, np.float64) if objective else None
    _check(L.gtt_os_sart_weighted(ctypes.byref(g), _p(a), _p(o), a.size, _p(idx), _p(sizes), nb,
                                  _p(p), _p(w), nsubsets, iterations, beta, gamma, int(nonneg),
                                  _p(out), None if obj is None else
                                  obj.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return (out, obj) if objective else out


def os_mom(g: capi.Geom, angles, offsets, projections, nsubsets, iterations, tv_weight=0.0,
           tau=1e-3, reg_steps=1, beta=1.0, gamma=0.0, nonneg=False, fuse_tv=True, bins=None,
           weights=None):
    """osMOMSolverD (+ isotropic TV) over CBSubsetOperator, or
    CBSubsetWeightOperator with per-pixel weights; returns z."""
    L = load()
    a, o = _views(angles, offsets)
    idx, sizes, nb = _bins(bins)
    p = np.ascontiguousarray(projections, np.float32)
    w = _weights(weights, p)
    out = np.zeros(_vol_shape(g, nb), np.float32)
    _check(L.gtt_os_mom_weighted(ctypes.byref(g), _p(a), _p(o), a.size, _p(idx), _p(sizes), nb,
                                 _p(p), _p(w), nsubsets, iterations, tv_weight, tau, reg_steps,
                                 beta, gamma, int(nonneg), int(fuse_tv), _p(out)))
    return out


class CBSubsetOperator:
    """CBSubsetOperator<cuNDArray> on torch device tensors (default stream)."""

    def __init__(self, nsubsets):
        self.L = load()
        h = ctypes.c_void_p()
        _check(self.L.gtt_subset_op_create(int(nsubsets), ctypes.byref(h)))
        self.h = h
        self.nsubsets = nsubsets

    def __del__(self):
        try:
            if self.h:
                self.L.gtt_subset_op_destroy(self.h)
        except Exception:
            pass

    def setup(self, g: capi.Geom, angles, offsets, projections=None, bins=None, weights=None):
        """Returns (permuted projections or None, permutation).  With weights
        (the projections' shape) the operator is a CBSubsetWeightOperator and
        the returned projections are the weighted ones."""
        self.g = g
        a, o = _views(angles, offsets)
        idx, sizes, nb = _bins(bins)
        self.nbins = max(nb, 1)
        p = None if projections is None else np.ascontiguousarray(projections, np.float32)
        if weights is not None and p is None:
            raise ValueError("weighted setup needs the projections")
        w = None if weights is None else _weights(weights, p)
        perm_p = None if p is None else np.empty_like(p)
        perm = np.zeros(a.size, np.uint32)
        _check(self.L.gtt_subset_op_setup_weighted(self.h, ctypes.byref(g), _p(a), _p(o), a.size,
                                                   _p(idx), _p(sizes), nb, _p(p), _p(w),
                                                   _p(perm_p), _p(perm)))
        return perm_p, perm

    def views(self, subset=-1):
        return self.L.gtt_subset_op_views(self.h, int(subset))

    def mult_M(self, vol, proj, subset=-1, accumulate=False):
        _check(self.L.gtt_subset_op_mult_M(self.h, ctypes.c_void_p(vol.data_ptr()),
                                           ctypes.c_void_p(proj.data_ptr()), int(subset), int(accumulate)))

    def mult_MH(self, proj, vol, subset=-1, accumulate=False):
        _check(self.L.gtt_subset_op_mult_MH(self.h, ctypes.c_void_p(proj.data_ptr()),
                                            ctypes.c_void_p(vol.data_ptr()), int(subset), int(accumulate)))


def _bind_extra(L):
    P, I, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
    G = ctypes.POINTER(capi.Geom)
    IP = ctypes.POINTER(ctypes.c_int)
    for name, args in {
        "gtt_cu_op_create": [ctypes.POINTER(P)], "gtt_cu_op_destroy": [P],
        "gtt_cu_op_setup": [P, G, P, P, I, P, P, I, I, I],
        "gtt_cu_op_mult_M": [P, P, P, I], "gtt_cu_op_mult_MH": [P, P, P, I],
        "gtt_cu_op_calculate_mask": [P, P, F, P], "gtt_cu_op_set_mask": [P, P],
        "gtt_cu_op_flags": [P, IP, IP],
        "gtt_ho_op_create": [ctypes.POINTER(P)], "gtt_ho_op_destroy": [P],
        "gtt_ho_op_setup": [P, G, P, P, I, P, P, I, I, I, P, F],
        "gtt_ho_op_mult_M": [P, P, P, I], "gtt_ho_op_mult_MH": [P, P, P, I],
        "gtt_ho_op_flags": [P, IP, IP],
        "gtt_write_real": [ctypes.c_char_p, I, ctypes.POINTER(ctypes.c_size_t), P],
        "gtt_cbct_forwards_projection": [G, F, F, I, P, P, P, I, F, P, P, P],
        "gtt_read_real_header": [ctypes.c_char_p, IP, ctypes.POINTER(ctypes.c_size_t)],
        "gtt_read_real": [ctypes.c_char_p, P, ctypes.c_size_t],
    }.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = I
    return L


def _dptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class _OpBase:
    _prefix = ""

    def __init__(self):
        self.L = _bind_extra(load())
        h = ctypes.c_void_p()
        _check(getattr(self.L, self._prefix + "create")(ctypes.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if self.h:
                getattr(self.L, self._prefix + "destroy")(self.h)
                self.h = None
        except Exception:
            pass

    def flags(self):
        """(short_scan, use_offset_correction) as setup decided them."""
        a, b = ctypes.c_int(), ctypes.c_int()
        _check(getattr(self.L, self._prefix + "flags")(self.h, ctypes.byref(a), ctypes.byref(b)))
        return bool(a.value), bool(b.value)


class CuConebeamProjectionOperator(_OpBase):
    """cuConebeamProjectionOperator on torch device tensors (default stream).
    use_offset_correction: None = decided by setup, else forced."""
    _prefix = "gtt_cu_op_"

    def setup(self, g: capi.Geom, angles, offsets=None, bins=None, transform_angles=True,
              use_offset_correction=None):
        a, o = _views(angles, offsets)
        idx, sizes, nb = _bins(bins)
        uoc = -1 if use_offset_correction is None else int(bool(use_offset_correction))
        _check(self.L.gtt_cu_op_setup(self.h, ctypes.byref(g), _p(a), _p(o), a.size, _p(idx),
                                      _p(sizes), nb, int(bool(transform_angles)), uoc))
        self.g = g
        return self

    def mult_M(self, vol, proj, accumulate=False):
        _check(self.L.gtt_cu_op_mult_M(self.h, _dptr(vol), _dptr(proj), int(accumulate)))

    def mult_MH(self, proj, vol, accumulate=False):
        _check(self.L.gtt_cu_op_mult_MH(self.h, _dptr(proj), _dptr(vol), int(accumulate)))

    def calculate_mask(self, proj, limit, mask):
        """mask: uint8 device tensor of the domain's shape."""
        _check(self.L.gtt_cu_op_calculate_mask(self.h, _dptr(proj), float(limit), _dptr(mask)))
        return mask

    def set_mask(self, mask):
        _check(self.L.gtt_cu_op_set_mask(self.h, _dptr(mask)))


class HoCuConebeamProjectionOperator(_OpBase):
    """hoCuConebeamProjectionOperator on host numpy arrays."""
    _prefix = "gtt_ho_op_"

    def setup(self, g: capi.Geom, angles, offsets=None, bins=None, use_fbp=False,
              use_offset_correction=None, filt=None, samples_per_pixel=1.5):
        a, o = _views(angles, offsets)
        idx, sizes, nb = _bins(bins)
        uoc = -1 if use_offset_correction is None else int(bool(use_offset_correction))
        F = None if filt is None else np.ascontiguousarray(filt, np.float32)
        if F is not None and F.size != g.nu + 1:
            raise ValueError("the frequency filter holds nu+1 taps")
        _check(self.L.gtt_ho_op_setup(self.h, ctypes.byref(g), _p(a), _p(o), a.size, _p(idx),
                                      _p(sizes), nb, int(bool(use_fbp)), uoc, _p(F),
                                      float(samples_per_pixel)))
        self.g, self.nviews, self.nbins = g, a.size, max(nb, 1)
        return self

    def _vol_shape(self):
        return _vol_shape(self.g, self.nbins)

    def mult_M(self, image, proj, accumulate=False):
        x = np.ascontiguousarray(image, np.float32)
        if x.shape != self._vol_shape() or proj.dtype != np.float32 or not proj.flags.c_contiguous:
            raise ValueError("mult_M: image %s and float32 projections expected" % (self._vol_shape(),))
        _check(self.L.gtt_ho_op_mult_M(self.h, _p(x), _p(proj), int(accumulate)))
        return proj

    def mult_MH(self, proj, image, accumulate=False):
        y = np.ascontiguousarray(proj, np.float32)
        if image.shape != self._vol_shape() or image.dtype != np.float32 or not image.flags.c_contiguous:
            raise ValueError("mult_MH: float32 image %s expected" % (self._vol_shape(),))
        _check(self.L.gtt_ho_op_mult_MH(self.h, _p(y), _p(image), int(accumulate)))
        return image


def write_real(path, array):
    """Write a numpy array as a .real file (C order (..., ny, nx) = sizes
    listed first-index-fastest in the file)."""
    L = _bind_extra(load())
    a = np.ascontiguousarray(array, np.float32)
    dims = (ctypes.c_size_t * a.ndim)(*a.shape[::-1])
    _check(L.gtt_write_real(os.fsencode(path), a.ndim, dims, _p(a)))


def read_real(path):
    L = _bind_extra(load())
    nd = ctypes.c_int()
    dims = (ctypes.c_size_t * 16)()
    _check(L.gtt_read_real_header(os.fsencode(path), ctypes.byref(nd), dims))
    shape = tuple(dims[i] for i in range(nd.value))[::-1]
    out = np.empty(shape, np.float32)
    _check(L.gtt_read_real(os.fsencode(path), _p(out), out.size))
    return out


def cbct_forwards_projection(g: capi.Geom, image, start_angle, angular_spacing, nviews,
                             offsets=None, bins=None, samples_per_pixel=1.5):
    """CBCT_forwards_projection's computation: returns (projections
    (nviews, nv, nu), angles) for a host image (nz, ny, nx) or
    (nbins, nz, ny, nx)."""
    L = _bind_extra(load())
    _, o = _views(np.zeros(nviews, np.float32), offsets)
    idx, sizes, nb = _bins(bins)
    x = np.ascontiguousarray(image, np.float32)
    if x.shape != _vol_shape(g, max(nb, 1)):
        raise ValueError("image must be %s" % (_vol_shape(g, max(nb, 1)),))
    proj = np.empty((nviews, g.nv, g.nu), np.float32)
    angles = np.empty(nviews, np.float32)
    _check(L.gtt_cbct_forwards_projection(ctypes.byref(g), float(start_angle), float(angular_spacing),
                                          int(nviews), _p(o), _p(idx), _p(sizes), nb,
                                          float(samples_per_pixel), _p(x), _p(proj), _p(angles)))
    return proj, angles
