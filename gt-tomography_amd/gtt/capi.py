"""ctypes binding of libcbct.so (include/cbct.h).

Device buffers are torch tensors on a ROCm device; only their data pointers
and the current HIP stream cross the boundary.  There is no CPU fallback:
every call goes to the gfx950 library and raises if it is missing.
"""
from __future__ import annotations

import ctypes
import os
import re

_PKG = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.dirname(_PKG)
ROOT = os.path.dirname(LIBDIR)
HEADER = os.path.join(ROOT, "include", "cbct.h")

CBCT_F32 = 0
CBCT_BF16 = 1
CBCT_DOT, CBCT_NRM2, CBCT_ASUM, CBCT_SUM = 0, 1, 2, 3


class CbctError(RuntimeError):
    """Raised for a non-zero status from libcbct (message from cbct_last_error)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class Geom(ctypes.Structure):
    """cbct_geom: volume voxels/extent, detector pixels/extent, SAD, SDD."""
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("lx", ctypes.c_float), ("ly", ctypes.c_float), ("lz", ctypes.c_float),
                ("nu", ctypes.c_int), ("nv", ctypes.c_int),
                ("fov_u", ctypes.c_float), ("fov_v", ctypes.c_float),
                ("sad", ctypes.c_float), ("sdd", ctypes.c_float)]

    @classmethod
    def make(cls, vol_shape_xyz, vol_mm, det_shape_uv, det_mm, sad, sdd):
        nx, ny, nz = vol_shape_xyz
        lx, ly, lz = vol_mm
        nu, nv = det_shape_uv
        fu, fv = det_mm
        return cls(nx, ny, nz, lx, ly, lz, nu, nv, fu, fv, sad, sdd)

    @property
    def vol_shape(self):
        return (self.nz, self.ny, self.nx)

    def proj_shape(self, nviews):
        return (nviews, self.nv, self.nu)

    def __repr__(self):
        return ("Geom(vol=%dx%dx%d @ %gx%gx%g mm, det=%dx%d @ %gx%g mm, SAD=%g, SDD=%g)" %
                (self.nx, self.ny, self.nz, self.lx, self.ly, self.lz, self.nu, self.nv,
                 self.fov_u, self.fov_v, self.sad, self.sdd))


_lib = None


def lib_path():
    return os.path.join(LIBDIR, "libcbct.so")


def declared_symbols():
    """Function names declared in include/cbct.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cbct_[a-z0-9_]+)\s*\(", txt)))


def load():
    """Load libcbct.so; raises if the HIP library has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise ImportError("libcbct.so not built (run __graft_entry__.build() or make -C "
                          "gt-tomography_amd); there is no CPU fallback")
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    I = ctypes.c_int
    F = ctypes.c_float
    S = ctypes.c_size_t
    G = ctypes.POINTER(Geom)
    sig = {
        "cbct_ctx_create": (I, [I, ctypes.POINTER(P)]),
        "cbct_ctx_destroy": (I, [P]),
        "cbct_last_error": (ctypes.c_char_p, [P]),
        "cbct_workspace_bytes": (S, [P]),
        "cbct_ctx_trim": (I, [P]),
        "cbct_ctx_share_workspace": (I, [P, P]),
        "cbct_ctx_reset": (I, [P]),
        "cbct_profile_enable": (I, [P, I]),
        "cbct_profile_read": (I, [P, ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double)]),
        "cbct_views_set": (I, [P, P, P, I]),
        "cbct_views_count": (I, [P]),
        "cbct_fp": (I, [P, G, I, I, P, I, P, F, I, P]),
        "cbct_fp_noalloc": (I, [P, G, I, I, P, I, P, F, I, P]),
        "cbct_bp": (I, [P, G, I, I, P, I, P, I, I, P]),
        "cbct_fp_transpose": (I, [P, G, I, I, P, P, F, I, P]),
        "cbct_fp_engine": (I, [P, I]),
        "cbct_fp_kernel": (ctypes.c_char_p, [P, G, I, I, I, F]),
        "cbct_fp_autotune": (I, [P, G, I, I, I, F, I, P, ctypes.POINTER(ctypes.c_double)]),
        "cbct_fp_autotune_clear": (I, []),
        "cbct_fp_tile": (I, [P, I]),
        "cbct_fp_view_group": (I, [P, I]),
        "cbct_bp_engine": (I, [P, I]),
        "cbct_bp_kernel": (ctypes.c_char_p, [P, G, I, I]),
        "cbct_bp_kz": (I, [P, I]),
        "cbct_storage": (I, [P, I]),
        "cbct_spacecarver": (I, [P, G, I, I, P, P, F, P]),
        "cbct_apply_mask": (I, [P, S, P, P, P]),
        "cbct_cosine_weights": (I, [G, P]),
        "cbct_default_filter": (I, [I, P]),
        "cbct_filter_set": (I, [P, I, P]),
        "cbct_filter_engine": (I, [P, I]),
        "cbct_fdk_prefilter": (I, [P, G, I, I, P, P, I, I, P]),
        "cbct_ramp_filter": (I, [P, I, S, P, P, P]),
        "cbct_offset_weights": (I, [P, G, I, I, P, I, P]),
        "cbct_parker_weights": (I, [P, G, I, I, P, P]),
        "cbct_tv_primal_dual": (I, [P, I, I, I, I, P, P, F, F, F, I, P]),
        "cbct_update_g": (I, [P, S, P, P, P, P, F, F, P]),
        "cbct_tv_update_g": (I, [P, I, I, I, I, P, P, P, P, F, F, F, F, F, P]),
        "cbct_sart_residual": (I, [P, S, P, P, P, P]),
        "cbct_sart_update": (I, [P, S, P, P, P, F, I, P]),
        "cbct_mom_gradient_step": (I, [P, S, P, P, P, F, P]),
        "cbct_mom_momentum": (I, [P, S, P, P, P, F, I, P]),
        "cbct_fill": (I, [P, S, P, F, P]),
        "cbct_scale": (I, [P, S, P, F, P]),
        "cbct_add_scalar": (I, [P, S, P, F, P]),
        "cbct_axpy": (I, [P, S, F, P, P, P]),
        "cbct_mul": (I, [P, S, P, P, P]),
        "cbct_div": (I, [P, S, P, P, P]),
        "cbct_sub": (I, [P, S, P, P, P]),
        "cbct_add": (I, [P, S, P, P, P]),
        "cbct_clamp_min": (I, [P, S, P, F, P]),
        "cbct_reciprocal": (I, [P, S, P, P]),
        "cbct_mul_periodic": (I, [P, S, P, S, P, P]),
        "cbct_to_bf16": (I, [P, S, P, P, P]),
        "cbct_from_bf16": (I, [P, S, P, P, P]),
        "cbct_reduce_f64": (I, [P, I, S, P, P, ctypes.POINTER(ctypes.c_double), P]),
        "cbct_mssim": (I, [P, I, I, I, P, P, F, F, F, I, ctypes.POINTER(ctypes.c_double), P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _torch():
    import torch
    return torch


def _is_number(v):
    return isinstance(v, (int, float)) or (hasattr(v, "ndim") and getattr(v, "ndim") == 0)


def _ptr(t):
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream=None):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def _dtype_code(t):
    import torch
    if t.dtype == torch.float32:
        return CBCT_F32
    if t.dtype == torch.bfloat16:
        return CBCT_BF16
    raise TypeError("unsupported dtype %s (float32 or bfloat16)" % t.dtype)


class Context:
    """Owns a cbct_ctx (device, view table, scratch)."""

    def __init__(self, device=0):
        self.lib = load()
        h = ctypes.c_void_p()
        rc = self.lib.cbct_ctx_create(int(device), ctypes.byref(h))
        if rc != 0:
            raise CbctError(rc, self.lib.cbct_last_error(None).decode())
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            self.lib.cbct_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise CbctError(rc, self.lib.cbct_last_error(self.h).decode())

    @property
    def workspace_bytes(self):
        return self.lib.cbct_workspace_bytes(self.h)

    def trim(self):
        self._check(self.lib.cbct_ctx_trim(self.h))

    def share_workspace(self, owner: "Context"):
        """Use owner's scratch from now on (one host thread at a time)."""
        self._check(self.lib.cbct_ctx_share_workspace(self.h, owner.h))

    def reset(self):
        """Settings back to their defaults; views, filter and scratch kept."""
        self._check(self.lib.cbct_ctx_reset(self.h))

    def profile(self, on=True):
        """Bracket the FP / BP / filter kernels with HIP events on their stream."""
        self._check(self.lib.cbct_profile_enable(self.h, int(on)))

    def profile_read(self):
        """(fp_ms, bp_ms, filter_ms) of the last profiled launches (-1 if none)."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.cbct_profile_read(self.h, ctypes.byref(a), ctypes.byref(b),
                                               ctypes.byref(c)))
        return a.value, b.value, c.value

    # ------------------------------------------------------------- views
    def set_views(self, angles_deg, offsets=None):
        import numpy as np
        a = np.ascontiguousarray(angles_deg, dtype=np.float32).ravel()
        o = None if offsets is None else np.ascontiguousarray(offsets, dtype=np.float32).reshape(-1)
        if o is not None and o.size != 2 * a.size:
            raise ValueError("offsets must hold (u, v) per view")
        self._check(self.lib.cbct_views_set(
            self.h, a.ctypes.data_as(ctypes.c_void_p),
            None if o is None else o.ctypes.data_as(ctypes.c_void_p), int(a.size)))
        self._angles, self._offsets = a, o

    @property
    def nviews(self):
        return self.lib.cbct_views_count(self.h)

    # -------------------------------------------------------- projectors
    def fp(self, g: Geom, vol, proj, view0=0, nviews=None, samples_per_pixel=1.5,
           accumulate=False, stream=None):
        nviews = proj.shape[0] if nviews is None else nviews
        self._check(self.lib.cbct_fp(self.h, ctypes.byref(g), view0, nviews, _ptr(vol),
                                     _dtype_code(vol), _ptr(proj), samples_per_pixel,
                                     int(accumulate), _stream(stream)))
        return proj

    def fp_noalloc(self, g: Geom, vol, proj, view0=0, nviews=None, samples_per_pixel=1.5,
                   accumulate=False, stream=None):
        """cbct_fp_noalloc: fp() with the working copy in the library's fixed
        per-GPU scratch (64 MiB); allocates nothing."""
        nviews = proj.shape[0] if nviews is None else nviews
        self._check(self.lib.cbct_fp_noalloc(self.h, ctypes.byref(g), view0, nviews, _ptr(vol),
                                             _dtype_code(vol), _ptr(proj), samples_per_pixel,
                                             int(accumulate), _stream(stream)))
        return proj

    def fp_kernel(self, g: Geom, view0=0, nviews=None, dtype="f32", samples_per_pixel=1.5):
        """Name of the kernel fp() would launch (for profiler output)."""
        n = self.nviews - view0 if nviews is None else nviews
        code = CBCT_BF16 if dtype in ("bf16", "bfloat16") else CBCT_F32
        return self.lib.cbct_fp_kernel(self.h, ctypes.byref(g), view0, n, code,
                                       samples_per_pixel).decode()

    def bp_kernel(self, g: Geom, view0=0, nviews=None):
        """Name of the kernel bp() would launch (for profiler output)."""
        n = self.nviews - view0 if nviews is None else nviews
        return self.lib.cbct_bp_kernel(self.h, ctypes.byref(g), view0, n).decode()

    def fp_autotune(self, g: Geom, view0=0, nviews=None, dtype="f32", samples_per_pixel=1.5,
                    retune=False, stream=None):
        """Time the column-engine candidates for this geometry (unless it was
        tuned before); engine 1 then uses the fastest.  Returns its time in ms."""
        n = self.nviews - view0 if nviews is None else nviews
        code = CBCT_BF16 if dtype in ("bf16", "bfloat16") else CBCT_F32
        out = ctypes.c_double()
        self._check(self.lib.cbct_fp_autotune(self.h, ctypes.byref(g), view0, n, code, samples_per_pixel,
                                              int(retune), _stream(stream), ctypes.byref(out)))
        return out.value

    def fp_engine(self, engine):
        """0 = one thread per ray, >= 1 = detector-column sharing variants (1 default)."""
        self._check(self.lib.cbct_fp_engine(self.h, int(engine)))

    def fp_tile(self, max_rows_per_lane):
        """Column engine: most detector rows per lane (1..12)."""
        self._check(self.lib.cbct_fp_tile(self.h, int(max_rows_per_lane)))

    def fp_view_group(self, views):
        """Column engine: adjacent views launched together per detector column (1..64)."""
        self._check(self.lib.cbct_fp_view_group(self.h, int(views)))

    def bp_engine(self, engine):
        """2 = default (LDS-staged kernel where its tile holds the blocks'
        detector rectangles, else as 1), 3 = LDS-staged kernel always, 1 =
        quad-texel projection copy, 0 = plain padded copy."""
        self._check(self.lib.cbct_bp_engine(self.h, int(engine)))

    def bp_kz(self, kz):
        """Backprojector z-voxels per thread: 0 = by size (default), 8 or 16."""
        self._check(self.lib.cbct_bp_kz(self.h, int(kz)))

    def storage(self, dtype):
        """Working-copy storage of the projectors for FP32 arrays: 'f32' or 'bf16'."""
        self._check(self.lib.cbct_storage(self.h, CBCT_BF16 if dtype == "bf16" else CBCT_F32))

    def bp(self, g: Geom, proj, vol, view0=0, nviews=None, fbp=False, accumulate=False,
           stream=None):
        nviews = proj.shape[0] if nviews is None else nviews
        self._check(self.lib.cbct_bp(self.h, ctypes.byref(g), view0, nviews, _ptr(proj),
                                     _dtype_code(proj), _ptr(vol), int(fbp), int(accumulate),
                                     _stream(stream)))
        return vol

    def fp_transpose(self, g: Geom, proj, vol, view0=0, nviews=None, samples_per_pixel=1.5,
                     accumulate=False, stream=None):
        nviews = proj.shape[0] if nviews is None else nviews
        self._check(self.lib.cbct_fp_transpose(self.h, ctypes.byref(g), view0, nviews,
                                               _ptr(proj), _ptr(vol), samples_per_pixel,
                                               int(accumulate), _stream(stream)))
        return vol

    def spacecarver(self, g: Geom, proj, mask, limit, view0=0, nviews=None, stream=None):
        nviews = proj.shape[0] if nviews is None else nviews
        self._check(self.lib.cbct_spacecarver(self.h, ctypes.byref(g), view0, nviews, _ptr(proj),
                                              _ptr(mask), float(limit), _stream(stream)))
        return mask

    def apply_mask(self, image, mask, stream=None):
        """image *= mask (uint8 device tensor of the image's size)."""
        if mask.numel() != image.numel():
            raise ValueError("apply_mask: mask and image sizes differ")
        self._check(self.lib.cbct_apply_mask(self.h, image.numel(), _ptr(image), _ptr(mask),
                                             _stream(stream)))
        return image

    # ------------------------------------------------------------- FDK
    def set_filter(self, F):
        import numpy as np
        F = np.ascontiguousarray(F, dtype=np.float32)
        self._check(self.lib.cbct_filter_set(self.h, int(F.size - 1),
                                             F.ctypes.data_as(ctypes.c_void_p)))

    def filter_engine(self, engine):
        self._check(self.lib.cbct_filter_engine(self.h, int(engine)))

    def fdk_prefilter(self, g: Geom, proj_in, proj_out, view0=0, nviews=None, short_scan=False,
                      offset_correct=False, stream=None):
        nviews = proj_in.shape[0] if nviews is None else nviews
        self._check(self.lib.cbct_fdk_prefilter(self.h, ctypes.byref(g), view0, nviews,
                                                _ptr(proj_in), _ptr(proj_out), int(short_scan),
                                                int(offset_correct), _stream(stream)))
        return proj_out

    def ramp_filter(self, nu, data_in, data_out, stream=None):
        nrows = data_in.numel() // nu
        self._check(self.lib.cbct_ramp_filter(self.h, int(nu), nrows, _ptr(data_in),
                                              _ptr(data_out), _stream(stream)))
        return data_out

    def offset_weights(self, g: Geom, proj, view0=0, nviews=None, sqrt_variant=False,
                       stream=None):
        nviews = proj.shape[0] if nviews is None else nviews
        self._check(self.lib.cbct_offset_weights(self.h, ctypes.byref(g), view0, nviews,
                                                 _ptr(proj), int(sqrt_variant), _stream(stream)))
        return proj

    def parker_weights(self, g: Geom, proj, view0=0, nviews=None, stream=None):
        nviews = proj.shape[0] if nviews is None else nviews
        self._check(self.lib.cbct_parker_weights(self.h, ctypes.byref(g), view0, nviews,
                                                 _ptr(proj), _stream(stream)))
        return proj

    # ----------------------------------------------------- TV / solver
    def tv_primal_dual(self, x, out, sigma, alpha=0.0, weight=1.0, accumulate=True,
                       stream=None):
        nz, ny, nx = x.shape[-3:]
        nvol = x.numel() // (nx * ny * nz)
        self._check(self.lib.cbct_tv_primal_dual(self.h, nx, ny, nz, nvol, _ptr(x), _ptr(out),
                                                 sigma, alpha, weight, int(accumulate),
                                                 _stream(stream)))
        return out

    def tv_update_g(self, x_in, x_out, s, precon, sigma, alpha, weight, tau, avg_lambda,
                    stream=None):
        nz, ny, nx = x_in.shape[-3:]
        nvol = x_in.numel() // (nx * ny * nz)
        self._check(self.lib.cbct_tv_update_g(self.h, nx, ny, nz, nvol, _ptr(x_in), _ptr(x_out),
                                              _ptr(s), _ptr(precon), sigma, alpha, weight, tau,
                                              avg_lambda, _stream(stream)))
        return x_out

    def update_g(self, g, x, s, precon, tau, avg_lambda, stream=None):
        self._check(self.lib.cbct_update_g(self.h, x.numel(), _ptr(g), _ptr(x), _ptr(s),
                                           _ptr(precon), tau, avg_lambda, _stream(stream)))
        return x

    def sart_residual(self, r, b, w, stream=None):
        self._check(self.lib.cbct_sart_residual(self.h, r.numel(), _ptr(r), _ptr(b), _ptr(w),
                                                _stream(stream)))

    def sart_update(self, x, img, wx, step, nonneg, stream=None):
        self._check(self.lib.cbct_sart_update(self.h, x.numel(), _ptr(x), _ptr(img), _ptr(wx),
                                              step, int(nonneg), _stream(stream)))

    def mom_gradient_step(self, z, img, precon, step, stream=None):
        self._check(self.lib.cbct_mom_gradient_step(self.h, z.numel(), _ptr(z), _ptr(img),
                                                    _ptr(precon), step, _stream(stream)))

    def mom_momentum(self, x, xold, z, c, nonneg, stream=None):
        self._check(self.lib.cbct_mom_momentum(self.h, x.numel(), _ptr(x), _ptr(xold), _ptr(z),
                                               c, int(nonneg), _stream(stream)))

    def ew(self, name, *args, stream=None):
        """Element-wise primitive by name: fill/scale/add_scalar (x, a),
        axpy (a, x, y), mul/div/sub/add (a, y), clamp_min (x, lo), reciprocal (x)."""
        fn = getattr(self.lib, "cbct_" + name)
        if name in ("fill", "scale", "add_scalar", "clamp_min"):
            x, a = args
            rc = fn(self.h, x.numel(), _ptr(x), float(a), _stream(stream))
        elif name == "axpy":
            a, x, y = args
            rc = fn(self.h, y.numel(), float(a), _ptr(x), _ptr(y), _stream(stream))
        elif name == "reciprocal":
            (x,) = args
            rc = fn(self.h, x.numel(), _ptr(x), _stream(stream))
        else:
            a, y = args
            rc = fn(self.h, y.numel(), _ptr(a), _ptr(y), _stream(stream))
        self._check(rc)

    def to_bf16(self, x, y, stream=None):
        """y (bfloat16) = x (float32) rounded to nearest even."""
        self._check(self.lib.cbct_to_bf16(self.h, x.numel(), _ptr(x), _ptr(y), _stream(stream)))
        return y

    def from_bf16(self, x, y, stream=None):
        """y (float32) = x (bfloat16)."""
        self._check(self.lib.cbct_from_bf16(self.h, x.numel(), _ptr(x), _ptr(y), _stream(stream)))
        return y

    def reduce(self, op, a, b=None, stream=None):
        out = ctypes.c_double()
        self._check(self.lib.cbct_reduce_f64(self.h, int(op), a.numel(), _ptr(a), _ptr(b),
                                             ctypes.byref(out), _stream(stream)))
        return out.value

    def dot(self, a, b, stream=None):
        return self.reduce(CBCT_DOT, a, b, stream)

    def nrm2(self, a, stream=None):
        return self.reduce(CBCT_NRM2, a, None, stream)

    def asum(self, a, stream=None):
        return self.reduce(CBCT_ASUM, a, None, stream)

    def mssim(self, image, reference, sigma=1.5, scales=3, stream=None):
        """Multi-scale SSIM of a device volume (nz, ny, nx) against a reference
        volume of the same shape (xray/mssim.cpp:77-121).  sigma: a number or
        (sx, sy, sz) in voxels."""
        if image.shape != reference.shape or image.dim() != 3:
            raise ValueError("mssim: image and reference must be 3-D volumes of one shape")
        for t in (image, reference):
            if t.dtype != _torch().float32 or not t.is_contiguous() or not t.is_cuda:
                raise ValueError("mssim: contiguous float32 device volumes expected")
        sx, sy, sz = (sigma,) * 3 if _is_number(sigma) else tuple(sigma)
        nz, ny, nx = image.shape
        out = ctypes.c_double()
        self._check(self.lib.cbct_mssim(self.h, nx, ny, nz, _ptr(image), _ptr(reference), sx, sy, sz,
                                        int(scales), ctypes.byref(out), _stream(stream)))
        return out.value


def fp_autotune_clear():
    load().cbct_fp_autotune_clear()


def cosine_weights(g: Geom):
    import numpy as np
    w = np.zeros((g.nv, g.nu), np.float32)
    rc = load().cbct_cosine_weights(ctypes.byref(g), w.ctypes.data_as(ctypes.c_void_p))
    if rc:
        raise CbctError(rc, "cbct_cosine_weights failed")
    return w


def default_filter(nu):
    import numpy as np
    F = np.zeros(nu + 1, np.float32)
    rc = load().cbct_default_filter(int(nu), F.ctypes.data_as(ctypes.c_void_p))
    if rc:
        raise CbctError(rc, "cbct_default_filter failed")
    return F
