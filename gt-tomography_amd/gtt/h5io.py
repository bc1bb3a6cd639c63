"""ctypes binding of libgtt_h5.so (include/gtt_h5.h): the reference's HDF5
acquisition, binning and image files.  The HDF5 1.10 shared libraries are
loaded first from H5LIB (default /opt/conda/lib) so that libgtt_h5 resolves
them without a run path into that directory."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import capi

H5LIB = os.environ.get("GTT_HDF5_LIBDIR", "/opt/conda/lib")
_lib = None


def lib_path():
    return os.path.join(capi.LIBDIR, "libgtt_h5.so")


def available():
    return os.path.exists(lib_path()) and os.path.exists(os.path.join(H5LIB, "libhdf5_hl.so"))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(lib_path()):
        raise ImportError("libgtt_h5.so not built (HDF5 headers missing at build time)")
    for name in ("libhdf5.so", "libhdf5_hl.so"):
        ctypes.CDLL(os.path.join(H5LIB, name), mode=ctypes.RTLD_GLOBAL)
    L = ctypes.CDLL(lib_path())
    P, I = ctypes.c_void_p, ctypes.c_int
    IP, FP = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)
    L.gtt_h5_last_error.restype = ctypes.c_char_p
    L.gtt_h5_acquisition_info.argtypes = [ctypes.c_char_p, IP, IP, IP, IP]
    L.gtt_h5_acquisition_read.argtypes = [ctypes.c_char_p, FP, FP, P, P, P, P, P]
    L.gtt_h5_acquisition_write.argtypes = [ctypes.c_char_p, ctypes.c_float, ctypes.c_float, P, I, I,
                                           I, P, P, P]
    L.gtt_h5_binning_info.argtypes = [ctypes.c_char_p, IP, IP]
    L.gtt_h5_binning_read.argtypes = [ctypes.c_char_p, P, P]
    L.gtt_h5_binning_write.argtypes = [ctypes.c_char_p, I, P, P]
    L.gtt_h5_image_write.argtypes = [ctypes.c_char_p, P, I, I, I, P, P, ctypes.c_char_p,
                                     ctypes.c_uint]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise RuntimeError(load().gtt_h5_last_error().decode())


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def load_acquisition(path):
    """Returns dict(projections (nviews, nv, nu), angles, offsets (nviews, 2),
    sad, sdd, fov (2,), airscan or None)."""
    L = load()
    b = os.fsencode(path)
    nu, nv, nviews = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    ad = (ctypes.c_int * 3)()
    _check(L.gtt_h5_acquisition_info(b, ctypes.byref(nu), ctypes.byref(nv), ctypes.byref(nviews), ad))
    proj = np.empty((nviews.value, nv.value, nu.value), np.float32)
    angles = np.empty(nviews.value, np.float32)
    offsets = np.empty((nviews.value, 2), np.float32)
    fov = np.empty(2, np.float32)
    air = np.empty((ad[2], ad[1], ad[0]), np.float32) if ad[0] else None
    sad, sdd = ctypes.c_float(), ctypes.c_float()
    _check(L.gtt_h5_acquisition_read(b, ctypes.byref(sad), ctypes.byref(sdd), _p(fov), _p(angles),
                                     _p(offsets), _p(proj), _p(air)))
    return {"projections": proj, "angles": angles, "offsets": offsets, "sad": sad.value,
            "sdd": sdd.value, "fov": fov, "airscan": air}


def save_acquisition(path, projections, angles, sad, sdd, fov, offsets=None):
    p = np.ascontiguousarray(projections, np.float32)
    a = np.ascontiguousarray(angles, np.float32)
    o = None if offsets is None else np.ascontiguousarray(offsets, np.float32)
    f = np.ascontiguousarray(fov, np.float32)
    nviews, nv, nu = p.shape
    _check(load().gtt_h5_acquisition_write(os.fsencode(path), sad, sdd, _p(f), nu, nv, nviews,
                                           _p(a), _p(o), _p(p)))


def load_binning(path):
    L = load()
    b = os.fsencode(path)
    nb, tot = ctypes.c_int(), ctypes.c_int()
    _check(L.gtt_h5_binning_info(b, ctypes.byref(nb), ctypes.byref(tot)))
    sizes = np.empty(nb.value, np.int32)
    idx = np.empty(tot.value, np.uint32)
    _check(L.gtt_h5_binning_read(b, _p(sizes), _p(idx)))
    out, k = [], 0
    for s in sizes:
        out.append(idx[k:k + s].copy())
        k += s
    return out


def save_binning(path, bins):
    sizes = np.ascontiguousarray([len(b) for b in bins], np.int32)
    idx = np.ascontiguousarray(np.concatenate([np.asarray(b, np.uint32) for b in bins])
                               if bins else np.zeros(0, np.uint32), np.uint32)
    _check(load().gtt_h5_binning_write(os.fsencode(path), len(bins), _p(sizes), _p(idx)))


def save_image(path, image, dims_mm, origin=(0.0, 0.0, 0.0), arguments="", iterations=0):
    """image: (nz, ny, nx) float32; dims_mm, origin: (x, y, z)."""
    im = np.ascontiguousarray(image, np.float32)
    nz, ny, nx = im.shape
    d = np.ascontiguousarray(dims_mm, np.float32)
    o = np.ascontiguousarray(origin, np.float32)
    _check(load().gtt_h5_image_write(os.fsencode(path), _p(im), nx, ny, nz, _p(d), _p(o),
                                     arguments.encode(), int(iterations)))
