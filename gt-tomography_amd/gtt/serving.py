"""Reconstruction serving on one GPU: ctypes binding of the gtt_server_* C
API (include/gtt_c.h), the uneven-scan workload, and the arrival-stream
driver used for latency measurements.

Workload (uneven scans): per scan a cube of edge N ~ U{96..320} voxels at
1 mm, views ~ U{90..540} over 360 degrees, a detector of width
nu ~ U{256..768} pixels with 4:3 aspect whose width in mm is 1.56 x the
volume width, OS-SART with subsets ~ U{5..12} and 10 iterations.  Mixed
request streams draw FDK with probability 0.6 and 10-iteration OS-SART with
probability 0.4 at the same sizes, arriving as a Poisson process.
Projections are the forward projection (this library's FP) of the seeded
ellipsoid phantom.
"""
from __future__ import annotations

import ctypes
import dataclasses
import time

import numpy as np

from . import capi
from . import operators as ops
from . import phantom as PH

REQ_FDK = 0
REQ_OS_SART = 1
GTT_WAIT_TIMEOUT = 2


class _Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("geom", capi.Geom), ("nviews", ctypes.c_int),
                ("angles_deg", ctypes.c_void_p), ("offsets", ctypes.c_void_p),
                ("projections", ctypes.c_void_p), ("filter", ctypes.c_void_p),
                ("samples_per_pixel", ctypes.c_float), ("subsets", ctypes.c_int),
                ("iterations", ctypes.c_int), ("beta", ctypes.c_float), ("gamma", ctypes.c_float),
                ("nonneg", ctypes.c_int), ("volume_out", ctypes.c_void_p),
                ("storage", ctypes.c_int)]


class _Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("queue_ms", ctypes.c_double),
                ("run_ms", ctypes.c_double), ("latency_ms", ctypes.c_double),
                ("scale", ctypes.c_double), ("message", ctypes.c_char * 256)]


_bound = False


def _lib():
    global _bound
    L = ops.load()
    if not _bound:
        P, I = ctypes.c_void_p, ctypes.c_int
        L.gtt_server_create.argtypes = [I, I, ctypes.c_size_t, ctypes.POINTER(P)]
        L.gtt_server_destroy.argtypes = [P]
        L.gtt_server_submit.argtypes = [P, ctypes.POINTER(_Request), ctypes.POINTER(ctypes.c_longlong)]
        L.gtt_server_wait.argtypes = [P, ctypes.c_longlong, ctypes.POINTER(_Result)]
        L.gtt_server_wait_for.argtypes = [P, ctypes.c_longlong, ctypes.c_longlong, ctypes.POINTER(_Result)]
        L.gtt_server_wait_for.restype = I
        L.gtt_server_last_error.argtypes = [P]
        L.gtt_server_last_error.restype = ctypes.c_char_p
        L.gtt_server_set_autotune.argtypes = [P, I]
        L.gtt_server_set_autotune.restype = I
        L.gtt_server_set_graphs.argtypes = [P, I]
        L.gtt_server_set_graphs.restype = I
        L.gtt_request_footprint.argtypes = [ctypes.POINTER(_Request)]
        L.gtt_request_footprint.restype = ctypes.c_size_t
        for n in ("gtt_server_create", "gtt_server_destroy", "gtt_server_submit", "gtt_server_wait"):
            getattr(L, n).restype = I
        _bound = True
    return L


@dataclasses.dataclass
class Request:
    """One reconstruction: FDK or OS-SART of host projections (nviews, nv, nu)."""
    kind: int
    geom: capi.Geom
    angles: np.ndarray
    projections: np.ndarray
    offsets: np.ndarray | None = None
    filter: np.ndarray | None = None
    samples_per_pixel: float = 1.5
    subsets: int = 10
    iterations: int = 10
    beta: float = 1.0
    gamma: float = 0.0
    nonneg: bool = True
    storage: str = "f32"                 # projectors' working copies: "f32" or "bf16"
    volume: np.ndarray | None = None     # filled on completion

    def updates(self):
        """Voxel-view updates (one per voxel and view for each FP and BP)."""
        g = self.geom
        vv = g.nx * g.ny * g.nz * len(self.angles)
        if self.kind == REQ_FDK:
            return 2 * vv                       # FBP + rescale FP
        return 2 * vv * (self.iterations + 1)   # normalisers + one FP/BP sweep per iteration

    def _struct(self):
        self.angles = np.ascontiguousarray(self.angles, np.float32)
        self.projections = np.ascontiguousarray(self.projections, np.float32)
        if self.offsets is not None:
            self.offsets = np.ascontiguousarray(self.offsets, np.float32)
        if self.filter is not None:
            self.filter = np.ascontiguousarray(self.filter, np.float32)
        g = self.geom
        if self.volume is None:
            self.volume = np.empty((g.nz, g.ny, g.nx), np.float32)
        p = lambda a: None if a is None else a.ctypes.data
        return _Request(self.kind, g, len(self.angles), p(self.angles), p(self.offsets),
                        p(self.projections), p(self.filter), self.samples_per_pixel, self.subsets,
                        self.iterations, self.beta, self.gamma, int(self.nonneg), p(self.volume),
                        1 if self.storage == "bf16" else 0)

    def footprint(self):
        return int(_lib().gtt_request_footprint(ctypes.byref(self._struct())))


class Server:
    """gtt_server: `lanes` worker lanes (own HIP stream each) on one device."""

    def __init__(self, lanes=4, hbm_budget_bytes=0, device=0, autotune=False, graphs=False):
        L = _lib()
        h = ctypes.c_void_p()
        if L.gtt_server_create(device, lanes, hbm_budget_bytes, ctypes.byref(h)) != 0:
            raise RuntimeError("gtt_server_create failed")
        self.h = h
        if autotune:
            L.gtt_server_set_autotune(h, 1)
        if graphs:
            L.gtt_server_set_graphs(h, 1)
        self.lanes = lanes
        self._live = {}

    def submit(self, req: Request):
        s = req._struct()
        rid = ctypes.c_longlong()
        if _lib().gtt_server_submit(self.h, ctypes.byref(s), ctypes.byref(rid)) != 0:
            raise RuntimeError(_lib().gtt_server_last_error(self.h).decode())
        self._live[rid.value] = req      # keeps the buffers alive until wait()
        return rid.value

    def wait(self, rid, timeout=None):
        """Result of request rid; with timeout (seconds), TimeoutError when it
        has not finished by then (it stays pending and may be waited for
        again)."""
        r = _Result()
        if timeout is None:
            rc = _lib().gtt_server_wait(self.h, rid, ctypes.byref(r))
        else:
            rc = _lib().gtt_server_wait_for(self.h, rid, int(timeout * 1000), ctypes.byref(r))
            if rc == GTT_WAIT_TIMEOUT:
                raise TimeoutError(_lib().gtt_server_last_error(self.h).decode())
        req = self._live.pop(rid, None)
        if rc != 0:
            raise RuntimeError(r.message.decode() or _lib().gtt_server_last_error(self.h).decode())
        return {"queue_ms": r.queue_ms, "run_ms": r.run_ms, "latency_ms": r.latency_ms,
                "scale": r.scale, "volume": None if req is None else req.volume}

    def close(self):
        if self.h:
            _lib().gtt_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# ------------------------------------------------------------------ workload
@dataclasses.dataclass
class ScanSpec:
    n: int            # cube edge (voxels, 1 mm)
    nviews: int
    nu: int
    nv: int
    subsets: int
    kind: int
    seed: int

    def geom(self):
        L = float(self.n)
        fu = 1.56 * L
        return capi.Geom.make((self.n,) * 3, (L, L, L), (self.nu, self.nv),
                              (fu, fu * self.nv / self.nu), 1000.0, 1500.0)

    def angles(self):
        return (np.arange(self.nviews, dtype=np.float32) * np.float32(360.0 / self.nviews)).astype(np.float32)


def uneven_scans(count, seed=1003, p_fdk=0.0, scale=1.0):
    """Scan sizes of the uneven-batch distribution (scale < 1 shrinks edges,
    views and detectors for small tests)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(96, 321))
        nviews = int(rng.integers(90, 541))
        nu = int(rng.integers(256, 769))
        subsets = int(rng.integers(5, 13))
        kind = REQ_FDK if rng.random() < p_fdk else REQ_OS_SART
        if scale != 1.0:
            n, nviews, nu = (max(8, int(round(n * scale))), max(subsets, int(round(nviews * scale))),
                             max(8, int(round(nu * scale))))
        nv = max(4, int(round(nu * 0.75)))
        out.append(ScanSpec(n, nviews, nu, nv, subsets, kind, seed * 1000 + i))
    return out


def make_request(spec: ScanSpec, ctx: capi.Context, iterations=10):
    """Host projections of the seeded phantom (FP on the GPU), as a Request."""
    import torch
    g = spec.geom()
    angles = spec.angles()
    L = float(spec.n)
    vol = PH.voxelize(PH.make_phantom((L, L, L), spec.seed), (spec.n,) * 3, (L, L, L), device="cuda")
    ctx.set_views(angles)
    proj = torch.empty(g.proj_shape(len(angles)), device="cuda")
    ctx.fp(g, vol, proj)
    host = proj.cpu().numpy()
    return Request(kind=spec.kind, geom=g, angles=angles, projections=host,
                   subsets=spec.subsets, iterations=iterations)


def poisson_arrivals(count, rate_per_s, seed=7):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, count))


def run_stream(server: Server, requests, arrivals_s=None, timeout=None):
    """Submit each request at its arrival time (all at once if None), wait
    for all (each wait at most `timeout` seconds when given); returns
    per-request results and the wall time."""
    n = len(requests)
    arrivals_s = np.zeros(n) if arrivals_s is None else np.asarray(arrivals_s)
    t0 = time.perf_counter()
    ids = []
    for req, ta in zip(requests, arrivals_s):
        dt = t0 + ta - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        ids.append(server.submit(req))
    res = [server.wait(i, timeout) for i in ids]
    wall = time.perf_counter() - t0
    return res, wall


def summarize(requests, res, wall):
    lat = np.array([r["latency_ms"] for r in res])
    upd = float(sum(q.updates() for q in requests))
    return {"requests": len(requests), "p50_ms": round(float(np.percentile(lat, 50)), 2),
            "p99_ms": round(float(np.percentile(lat, 99)), 2),
            "mean_ms": round(float(lat.mean()), 2), "wall_s": round(wall, 3),
            "gups": round(upd / wall / 1e9, 2)}
