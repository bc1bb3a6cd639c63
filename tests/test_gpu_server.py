"""GPU tests of the reconstruction server (gtt_server_*, through libgtt).

* Served results equal alone-run results (the plain gtt_fdk_reconstruct_3d /
  gtt_os_sart entry points, one at a time) bitwise, whatever the number of
  lanes and the mix of concurrent requests.
* An OS-SART request through the server meets the solver parity bound
  against the FP64 oracle.
* Admission by device footprint: a request larger than the budget is
  refused at submit; with a budget of one request, requests run one at a
  time in arrival order.
"""
import dataclasses

import numpy as np
import pytest

import oracle as O
from gtt import capi
from gtt import operators as ops
from gtt import serving as S
from helpers import rel_l2, to_oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def print_case(*a):
    print("CASE", *a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def small_requests(ctx):
    specs = S.uneven_scans(6, seed=2024, p_fdk=0.5, scale=0.15)
    return [S.make_request(s, ctx, iterations=3) for s in specs]


def alone(req):
    if req.kind == S.REQ_FDK:
        vol, _ = ops.fdk_reconstruct_3d(req.geom, req.angles, None, req.projections)
        return vol
    return ops.os_sart(req.geom, req.angles, None, req.projections, req.subsets, req.iterations,
                       beta=req.beta, gamma=req.gamma, nonneg=req.nonneg)


@pytest.mark.parametrize("lanes", [1, 3])
def test_served_equals_alone(small_requests, lanes):
    print_case("server served == alone, lanes", lanes,
               [(r.kind, r.geom.nx, len(r.angles), r.geom.nu) for r in small_requests])
    ref = [alone(r) for r in small_requests]
    for r in small_requests:
        r.volume = None
    with S.Server(lanes=lanes) as srv:
        res, wall = S.run_stream(srv, small_requests)
    for r, out, want in zip(small_requests, res, ref):
        assert np.all(np.isfinite(out["volume"]))
        assert np.array_equal(out["volume"], want), (r.kind, r.geom.nx, rel_l2(out["volume"], want))
        assert out["latency_ms"] >= out["run_ms"] >= 0.0


def test_served_os_sart_matches_oracle(ctx):
    spec = S.ScanSpec(n=20, nviews=30, nu=40, nv=30, subsets=5, kind=S.REQ_OS_SART, seed=77)
    print_case("server os-sart vs oracle", spec)
    req = S.make_request(spec, ctx, iterations=3)
    g = req.geom
    op = O.CBSubsetOperator(spec.subsets)
    pb = op.setup(to_oracle(g), req.angles, np.zeros((spec.nviews, 2), np.float32),
                  req.projections, O.Binning.default_3d(spec.nviews))
    ref = O.os_sart(op, pb, 3, beta=1.0, gamma=0.0, nonneg=True)
    with S.Server(lanes=2) as srv:
        out = srv.wait(srv.submit(req))
    err = rel_l2(out["volume"], ref)
    print("server os-sart rel-L2", err)
    assert err <= 1e-4


def test_admission_by_footprint(small_requests):
    reqs = small_requests[:4]
    fps = [r.footprint() for r in reqs]
    print_case("server admission", fps)
    with S.Server(lanes=2, hbm_budget_bytes=min(fps) - 1) as srv:
        with pytest.raises(RuntimeError, match="budget"):
            srv.submit(reqs[int(np.argmin(fps))])
    for r in reqs:
        r.volume = None
    with S.Server(lanes=3, hbm_budget_bytes=max(fps)) as srv:
        res, _ = S.run_stream(srv, reqs)
    # one at a time, in arrival order: each starts after the previous ended
    for a, b in zip(res, res[1:]):
        assert b["queue_ms"] >= a["queue_ms"] + a["run_ms"] - 1.0


def test_failed_request_reports_and_server_continues(small_requests):
    bad = S.Request(kind=S.REQ_OS_SART, geom=small_requests[0].geom, angles=small_requests[0].angles,
                    projections=small_requests[0].projections, subsets=0, iterations=1)
    good = small_requests[0]
    good.volume = None
    print_case("server error path")
    with S.Server(lanes=2) as srv:
        ib = srv.submit(bad)
        ig = srv.submit(good)
        with pytest.raises(RuntimeError, match="subsets"):
            srv.wait(ib)
        assert np.all(np.isfinite(srv.wait(ig)["volume"]))


def test_wait_with_time_limit(small_requests):
    """gtt_server_wait_for: a request not finished within the limit reports
    GTT_WAIT_TIMEOUT and stays pending; waited for again it returns the same
    result as gtt_server_wait."""
    req = next(r for r in small_requests if r.kind == S.REQ_OS_SART)
    print_case("server wait with a time limit", req.geom.nx, len(req.angles))
    ref = alone(req)
    req.volume = None
    with S.Server(lanes=1) as srv:
        rid = srv.submit(req)
        with pytest.raises(TimeoutError, match="not finished"):
            srv.wait(rid, timeout=0.0)
        out = srv.wait(rid, timeout=120.0)
        assert np.array_equal(out["volume"], ref)
        with pytest.raises(RuntimeError, match="unknown request"):
            srv.wait(rid, timeout=1.0)


def test_served_bf16_storage_equals_alone(small_requests):
    """A request with BF16 working copies, served beside FP32 requests,
    equals the same reconstruction run alone with BF16 working copies
    (bitwise), and stays within the BF16 bound of the FP32 result."""
    reqs = [r for r in small_requests if r.kind == S.REQ_OS_SART][:2] + \
           [r for r in small_requests if r.kind == S.REQ_FDK][:1]
    print_case("server bf16 storage", [(r.kind, r.geom.nx) for r in reqs])
    ref32 = [alone(r) for r in reqs]
    ops.set_default_storage("bf16")
    try:
        ref16 = [alone(r) for r in reqs]
    finally:
        ops.set_default_storage("f32")
    mixed = []
    for r in reqs:
        r.volume = None
        r.storage = "bf16"
        mixed.append(r)
    # an FP32 request served beside them (its own copy: the others are BF16)
    extra = dataclasses.replace(small_requests[0], volume=None, storage="f32")
    try:
        with S.Server(lanes=3) as srv:
            res, _ = S.run_stream(srv, mixed + [extra])
    finally:
        for r in reqs:
            r.storage = "f32"
    for r, out, w16, w32 in zip(mixed, res, ref16, ref32):
        assert np.array_equal(out["volume"], w16), r.kind
        err = rel_l2(out["volume"], w32)
        print("bf16 working copies vs fp32, kind", r.kind, "rel-L2", err)
        assert err <= 2e-2
    assert np.array_equal(res[-1]["volume"], alone(extra))


def test_served_autotuned_equals_alone(small_requests):
    """With autotuning on, every OS-SART request tunes the forward projector
    for its geometry first; the results are still those of the alone runs."""
    print_case("server autotune", [(r.kind, r.geom.nx) for r in small_requests])
    capi.fp_autotune_clear()
    ref = [alone(r) for r in small_requests]
    for r in small_requests:
        r.volume = None
    try:
        with S.Server(lanes=2, autotune=True) as srv:
            res, _ = S.run_stream(srv, small_requests)
    finally:
        capi.fp_autotune_clear()
    for r, out, want in zip(small_requests, res, ref):
        assert np.array_equal(out["volume"], want), (r.kind, r.geom.nx)


@pytest.mark.parametrize("kind", [S.REQ_OS_SART, S.REQ_FDK])
def test_footprint_covers_measured_peak(ctx, kind):
    """The admission estimate covers the device memory a request really
    takes: one request (12 subsets for OS-SART) on a fresh one-lane server,
    device free memory polled while it runs; the peak use (free memory
    before minus the lowest seen) stays within gtt_request_footprint."""
    import threading
    spec = S.ScanSpec(n=200, nviews=240, nu=400, nv=300, subsets=12, kind=kind, seed=77)
    req = S.make_request(spec, ctx, iterations=2)
    est = req.footprint()
    print_case("footprint vs measured peak, kind", kind, "estimate", est)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    low = [free0]
    stop = threading.Event()

    def poll():
        while not stop.is_set():
            low[0] = min(low[0], torch.cuda.mem_get_info()[0])

    t = threading.Thread(target=poll)
    t.start()
    try:
        with S.Server(lanes=1) as srv:
            rid = srv.submit(req)
            out = srv.wait(rid)
    finally:
        stop.set()
        t.join()
    peak = free0 - low[0]
    print("kind", kind, "peak", peak, "estimate", est, "ratio %.2f" % (peak / est))
    assert np.all(np.isfinite(out["volume"]))
    assert 0 < peak <= est


def test_served_with_graphs_equals_alone(small_requests):
    """OS-SART requests served with their iterations replayed as captured HIP
    graphs (gtt_server_set_graphs) give the alone results bit for bit."""
    print_case("server graphs", [(r.kind, r.geom.nx) for r in small_requests])
    ref = [alone(r) for r in small_requests]
    for r in small_requests:
        r.volume = None
    with S.Server(lanes=3, graphs=True) as srv:
        res, _ = S.run_stream(srv, small_requests)
    for r, out, want in zip(small_requests, res, ref):
        assert np.array_equal(out["volume"], want), (r.kind, r.geom.nx)


def test_graph_captures_beside_other_lanes_setup(ctx):
    """Lanes that capture OS-SART iterations while other lanes set up their
    operators (view-table and filter uploads) and run FDK requests: the
    uploads go through the library's own non-blocking stream, never the
    legacy default stream, which a capture on another stream forbids.
    Results equal the alone runs bit for bit."""
    specs = S.uneven_scans(16, seed=77, p_fdk=0.4, scale=0.35)
    reqs = [S.make_request(s, ctx, iterations=4) for s in specs]
    print_case("graphs beside setups", [(r.kind, r.geom.nx, len(r.angles)) for r in reqs])
    ref = [alone(r) for r in reqs]
    for r in reqs:
        r.volume = None
    with S.Server(lanes=4, graphs=True) as srv:
        res, _ = S.run_stream(srv, reqs)
    for r, out, want in zip(reqs, res, ref):
        assert np.array_equal(out["volume"], want), (r.kind, r.geom.nx)
