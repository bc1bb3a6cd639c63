"""CPU checks of the C ABI: the library loads, exports every symbol the
header declares, and the host-only entry points work (no GPU calls)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from gtt import capi
from helpers import geom, to_oracle


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    declared = capi.declared_symbols()
    assert len(declared) >= 40
    out = subprocess.run(["nm", "-D", "--defined-only", capi.lib_path()], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    for s in declared:
        assert hasattr(lib, s)


def test_ctx_create_without_gpu_reports_error():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("GPU present")
    except Exception:
        pass
    lib = capi.load()
    h = ctypes.c_void_p()
    rc = lib.cbct_ctx_create(0, ctypes.byref(h))
    assert rc != 0
    assert b"device" in lib.cbct_last_error(None)


def test_host_cosine_weights_match_oracle():
    g = geom(8, 8, 8, 40, 30, fov=(60.0, 45.0))
    w = capi.cosine_weights(g)
    ref = O.cosine_weights(to_oracle(g))
    assert np.abs(w - ref).max() <= 1e-7


def test_host_default_filter_matches_oracle():
    for nu in (16, 200, 1024):
        F = capi.default_filter(nu)
        ref = O.guo_filter(nu)
        assert np.abs(F - ref).max() <= 1e-6 * np.abs(ref).max()


def test_declared_entry_points_cite_reference():
    """Every entry point in the header says which reference interface it
    replaces, or that it has no counterpart."""
    txt = open(capi.HEADER).read()
    blocks = txt.split("int cbct_")
    assert "conebeam_forwards_projection" in txt
    assert "conebeam_backwards_projection" in txt
    assert "cuTVPrimalDualOperator" in txt
    assert "osMOMSolverD" in txt and "osSARTSolver" in txt


def test_libgtt_exports_every_declared_symbol():
    import re

    from gtt import operators as ops
    hdr = os.path.join(os.path.dirname(capi.HEADER), "gtt_c.h")
    declared = sorted(set(re.findall(r"\b(gtt_\w+)\s*\(", open(hdr).read())))
    assert "gtt_server_submit" in declared and len(declared) >= 15
    out = subprocess.run(["nm", "-D", "--defined-only", ops.lib_path()], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in declared if s not in exported]
    assert not missing, missing


def test_server_wait_for_argument_errors():
    """gtt_server_wait_for refuses a null server (host-only check; the time
    limit itself is tested on the GPU)."""
    from gtt import serving as S
    L = S._lib()
    assert L.gtt_server_wait_for(None, 1, 10, None) == -1
    assert L.gtt_server_wait(None, 1, None) == -1
    assert S.GTT_WAIT_TIMEOUT == 2


def test_request_footprint_and_server_without_gpu():
    """Host-only parts of the server API: the footprint estimate (volume- and
    padded-projection-sized arrays, by request kind and subset count, plus
    slack) and a clean error from gtt_server_create when no device can be
    selected."""
    from gtt import serving as S
    spec = S.ScanSpec(n=40, nviews=30, nu=64, nv=48, subsets=5, kind=S.REQ_OS_SART, seed=1)
    g = spec.geom()
    req = S.Request(kind=S.REQ_OS_SART, geom=g, angles=spec.angles(),
                    projections=np.zeros(g.proj_shape(30), np.float32))
    nvox, nproj = 40 ** 3, (64 + 3) * (48 + 3) * 30
    blocks = 3 * 3 * 5                                  # 16 x 16 x 8 voxel blocks
    # OS-SART: S + 2 volumes, the shared padded volume copy (<= 1.3), four
    # sinograms, one subset's quad-texel copy and the backprojector's
    # rectangle table for one subset; FDK: 2.3 volumes, 8 sinograms, the
    # rectangle table for all views
    for S_ in (1, 5, 12):
        req.subsets = S_
        want = (4 * (nvox * (10 * S_ + 33) // 10 + 4 * nproj + 4 * (-(-nproj // S_)))
                + 8 * blocks * (-(-30 // S_)) + (64 << 20))
        assert req.footprint() == want
    req.kind = S.REQ_FDK
    assert req.footprint() == 4 * (nvox * 23 // 10 + 8 * nproj) + 8 * blocks * 30 + (64 << 20)
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("GPU present")
    except Exception:
        pass
    with pytest.raises(RuntimeError, match="gtt_server_create"):
        S.Server(lanes=2)


def test_real_files_round_trip(tmp_path):
    """.real arrays (write_nd_array; python/readReal.py:41-76): uint32 ndims,
    the sizes first-index-fastest, float32 data.  libgtt's writer is read by
    a numpy restatement of the reference's reader, and libgtt's reader reads
    a file laid out as the reference's saveReal writes it."""
    from gtt import operators as ops
    a = np.random.default_rng(3).random((4, 5, 6)).astype(np.float32)
    path = str(tmp_path / "a.real")
    ops.write_real(path, a)
    raw = np.fromfile(path, dtype=np.uint8)
    hdr = raw[:16].view(np.uint32)
    assert list(hdr) == [3, 6, 5, 4]                      # ndims, nx, ny, nz
    assert np.array_equal(raw[16:].view(np.float32).reshape(4, 5, 6), a)
    b = np.random.default_rng(4).random((2, 3, 7, 5)).astype(np.float32)
    p2 = str(tmp_path / "b.real")
    with open(p2, "wb") as f:                              # saveReal's layout
        np.array(b.ndim, np.uint32).tofile(f)
        np.array(b.shape[::-1], np.uint32).tofile(f)
        b.tofile(f)
    assert np.array_equal(ops.read_real(p2), b)
    with open(p2, "r+b") as f:                             # truncated data
        f.truncate(4 + 4 * b.ndim + 4 * (b.size - 1))
    with pytest.raises(RuntimeError, match="truncated"):
        ops.read_real(p2)
