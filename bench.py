#!/usr/bin/env python3
"""fwd+back GUPS of the gfx950 cone-beam projector pair on one MI355X.

Workload (default): 512^3 volume at 0.5 mm, 720 views over 360 degrees,
1024x768 flat detector at 0.39 mm, SAD 1000 mm, SDD 1500 mm, seeded
ellipsoid phantom (synthetic); the projectors' working copies of the volume
and the sinogram in BF16, arithmetic in FP32 (--storage f32 for FP32
copies).  One step = one forward projection of the volume onto all 720
views + one backprojection of all 720 views, inputs resident in HBM.
GUPS = 2 * Nvox * Nviews / t_step / 1e9.  The plain FP32 path is timed in
the same run.

Besides the headline step the line carries:
  roofline      the dominant kernel against the resource that bounds it
                (issued VALU lane-instructions for the forward projector,
                measured HBM bytes for the backprojector, MFMA flops for the
                ramp filter), from the live kernel time (HIP events on the
                kernel's stream) and the counters committed under profiles/;
  cpu_baseline  the FP32 OpenMP restatement and the FP64 oracle timed on
                this host (bounded samples), and each stage's error against
                the FP64 oracle on a deterministic subsample;
  workloads     the 128^3 FDK and 256^3 OS-SART / osMOMSolverD workloads;
  latency       the reconstruction batch (64 uneven scans) run plain, its
                served run with --served-batch, and the served-vs-plain
                Poisson request stream.
"""
import argparse
import hashlib
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gt-tomography_amd"))

# MI355X published peaks
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
VALU_PEAK_TOPS = FP32_PEAK_TFLOPS / 2.0   # lane-instructions/s: one FMA (2 flops) per lane
FP16_MFMA_PEAK_TFLOPS = 2500.0

FP_SOURCES = ("gt-tomography_amd/csrc/fp.hip", "gt-tomography_amd/csrc/cbct_internal.h")
BP_SOURCES = ("gt-tomography_amd/csrc/bp.hip", "gt-tomography_amd/csrc/cbct_internal.h")
FILTER_SOURCES = ("gt-tomography_amd/csrc/filter.hip", "gt-tomography_amd/csrc/cbct_internal.h")


def sources_hash(paths):
    h = hashlib.sha256()
    for p in paths:
        with open(os.path.join(ROOT, p), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def footprint_bytes(kind, nvox, nviews, nrays, nsamp, s_vol, s_proj):
    """Footprint-convention bytes of one launch (an effective-bandwidth
    figure: on-chip reuse lets it exceed HBM peak).
    FP: every sample reads its 8-tap footprint, every ray writes one value.
    BP: every voxel-view reads its 4-tap footprint, every voxel is written."""
    if kind == "fp":
        return nrays * nsamp * 8 * s_vol + nrays * 4
    return nvox * nviews * 4 * s_proj + nvox * 4


# ---------------------------------------------------------------- counters
def read_pmc(kernel, workload, sources):
    """Per-launch counters of a kernel from profiles/pmc_<kernel>_<workload>.json,
    or None when absent or recorded for other kernel sources."""
    path = os.path.join(ROOT, "profiles", "pmc_%s_%s.json" % (kernel, workload))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if d.get("sources_sha") != sources_hash(sources):
        return None
    return d


def kernel_matches(kernel, name):
    """Whether a profiler kernel name (demangled, or an Itanium-mangled one
    such as _ZN12_GLOBAL__N_115k_toeplitz_mfmaE...) is `kernel` itself, not
    a longer name that starts with it."""
    import re
    if name.startswith("_Z"):
        return ("%d%s" % (len(kernel), kernel)) in name
    return re.search(r"(?<![A-Za-z0-9_])%s(?=[<(])" % re.escape(kernel), name) is not None


def summarize_pmc(kernel, workload, sources, dirs, out_path):
    """Collect the per-dispatch counters of `kernel` from rocprofv3 --pmc CSV
    output directories (one counter set per run) into one JSON summary.
    HBM bytes: FETCH_SIZE and WRITE_SIZE (KiB per dispatch) scaled by what
    they read for the 1 GiB device copy of the same run (--pmc-calibrate)."""
    import csv
    import glob
    counters, calib, cmds = {}, {}, []
    for d in dirs:
        for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(path) as f:
                rows = list(csv.DictReader(f))
            per = {}
            for r in rows:
                name = r["Kernel_Name"]
                key = (name, r["Dispatch_Id"])
                per.setdefault(key, {})
                per[key][r["Counter_Name"]] = per[key].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
            for (name, _), vals in per.items():
                if kernel_matches(kernel, name):
                    for c, v in vals.items():
                        counters.setdefault(c, []).append(v)
                elif "copyBuffer" in name:
                    for c, v in vals.items():
                        calib.setdefault(c, []).append(v)
        cmdf = os.path.join(d, "command.txt")
        if os.path.exists(cmdf):
            cmds.append(open(cmdf).read().strip())
    out = {"kernel": kernel, "workload": workload, "sources_sha": sources_hash(sources),
           "counters_per_launch": {c: sum(v) / len(v) for c, v in counters.items()},
           "launches": {c: len(v) for c, v in counters.items()}, "commands": cmds}
    gib_kib = float(1 << 20)
    c = out["counters_per_launch"]
    if "FETCH_SIZE" in c and calib.get("FETCH_SIZE"):
        f = gib_kib / max(calib["FETCH_SIZE"])
        out["fetch_bytes"] = c["FETCH_SIZE"] * 1024 * f
        out["fetch_factor"] = f
    if "WRITE_SIZE" in c and calib.get("WRITE_SIZE"):
        f = gib_kib / max(calib["WRITE_SIZE"])
        out["write_bytes"] = c["WRITE_SIZE"] * 1024 * f
        out["write_factor"] = f
    if "fetch_bytes" in out and "write_bytes" in out:
        out["bytes_per_launch"] = out["fetch_bytes"] + out["write_bytes"]
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


def kernel_rooflines(fp_kernel, fp_ms, bp_kernel, bp_ms, workload, nvox, nviews, nrays, nsamp,
                     s_store, filt):
    """Roofline of each projector kernel and the ramp filter."""
    out = {}
    pmc = read_pmc(fp_kernel, workload, FP_SOURCES)
    fp = {"kernel": fp_kernel, "kernel_ms": round(fp_ms, 3), "bound": "valu",
          "peak": VALU_PEAK_TOPS, "unit": "T VALU lane-instructions/s",
          "effective_gbs_footprint": round(footprint_bytes("fp", nvox, nviews, nrays, nsamp, s_store,
                                                           s_store) / (fp_ms / 1e3) / 1e9, 1)}
    if pmc:
        c = pmc["counters_per_launch"]
        if "SQ_INSTS_VALU" in c:
            lane_ops = c["SQ_INSTS_VALU"] * 64
            fp["achieved"] = round(lane_ops / (fp_ms / 1e3) / 1e12, 2)
            fp["frac"] = round(fp["achieved"] / VALU_PEAK_TOPS, 4)
            fp["valu_lane_ops_per_sample"] = round(lane_ops / (nrays * nsamp), 2)
        if "bytes_per_launch" in pmc:
            fp["traffic"] = pmc["bytes_per_launch"]
            fp["hbm_gbs_measured"] = round(pmc["bytes_per_launch"] / (fp_ms / 1e3) / 1e9, 1)
            fp["hbm_frac_measured"] = round(fp["hbm_gbs_measured"] / HBM_PEAK_GBS, 4)
    out["fp"] = fp
    pmc = read_pmc(bp_kernel, workload, BP_SOURCES)
    bp = {"kernel": bp_kernel, "kernel_ms": round(bp_ms, 3),
          "effective_gbs_footprint": round(footprint_bytes("bp", nvox, nviews, nrays, nsamp, s_store,
                                                           s_store) / (bp_ms / 1e3) / 1e9, 1)}
    if pmc:
        # the binding resource: measured HBM bytes or VALU issue, whichever
        # runs closer to its peak
        c = pmc["counters_per_launch"]
        hbm = valu = None
        if "bytes_per_launch" in pmc:
            gbs = pmc["bytes_per_launch"] / (bp_ms / 1e3) / 1e9
            hbm = {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": round(gbs / HBM_PEAK_GBS, 4)}
            bp["traffic"] = pmc["bytes_per_launch"]
            bp["hbm_gbs_measured"] = hbm["achieved"]
            bp["hbm_frac_measured"] = hbm["frac"]
            # compulsory: the sinogram once at storage precision, the volume written once
            bp["traffic_vs_compulsory"] = round(pmc["bytes_per_launch"] / (nrays * s_store + nvox * 4), 2)
        if "SQ_INSTS_VALU" in c:
            lane_ops = c["SQ_INSTS_VALU"] * 64
            tops = lane_ops / (bp_ms / 1e3) / 1e12
            valu = {"achieved": round(tops, 2), "peak": VALU_PEAK_TOPS, "unit": "T VALU lane-instructions/s",
                    "frac": round(tops / VALU_PEAK_TOPS, 4)}
            bp["valu_frac"] = valu["frac"]
            bp["valu_lane_ops_per_update"] = round(lane_ops / (nvox * nviews), 2)
        best = max((r for r in (("hbm", hbm), ("valu", valu)) if r[1]), key=lambda r: r[1]["frac"],
                   default=None)
        if best:
            bp["bound"] = best[0]
            bp.update(best[1])
    else:
        bp.update({"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s"})
    out["bp"] = bp
    if filt:
        # the default engine (FFT): with its counters, the measured HBM bytes
        # or VALU issue, whichever runs closer to its peak (as for BP); without
        # them, HBM on its algorithmic bytes (each row read and written once).
        # The MFMA engine beside it, against the MFMA peak.
        e2, e1 = filt["engine2"], filt["engine1"]
        f_ms = e2["kernel_ms"]
        fl = {"kernel": e2["kernel"], "kernel_ms": f_ms, "bound": "hbm", "peak": HBM_PEAK_GBS,
              "unit": "GB/s", "achieved": e2["gbs_algorithmic"],
              "frac": round(e2["gbs_algorithmic"] / HBM_PEAK_GBS, 4),
              "hbm_frac_algorithmic": round(e2["gbs_algorithmic"] / HBM_PEAK_GBS, 4),
              "traffic": e2["bytes_algorithmic"]}
        pmc = read_pmc(e2["kernel"], workload, FILTER_SOURCES)
        if pmc:
            c = pmc["counters_per_launch"]
            hbm = valu = None
            if "bytes_per_launch" in pmc:
                gbs = pmc["bytes_per_launch"] / (f_ms / 1e3) / 1e9
                hbm = {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(gbs / HBM_PEAK_GBS, 4)}
                fl["traffic"] = pmc["bytes_per_launch"]
                fl["hbm_frac_measured"] = hbm["frac"]
            if "SQ_INSTS_VALU" in c:
                tops = c["SQ_INSTS_VALU"] * 64 / (f_ms / 1e3) / 1e12
                valu = {"achieved": round(tops, 2), "peak": VALU_PEAK_TOPS,
                        "unit": "T VALU lane-instructions/s", "frac": round(tops / VALU_PEAK_TOPS, 4)}
                fl["valu_frac"] = valu["frac"]
            best = max((r for r in (("hbm", hbm), ("valu", valu)) if r[1]), key=lambda r: r[1]["frac"],
                       default=None)
            if best:
                fl["bound"] = best[0]
                fl.update(best[1])
        fl["mfma_engine"] = {"kernel": e1["kernel"], "kernel_ms": e1["kernel_ms"], "bound": "mfma",
                             "peak": FP16_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                             "achieved": e1["tflops_algorithmic"],
                             "frac": round(e1["tflops_algorithmic"] / FP16_MFMA_PEAK_TFLOPS, 4),
                             "issued_frac": e1["mfma_frac_of_2500_tflops"]}
        out["filter"] = fl
    return out


# ------------------------------------------------------------ CPU baseline
def host_info():
    model = platform.processor() or ""
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(cfg, vol_host, angles, proj_f32_dev, proj_bf16_dev, bp32_dev, bp16_dev,
                 filt_dev, ctx, budget_s=12.0):
    """CPU baseline leg: the FP32 OpenMP restatement (oracle/libcpucbct.so)
    and the FP64 oracle timed on this host on bounded samples of the same
    workload, and the error of each GPU stage (FP32 and BF16 working copies)
    against the FP64 oracle on a deterministic subsample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle as O
    g = cfg.geom()
    og = O.Geom(g.nx, g.ny, g.nz, g.lx, g.ly, g.lz, g.nu, g.nv, g.fov_u, g.fov_v, g.sad, g.sdd)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    nvox = g.nx * g.ny * g.nz

    def run(n, fp_fn, bp_fn, vol):
        a = angles[:n]
        t0 = time.perf_counter()
        p = fp_fn(og, a, None, vol)
        t1 = time.perf_counter()
        bp_fn(og, a, None, p)
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1

    tf, tb = run(1, O.cpu_fp_f32, O.cpu_bp_f32, vol_host)
    n = int(max(1, min(32, budget_s / max(tf + tb, 1e-3))))
    if n > 1:
        tf, tb = run(n, O.cpu_fp_f32, O.cpu_bp_f32, vol_host)
    gups = 2.0 * nvox * n / (tf + tb) / 1e9
    res = {"value": round(gups, 4), "unit": "GUPS", "cores": threads, "kind": "port",
           "sample": "FP + BP of the full 512^3 volume for the first %d of 720 views (FP32 "
                     "OpenMP restatement, %d threads; fp %.2fs, bp %.2fs), extrapolated per "
                     "voxel-view update" % (n, threads, tf, tb),
           "host": host_info()}
    # FP64 oracle, same kind of sample
    v64 = vol_host.astype(np.float64)
    tf, tb = run(1, O.fp, O.bp, v64)
    res["fp64_oracle"] = {"value": round(2.0 * nvox / (tf + tb) / 1e9, 4), "unit": "GUPS",
                          "sample": "FP + BP of the full volume for 1 view (fp %.2fs, bp %.2fs)"
                                    % (tf, tb)}
    # per-stage error against the FP64 oracle, deterministic subsample
    import torch
    sel = np.arange(0, len(angles), 90)
    uu, vv = np.meshgrid(np.arange(1, g.nu, 4), np.arange(2, g.nv, 4), indexing="xy")
    rays = np.stack([np.repeat(np.arange(len(sel)), uu.size), np.tile(uu.ravel(), len(sel)),
                     np.tile(vv.ravel(), len(sel))], -1)
    ref_fp = O.fp_rays(og, angles[sel], None, v64, rays)
    sel_t = torch.from_numpy(sel).cuda()

    def at_rays(p):
        h = p[sel_t].cpu().numpy().astype(np.float64)
        return h[rays[:, 0], rays[:, 2], rays[:, 1]]

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    vox = np.sort(np.random.default_rng(512).choice(nvox, nvox // 1000, replace=False))
    ph = proj_f32_dev.cpu().numpy().astype(np.float64)
    ref_bp = O.bp_voxels(og, angles, None, ph, vox)
    del ph
    F = np.asarray(filt_dev["F"], np.float32)
    ph_sel = proj_f32_dev[sel_t].cpu().numpy().astype(np.float64)
    ref_filt = O.fdk_prefilter(og, angles[sel], None, ph_sel, F)
    st = {"subsample": "FP: views 0, 90, .., 630 x every 4th detector column and row (%d rays); "
                       "BP: 0.1%% of the voxels (%d) over all views of the FP32 path's "
                       "projections; filter: the same 8 views' rows" % (len(rays), len(vox)),
          "fp_f32": rel(at_rays(proj_f32_dev), ref_fp),
          "bp_f32": rel(bp32_dev.cpu().numpy().ravel()[vox].astype(np.float64), ref_bp)}
    if proj_bf16_dev is not None:
        st["fp_bf16"] = rel(at_rays(proj_bf16_dev), ref_fp)
        st["bp_bf16"] = rel(bp16_dev.cpu().numpy().ravel()[vox].astype(np.float64), ref_bp)
    for e in (0, 1, 2):
        got = filt_dev["out%d" % e][sel_t].cpu().numpy().astype(np.float64)
        st["filter_engine%d" % e] = rel(got, ref_filt)
    res["stage_rel_l2_vs_fp64_oracle"] = st
    return res


def cpu_workloads(fdk_cfg, vol_host, angles):
    """CPU FP32 OpenMP restatement at the 128^3 FDK workload: full FP and BP."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    g = fdk_cfg.geom()
    og = O.Geom(g.nx, g.ny, g.nz, g.lx, g.ly, g.lz, g.nu, g.nv, g.fov_u, g.fov_v, g.sad, g.sdd)
    t0 = time.perf_counter()
    p = O.cpu_fp_f32(og, angles, None, vol_host)
    t1 = time.perf_counter()
    O.cpu_bp_f32(og, angles, None, p)
    t2 = time.perf_counter()
    nvv = g.nx * g.ny * g.nz * len(angles)
    return {"fp_s": round(t1 - t0, 3), "bp_s": round(t2 - t1, 3),
            "gups_fp": round(nvv / (t1 - t0) / 1e9, 3), "gups_bp": round(nvv / (t2 - t1) / 1e9, 3)}


# ---------------------------------------------------------------- GPU legs
def measure_filter(ctx, g, proj, nviews, reps=3):
    """FDK pre-filter (cosine weights + ramp filter) of the whole sinogram,
    all three engines.  Kernel time of the filter kernel from the library's own
    HIP events; whole-call time (including the MFMA engine's weighting / fp16
    split pass) from events on the same stream."""
    import torch

    from gtt import capi
    F = capi.default_filter(g.nu)
    ctx.set_filter(F)
    flops_alg = 2.0 * g.nu * g.nu * g.nv * nviews
    bytes_alg = 2.0 * 4 * g.nu * g.nv * nviews      # every row read and written once
    res = {"rows": g.nv * nviews, "nu": g.nu, "flops_algorithmic": flops_alg}
    outs = {"F": F}
    names = {0: "k_toeplitz_f32", 1: "k_toeplitz_mfma", 2: "k_ramp_fft"}
    for engine in (0, 1, 2):
        out = torch.empty_like(proj)
        ctx.filter_engine(engine)
        ctx.fdk_prefilter(g, proj, out)          # warm-up (builds workspaces)
        torch.cuda.synchronize()
        ctx.profile(True)
        kern, call = [], []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            ctx.fdk_prefilter(g, proj, out)
            e1.record()
            torch.cuda.synchronize()
            kern.append(ctx.profile_read()[2])
            call.append(e0.elapsed_time(e1))
        ctx.profile(False)
        outs["out%d" % engine] = out
        km = float(sorted(kern)[len(kern) // 2])
        res["engine%d" % engine] = {
            "kernel": names[engine],
            "kernel_ms": round(km, 3), "call_ms": round(float(sorted(call)[len(call) // 2]), 3),
            "tflops_algorithmic": round(flops_alg / (km / 1e3) / 1e12, 1)}
    res["engine1"]["tflops_issued"] = round(3 * flops_alg / (res["engine1"]["kernel_ms"] / 1e3) / 1e12, 1)
    res["engine1"]["mfma_frac_of_2500_tflops"] = round(res["engine1"]["tflops_issued"] / 2500.0, 4)
    res["engine2"]["bytes_algorithmic"] = bytes_alg
    res["engine2"]["gbs_algorithmic"] = round(bytes_alg / (res["engine2"]["kernel_ms"] / 1e3) / 1e9, 1)
    for e in (1, 2):
        res["engine%d" % e]["rel_l2_vs_engine0"] = (ctx.nrm2(outs["out%d" % e] - outs["out0"]) /
                                                    ctx.nrm2(outs["out0"]))
    ctx.filter_engine(2)
    return res, outs


def measured_peaks():
    """HBM copy bandwidth (read + write bytes) and dense FP16 GEMM rate on
    this device, for comparison with the published peaks."""
    import torch
    out = {}
    n = 1 << 29   # 2 GiB of float32
    a = torch.empty(n, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    out["hbm_copy_gbs"] = round(10 * 2 * 4 * n / (e0.elapsed_time(e1) / 1e3) / 1e9, 1)
    del a, b
    m = 8192
    x = torch.randn(m, m, device="cuda", dtype=torch.float16)
    y = torch.randn(m, m, device="cuda", dtype=torch.float16)
    z = x @ y
    torch.cuda.synchronize()
    e0.record()
    for _ in range(10):
        z = x @ y
    e1.record()
    torch.cuda.synchronize()
    out["fp16_gemm_tflops"] = round(10 * 2 * m ** 3 / (e0.elapsed_time(e1) / 1e3) / 1e12, 1)
    del x, y, z
    torch.cuda.empty_cache()
    return out


def workloads_leg(ctx):
    """The 128^3 FDK workload (projector pair, FDK_reconstruct_3d latency
    from host projections to host volume) and the 256^3 OS workload
    (20-iteration OS-SART and osMOMSolverD + TV over 10 subsets, wall time
    from host projections to host volume), FP32 working copies."""
    import numpy as np
    import torch

    from gtt import operators as ops
    from gtt import phantom as PH
    out = {}
    cfg = PH.FDK128
    g = cfg.geom()
    angles = cfg.angles()
    vol = PH.voxelize(cfg.phantom(), cfg.vol_shape, cfg.vol_mm, device="cuda")
    ctx.set_views(angles)
    proj = torch.empty(g.proj_shape(len(angles)), device="cuda")
    vout = torch.empty_like(vol)
    for _ in range(2):
        ctx.fp(g, vol, proj)
        ctx.bp(g, proj, vout)
    torch.cuda.synchronize()
    ctx.profile(True)
    fps, bps = [], []
    for _ in range(5):
        ctx.fp(g, vol, proj)
        ctx.bp(g, proj, vout)
        f, b, _ = ctx.profile_read()
        fps.append(f)
        bps.append(b)
    ctx.profile(False)
    nvv = g.nx * g.ny * g.nz * len(angles)
    fpm, bpm = float(np.median(fps)), float(np.median(bps))
    ph = proj.cpu().numpy()
    F = np.arange(g.nu + 1, dtype=np.float32)   # Ram-Lak
    ops.fdk_reconstruct_3d(g, angles, None, ph, filt=F)
    lat = []
    for _ in range(5):
        t0 = time.perf_counter()
        ops.fdk_reconstruct_3d(g, angles, None, ph, filt=F)
        lat.append((time.perf_counter() - t0) * 1e3)
    out["fdk128"] = {"fp_ms": round(fpm, 3), "bp_ms": round(bpm, 3),
                     "gups_fp": round(nvv / fpm / 1e6, 1), "gups_bp": round(nvv / bpm / 1e6, 1),
                     "gups_pair": round(2 * nvv / (fpm + bpm) / 1e6, 1),
                     "fdk_reconstruct_3d_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                                               "min": round(min(lat), 2)},
                     "fdk_note": "host projections in, host volume out (H2D, cosine weights, "
                                 "Ram-Lak filter, FBP, FP, LSQ rescale, D2H)"}
    cfg = PH.OS256
    g = cfg.geom()
    angles = cfg.angles()
    vol = PH.voxelize(cfg.phantom(), cfg.vol_shape, cfg.vol_mm, device="cuda")
    ctx.set_views(angles)
    proj = torch.empty(g.proj_shape(len(angles)), device="cuda")
    ctx.fp(g, vol, proj)
    clean = proj.cpu().numpy().astype(np.float64)
    counts = np.random.default_rng(1001).poisson(1e5 * np.exp(-clean))
    noisy = (-np.log(np.maximum(counts, 1) / 1e5)).astype(np.float32)
    del counts, clean, proj
    ops.os_sart(g, angles, None, noisy, 10, 1)    # warm-up
    t0 = time.perf_counter()
    ops.os_sart(g, angles, None, noisy, 10, 20)
    t_sart = time.perf_counter() - t0
    t0 = time.perf_counter()
    ops.os_mom(g, angles, None, noisy, 10, 20, tv_weight=1e-3, tau=1e-5, reg_steps=2, nonneg=True)
    t_mom = time.perf_counter() - t0
    nvv = g.nx * g.ny * g.nz * len(angles)
    out["os256"] = {"os_sart_20it_s": round(t_sart, 3), "os_mom_tv_20it_s": round(t_mom, 3),
                    "os_sart_gups": round(2 * nvv * 21 / t_sart / 1e9, 1),
                    "note": "10 subsets, Poisson noise at I0=1e5; wall time from host projections "
                            "to host volume, including the normaliser passes"}
    del vol
    torch.cuda.empty_cache()
    return out


def serving_leg(ctx, lanes=4, n_batch=64, n_stream=32, load=0.8, served_batch=False):
    """Reconstruction latency and throughput, served (worker lanes on their
    own streams) vs plain (one lane: requests strictly one after another).

    batch:  uneven OS-SART scans (10 iterations, sizes drawn per scan), all
            arriving at t=0.
    stream: FDK (p=0.6) / 10-iteration OS-SART (p=0.4) at the same sizes,
            Poisson arrivals at `load` x the plain server's capacity (rate from
            a calibration pass of the same requests).
    Served and plain results are compared bitwise."""
    import numpy as np

    from gtt import serving as S

    def run(reqs, lanes_, arrivals=None, graphs=False):
        for r in reqs:
            r.volume = None
        with S.Server(lanes=lanes_, graphs=graphs) as srv:
            res, wall = S.run_stream(srv, reqs, arrivals)
        return res, wall

    def same(a, b):
        return all(np.array_equal(x["volume"], y["volume"]) for x, y in zip(a, b))

    out = {"lanes": lanes}
    t0 = time.perf_counter()
    reqs = [S.make_request(s, ctx) for s in S.uneven_scans(n_batch, seed=1003)]
    gen_s = time.perf_counter() - t0
    rp, wp = run(reqs, 1)
    sp = S.summarize(reqs, rp, wp)
    out["batch"] = {"plain": sp,
                    "sizes": "%d OS-SART scans, edge %d..%d, %d..%d views" % (
                        len(reqs), min(r.geom.nx for r in reqs), max(r.geom.nx for r in reqs),
                        min(len(r.angles) for r in reqs), max(len(r.angles) for r in reqs))}
    if served_batch:
        vp = [dict(volume=r["volume"].copy()) for r in rp]
        rs, ws = run(reqs, lanes)
        # served, iterations 2.. replayed as captured HIP graphs
        rg, wg = run(reqs, lanes, graphs=True)
        # served with BF16 working copies, against the plain FP32 results
        for r in reqs:
            r.storage = "bf16"
        rb, wb = run(reqs, lanes)
        for r in reqs:
            r.storage = "f32"
        errs = [float(np.linalg.norm(b["volume"] - p["volume"]) / np.linalg.norm(p["volume"]))
                for b, p in zip(rb, vp)]
        sb = S.summarize(reqs, rb, wb)
        sb["rel_l2_vs_plain_fp32"] = {"max": max(errs), "mean": float(np.mean(errs))}
        out["batch"].update({"served": S.summarize(reqs, rs, ws),
                             "served_graphs": S.summarize(reqs, rg, wg),
                             "graphs_identical": same(vp, rg), "served_bf16": sb,
                             "served_bf16_gups_vs_plain_fp32": round(sb["gups"] / sp["gups"], 3),
                             "identical": same(vp, rs)})
        del rs, rb, rg, vp
    else:
        out["batch"]["served"] = ("not run: the served run of this batch hung once on an MI355X and "
                                  "its cause is not found (--served-batch runs it)")
    del reqs, rp
    t0 = time.perf_counter()
    reqs = [S.make_request(s, ctx) for s in S.uneven_scans(n_stream, seed=1004, p_fdk=0.6)]
    gen_s += time.perf_counter() - t0
    rc, _ = run(reqs, 1)
    mean_service_s = float(np.mean([r["run_ms"] for r in rc])) / 1e3
    rate = load / mean_service_s
    arr = S.poisson_arrivals(len(reqs), rate, seed=7)
    rp, wp = run(reqs, 1, arr)
    vp = [dict(volume=r["volume"].copy()) for r in rp]
    rs, ws = run(reqs, lanes, arr)
    out["stream"] = {"rate_per_s": round(rate, 3), "load_vs_plain": load,
                     "fdk": sum(r.kind == S.REQ_FDK for r in reqs),
                     "plain": S.summarize(reqs, rp, wp), "served": S.summarize(reqs, rs, ws),
                     "identical": same(vp, rs)}
    out["request_generation_s"] = round(gen_s, 2)
    return out


def serving_child(args, limit_s=480):
    """The serving leg in a child process under a time limit: a server that
    stops making progress costs the latency figures, not the line."""
    cmd = [sys.executable, os.path.abspath(__file__), "--serving-only"]
    if args.served_batch:
        cmd.append("--served-batch")
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=limit_s)
    except subprocess.TimeoutExpired:
        return {"error": "the serving leg ran past %d s and was stopped" % limit_s}
    for line in reversed(r.stdout.splitlines()):
        if line.startswith("{"):
            try:
                return json.loads(line)["latency"]
            except (ValueError, KeyError):
                break
    return {"error": "the serving leg exited with status %d" % r.returncode}


def batch_trace(ctx, lanes, graphs, storage="f32", n_batch=64):
    """One run of the uneven 64-scan OS-SART batch (for a kernel trace): the
    clock readings around it bound the window trace_busy looks at."""
    from gtt import serving as S
    reqs = [S.make_request(s, ctx) for s in S.uneven_scans(n_batch, seed=1003)]
    for r in reqs:
        r.storage = storage
    import torch
    torch.cuda.synchronize()
    time.sleep(0.5)
    t0 = (time.monotonic_ns(), time.clock_gettime_ns(time.CLOCK_BOOTTIME))
    with S.Server(lanes=lanes, graphs=graphs) as srv:
        res, wall = S.run_stream(srv, reqs)
    t1 = (time.monotonic_ns(), time.clock_gettime_ns(time.CLOCK_BOOTTIME))
    out = {"lanes": lanes, "graphs": graphs, "storage": storage, "summary": S.summarize(reqs, res, wall),
           "window_monotonic_ns": [t0[0], t1[0]], "window_boottime_ns": [t0[1], t1[1]]}
    print(json.dumps(out), flush=True)
    return out


def trace_busy(dirs, window=None):
    """Device occupancy of a kernel trace (rocprofv3 --kernel-trace CSV):
    the union of all kernels' [start, end) intervals against the span from
    the first start to the last end, the number of dispatches and the gaps
    between busy intervals."""
    import csv
    import glob
    iv = []
    for d in dirs:
        for path in glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True):
            with open(path) as f:
                for r in csv.DictReader(f):
                    s, e = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
                    if window is None or (s >= window[0] and e <= window[1]):
                        iv.append((s, e))
    if not iv:
        print(json.dumps({"dispatches": 0}))
        return None
    iv.sort()
    busy, gaps = 0, []
    cs, ce = iv[0]
    for s, e in iv[1:]:
        if s > ce:
            busy += ce - cs
            gaps.append(s - ce)
            cs, ce = s, e
        else:
            ce = max(ce, e)
    busy += ce - cs
    span = iv[-1][1] - iv[0][0]
    gaps.sort()
    out = {"dispatches": len(iv), "span_ms": round(span / 1e6, 3), "busy_ms": round(busy / 1e6, 3),
           "busy_frac": round(busy / span, 4), "idle_gaps": len(gaps),
           "gap_p50_us": round(gaps[len(gaps) // 2] / 1e3, 2) if gaps else 0.0,
           "gap_total_ms": round(sum(gaps) / 1e6, 3)}
    print(json.dumps(out))
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="prec512")
    ap.add_argument("--storage", choices=["f32", "bf16"], default="bf16",
                    help="storage type of the projectors' working copies of the volume and "
                         "sinogram (arithmetic is FP32)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the plain FP32 path, the filter and the other workloads")
    ap.add_argument("--served-batch", action="store_true",
                    help="also run the 64-scan batch served on four lanes (off by default: that run "
                         "hung once on an MI355X and its cause is not found)")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the served-vs-plain reconstruction latency measurement")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="only run this many FP+BP steps (for rocprof), no JSON extras")
    ap.add_argument("--pmc-calibrate", action="store_true",
                    help="with --profile-steps: first run a plain 1 GiB device copy whose "
                         "FETCH_SIZE/WRITE_SIZE scale the projector kernels' counters")
    ap.add_argument("--serving-only", action="store_true",
                    help="run only the served-vs-plain batch and stream legs (for tracing)")
    ap.add_argument("--trace-busy", nargs="+", metavar="DIR",
                    help="device occupancy of rocprofv3 --kernel-trace CSV output (no GPU work)")
    ap.add_argument("--window", nargs=2, type=int, metavar=("T0", "T1"),
                    help="with --trace-busy: only dispatches inside [T0, T1] (ns, the trace's clock)")
    ap.add_argument("--batch-trace", nargs="+", metavar="ARG",
                    help="LANES [graphs] [bf16]: run the 64-scan batch once (for rocprofv3 --kernel-trace)")
    ap.add_argument("--pmc-summary", nargs="+", metavar="ARG",
                    help="KERNEL WORKLOAD {fp,bp,filter} DIR...: collect rocprofv3 --pmc CSV "
                         "output into profiles/pmc_KERNEL_WORKLOAD.json (no GPU work)")
    args = ap.parse_args()

    if args.pmc_summary:
        kernel, workload, which = args.pmc_summary[:3]
        src = {"fp": FP_SOURCES, "bp": BP_SOURCES, "filter": FILTER_SOURCES}[which]
        summarize_pmc(kernel, workload, src, args.pmc_summary[3:],
                      os.path.join(ROOT, "profiles", "pmc_%s_%s.json" % (kernel, workload)))
        return

    if args.trace_busy:
        trace_busy(args.trace_busy, args.window)
        return

    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the path has no multi-GPU "
              "sharding here); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from gtt import capi
    from gtt import phantom as PH

    cfg = PH.CONFIGS[args.config]
    g = cfg.geom()
    angles = cfg.angles()
    nviews = len(angles)
    nvox = g.nx * g.ny * g.nz
    nrays = g.nu * g.nv * nviews
    nsamp = int(np.float32(1.5) * np.float32(g.nx))

    ctx = capi.Context(0)
    if args.serving_only:
        print(json.dumps({"latency": serving_leg(ctx, served_batch=args.served_batch)}), flush=True)
        ctx.close()
        return
    if args.batch_trace:
        batch_trace(ctx, int(args.batch_trace[0]), "graphs" in args.batch_trace,
                    "bf16" if "bf16" in args.batch_trace else "f32")
        ctx.close()
        return
    ctx.set_views(angles, cfg.offsets())
    vol = PH.voxelize(cfg.phantom(), cfg.vol_shape, cfg.vol_mm, device="cuda")
    torch.cuda.synchronize()

    def measure(storage, steps, warmup):
        """One step = FP of the volume onto all views + BP of all views.  With
        'bf16' the projectors' working copies of the volume and the sinogram
        (the arrays their kernels gather from) are BF16, converted inside the
        copy kernels each call; arithmetic and the caller's arrays are FP32."""
        ctx.storage(storage)
        proj = torch.empty(g.proj_shape(nviews), device="cuda", dtype=torch.float32)
        v_out = torch.empty_like(vol)

        def step():
            ctx.fp(g, vol, proj)
            ctx.bp(g, proj, v_out)

        for _ in range(warmup):
            step()
        ctx.profile(True)
        torch.cuda.synchronize()
        fp_ms, bp_ms = [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
            f, b, _ = ctx.profile_read()
            fp_ms.append(f)
            bp_ms.append(b)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ctx.profile(False)
        ctx.storage("f32")
        return (t1 - t0) / steps * 1e3, float(np.mean(fp_ms)), float(np.mean(bp_ms)), proj, v_out

    if args.profile_steps:
        if args.pmc_calibrate:
            src = torch.ones(1 << 28, device="cuda")   # 1 GiB
            dst = torch.empty_like(src)
            dst.copy_(src)
            torch.cuda.synchronize()
            del src, dst
        _, _, _, proj, _ = measure(args.storage, args.profile_steps, 0)
        # one FDK pre-filter of the whole sinogram per filter engine
        ctx.set_filter(capi.default_filter(g.nu))
        fout = torch.empty_like(proj)
        for engine in (0, 1, 2):
            ctx.filter_engine(engine)
            ctx.fdk_prefilter(g, proj, fout)
        torch.cuda.synchronize()
        print("profile run done", flush=True)
        return

    print("bench: %s %r, %d views, storage %s" % (cfg.name, g, nviews, args.storage),
          file=sys.stderr, flush=True)
    ms_step, fp_ms, bp_ms, proj, v_out = measure(args.storage, args.steps, args.warmup)
    value = 2.0 * nvox * nviews / (ms_step / 1e3) / 1e9
    print("bench: headline %.1f GUPS (%.2f ms per step)" % (value, ms_step), file=sys.stderr, flush=True)

    s_store = 4 if args.storage == "f32" else 2
    fp_kernel = ctx.fp_kernel(g, 0, nviews, args.storage)
    bp_kernel = ctx.bp_kernel(g, 0, nviews)
    extra = {"fp_ms": round(fp_ms, 3), "bp_ms": round(bp_ms, 3),
             "gups_fp": round(nvox * nviews / (fp_ms / 1e3) / 1e9, 2),
             "gups_bp": round(nvox * nviews / (bp_ms / 1e3) / 1e9, 2),
             "fp_samples_per_s": round(nrays * nsamp / (fp_ms / 1e3), 1)}
    p32 = v32 = filt = fouts = None
    if not args.no_variants:
        if args.storage == "bf16":
            # the plain FP32 path on the same inputs
            ms2, f2, b2, p32, v32 = measure("f32", max(1, args.steps - 1), 1)
            g32 = 2.0 * nvox * nviews / (ms2 / 1e3) / 1e9
            extra["plain_fp32"] = {"ms_per_step": round(ms2, 3), "fp_ms": round(f2, 3),
                                   "bp_ms": round(b2, 3), "gups": round(g32, 2),
                                   "speedup_of_bf16_storage": round(value / g32, 3)}
            # BF16 BP stage on the FP32 path's projections (for the oracle check)
            ctx.storage("bf16")
            vb = torch.empty_like(vol)
            ctx.bp(g, p32, vb)
            ctx.storage("f32")
            v_out = vb
        else:
            p32, v32 = proj, v_out
        filt, fouts = measure_filter(ctx, g, proj if p32 is None else p32, nviews)
        extra["fdk_prefilter"] = filt
        extra["measured_peaks"] = measured_peaks()
    roof = kernel_rooflines(fp_kernel, fp_ms, bp_kernel, bp_ms, "%s_%s" % (cfg.name, args.storage),
                            nvox, nviews, nrays, nsamp, s_store, filt)
    roofline = dict(roof["fp"] if fp_ms >= bp_ms else roof["bp"])
    roofline.setdefault("achieved", None)
    roofline.setdefault("frac", None)
    roofline.setdefault("traffic", None)
    extra["rooflines"] = roof

    cpu = None
    if not args.no_cpu_baseline and p32 is not None:
        vol_host = vol.cpu().numpy()
        bf16 = args.storage == "bf16"
        cpu = cpu_baseline(cfg, vol_host, angles, p32, proj if bf16 else None, v32,
                           v_out if bf16 else None, fouts, ctx)
        print("bench: cpu baseline done", file=sys.stderr, flush=True)
    del fouts
    torch.cuda.empty_cache()

    workloads = None
    if not args.no_variants:
        workloads = workloads_leg(ctx)
        if cpu is not None:
            fc = PH.FDK128
            fvol = PH.voxelize(fc.phantom(), fc.vol_shape, fc.vol_mm).numpy()
            workloads["fdk128"]["cpu_fp32_openmp"] = cpu_workloads(fc, fvol, fc.angles())
        ctx.set_views(angles, cfg.offsets())
        print("bench: workloads done", file=sys.stderr, flush=True)

    latency = None
    if not args.no_serving:
        del proj, v_out, p32, v32
        torch.cuda.empty_cache()
        latency = serving_child(args)

    line = {
        "metric": "fwd+back GUPS at 512^3 x 720 views",
        "value": round(value, 3),
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.storage,
        "data": "synthetic (seeded ellipsoid phantom)",
        "config": {"workload": "%s: %dx%dx%d volume @ %.2f mm, %d views over 360 deg, "
                               "%dx%d detector @ %.2f mm, SAD %.0f, SDD %.0f; step = FP + BP"
                               % (cfg.name, g.nx, g.ny, g.nz, g.lx / g.nx, nviews, g.nu, g.nv,
                                  g.fov_u / g.nu, g.sad, g.sdd),
                   "storage": "%s working copies of volume and sinogram" % args.storage.upper(),
                   "arithmetic": "FP32 (accumulation FP32; FP64 reductions)",
                   "samples_per_ray": nsamp},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "workloads": workloads,
        "latency": latency,
        "detail": extra,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
