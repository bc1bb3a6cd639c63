#!/usr/bin/env python3
"""fwd+back GUPS of the gfx950 cone-beam projector pair on one MI355X.

Workload (default): 512^3 volume at 0.5 mm, 720 views over 360 degrees,
1024x768 flat detector at 0.39 mm, SAD 1000 mm, SDD 1500 mm, seeded
ellipsoid phantom (synthetic); the projectors' working copies of the volume
and the sinogram in BF16, arithmetic in FP32 (--storage f32 for FP32
copies).  One step = one forward projection of the volume onto all 720
views + one backprojection of all 720 views, inputs resident in HBM.
GUPS = 2 * Nvox * Nviews / t_step / 1e9.  The plain FP32 path is timed in
the same run, with the error of each BF16 stage against it.

Prints one JSON line (rank 0) with the roofline of the dominant kernel and
a CPU baseline (FP32 OpenMP restatement timed on this host on a bounded
sample of views).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gt-tomography_amd"))

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(kind, nvox, nviews, nrays, nsamp, s_vol, s_proj):
    """Footprint-convention bytes of one launch.
    FP: every sample reads its 8-tap footprint, every ray writes one value.
    BP: every voxel-view reads its 4-tap footprint, every voxel is written."""
    if kind == "fp":
        return nrays * nsamp * 8 * s_vol + nrays * 4
    return nvox * nviews * 4 * s_proj + nvox * 4


def cpu_baseline(cfg, vol_host, angles, budget_s=12.0):
    """FP32 OpenMP restatement (oracle/libcpucbct.so) on a bounded sample:
    forward and back projection of the full volume for the first n views."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle as O
    g = cfg.geom()
    og = O.Geom(g.nx, g.ny, g.nz, g.lx, g.ly, g.lz, g.nu, g.nv, g.fov_u, g.fov_v, g.sad, g.sdd)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    def run(n):
        a = angles[:n]
        t0 = time.perf_counter()
        p = O.cpu_fp_f32(og, a, None, vol_host)
        t1 = time.perf_counter()
        O.cpu_bp_f32(og, a, None, p)
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1

    tf, tb = run(1)
    n = int(max(1, min(32, budget_s / max(tf + tb, 1e-3))))
    if n > 1:
        tf, tb = run(n)
    nvox = g.nx * g.ny * g.nz
    gups = 2.0 * nvox * n / (tf + tb) / 1e9
    return {"value": round(gups, 4), "unit": "GUPS", "cores": threads, "kind": "port",
            "sample": "FP + BP of the full 512^3 volume for the first %d of 720 views "
                      "(FP32 OpenMP restatement; fp %.2fs, bp %.2fs)" % (n, tf, tb)}


def measure_filter(ctx, g, proj, nviews, reps=3):
    """FDK pre-filter (cosine weights + ramp filter) of the whole sinogram,
    both engines.  Kernel time of the Toeplitz product from the library's own
    HIP events; whole-call time (including the weighting / fp16 split pass)
    from events on the same stream."""
    import torch

    from gtt import capi
    ctx.set_filter(capi.default_filter(g.nu))
    out = torch.empty_like(proj)
    flops_alg = 2.0 * g.nu * g.nu * g.nv * nviews
    res = {"rows": g.nv * nviews, "nu": g.nu, "flops_algorithmic": flops_alg}
    out0 = None
    for engine in (0, 1):
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
        if engine == 0:
            out0 = out.clone()
        km = float(sorted(kern)[len(kern) // 2])
        res["engine%d" % engine] = {
            "kernel": "k_toeplitz_f32" if engine == 0 else "k_toeplitz_mfma",
            "kernel_ms": round(km, 3), "call_ms": round(float(sorted(call)[len(call) // 2]), 3),
            "tflops_algorithmic": round(flops_alg / (km / 1e3) / 1e12, 1)}
    res["engine1"]["tflops_issued"] = round(3 * flops_alg / (res["engine1"]["kernel_ms"] / 1e3) / 1e12, 1)
    res["engine1"]["mfma_frac_of_2500_tflops"] = round(res["engine1"]["tflops_issued"] / 2500.0, 4)
    # the split-precision MFMA result against the FP32 engine (itself ~2e-7
    # from the FP64 oracle in the parity tests), FP64 device reductions
    res["engine1"]["rel_l2_vs_engine0"] = ctx.nrm2(out - out0) / ctx.nrm2(out0)
    ctx.filter_engine(1)
    return res


def serving_leg(ctx, lanes=4, n_batch=16, n_stream=24, load=0.8):
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

    def run(reqs, lanes_, arrivals=None):
        for r in reqs:
            r.volume = None
        with S.Server(lanes=lanes_) as srv:
            res, wall = S.run_stream(srv, reqs, arrivals)
        return res, wall

    def same(a, b):
        return all(np.array_equal(x["volume"], y["volume"]) for x, y in zip(a, b))

    out = {"lanes": lanes}
    t0 = time.perf_counter()
    reqs = [S.make_request(s, ctx) for s in S.uneven_scans(n_batch, seed=1003)]
    gen_s = time.perf_counter() - t0
    rp, wp = run(reqs, 1)
    vp = [dict(volume=r["volume"].copy()) for r in rp]
    rs, ws = run(reqs, lanes)
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
    sp = S.summarize(reqs, rp, wp)
    out["batch"] = {"plain": sp, "served": S.summarize(reqs, rs, ws), "served_bf16": sb,
                    "served_bf16_gups_vs_plain_fp32": round(sb["gups"] / sp["gups"], 3),
                    "identical": same(vp, rs),
                    "sizes": "%d OS-SART scans, edge %d..%d, %d..%d views" % (
                        len(reqs), min(r.geom.nx for r in reqs), max(r.geom.nx for r in reqs),
                        min(len(r.angles) for r in reqs), max(len(r.angles) for r in reqs))}
    del reqs, rp, rs, rb, vp
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


def read_traffic(kernel, workload):
    path = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (kernel, workload))
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    except Exception:
        return None


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
                    help="skip the extra bf16-storage and filter measurements")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the served-vs-plain reconstruction latency measurement")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="only run this many FP+BP steps (for rocprof), no JSON extras")
    ap.add_argument("--pmc-calibrate", action="store_true",
                    help="with --profile-steps: first run a plain 1 GiB device copy whose "
                         "FETCH_SIZE/WRITE_SIZE scale the projector kernels' counters")
    args = ap.parse_args()

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

    def rel_err(a, b):
        """||a - b|| / ||b|| with FP64 device reductions."""
        return ctx.nrm2(a - b) / ctx.nrm2(b)

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
        for engine in (0, 1):
            ctx.filter_engine(engine)
            ctx.fdk_prefilter(g, proj, fout)
        torch.cuda.synchronize()
        print("profile run done", flush=True)
        return

    print("bench: %s %r, %d views, storage %s" % (cfg.name, g, nviews, args.storage),
          file=sys.stderr, flush=True)
    ms_step, fp_ms, bp_ms, proj, v_out = measure(args.storage, args.steps, args.warmup)
    value = 2.0 * nvox * nviews / (ms_step / 1e3) / 1e9

    s_vol = 4 if args.storage == "f32" else 2
    kind = "fp" if fp_ms >= bp_ms else "bp"
    kms = fp_ms if kind == "fp" else bp_ms
    ab = algorithmic_bytes(kind, nvox, nviews, nrays, nsamp, s_vol, s_vol)
    achieved = ab / (kms / 1e3) / 1e9
    kname = ctx.fp_kernel(g, 0, nviews, args.storage) if kind == "fp" else "k_bp"
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": read_traffic(kname, "%s_%s" % (cfg.name, args.storage)),
                "kernel": kname, "kernel_ms": round(kms, 3), "algorithmic_bytes": ab,
                "convention": "footprint: FP 8 taps x storage bytes per sample + 4 B per ray; "
                              "BP 4 taps x storage bytes per voxel-view + 4 B per voxel"}
    bp_ab = algorithmic_bytes("bp", nvox, nviews, nrays, nsamp, s_vol, s_vol)

    extra = {"fp_ms": round(fp_ms, 3), "bp_ms": round(bp_ms, 3),
             "gups_fp": round(nvox * nviews / (fp_ms / 1e3) / 1e9, 2),
             "gups_bp": round(nvox * nviews / (bp_ms / 1e3) / 1e9, 2),
             "fp_samples_per_s": round(nrays * nsamp / (fp_ms / 1e3), 1),
             "bp_roofline_frac": round(bp_ab / (bp_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
    if args.storage == "bf16":
        # the plain FP32 path on the same inputs: its time, and the error of
        # each BF16 stage against it (the FP32 path itself agrees with the
        # FP64 oracle to ~1e-7 in the parity tests)
        ms2, f2, b2, p32, v32 = measure("f32", max(1, args.steps - 1), 1)
        g32 = 2.0 * nvox * nviews / (ms2 / 1e3) / 1e9
        extra["plain_fp32"] = {"ms_per_step": round(ms2, 3), "fp_ms": round(f2, 3),
                               "bp_ms": round(b2, 3), "gups": round(g32, 2),
                               "speedup_of_bf16_storage": round(value / g32, 3)}
        # FP stage: the step's projections; BP stage: BP of the FP32 sinogram
        # with BF16 working copies vs without
        ctx.storage("bf16")
        vb = torch.empty_like(vol)
        ctx.bp(g, p32, vb)
        ctx.storage("f32")
        torch.cuda.synchronize()
        extra["stage_rel_l2_vs_fp32_path"] = {"fp": rel_err(proj, p32), "bp": rel_err(vb, v32),
                                              "fp_then_bp": rel_err(v_out, v32)}
        del p32, v32, vb

    if not args.no_variants:
        extra["fdk_prefilter"] = measure_filter(ctx, g, proj, nviews)

    latency = None
    if not args.no_serving:
        del proj, v_out
        torch.cuda.empty_cache()
        try:
            latency = serving_leg(ctx)
        except Exception as e:  # reported in the line; the throughput figures stand
            latency = {"error": "%s: %s" % (type(e).__name__, e)}

    cpu = None
    if not args.no_cpu_baseline:
        vol_host = vol.cpu().numpy()
        cpu = cpu_baseline(cfg, vol_host, angles)

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
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom)",
        "config": {"workload": "%s: %dx%dx%d volume @ %.2f mm, %d views over 360 deg, "
                               "%dx%d detector @ %.2f mm, SAD %.0f, SDD %.0f; step = FP + BP"
                               % (cfg.name, g.nx, g.ny, g.nz, g.lx / g.nx, nviews, g.nu, g.nv,
                                  g.fov_u / g.nu, g.sad, g.sdd),
                   "storage": "%s working copies of volume and sinogram, FP32 arithmetic"
                              % args.storage.upper(),
                   "samples_per_ray": nsamp},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "latency": latency,
        "detail": extra,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
