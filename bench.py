#!/usr/bin/env python3
"""Benchmark of the GENFIRE reconstruction hot path on one MI355X.

A step is one whole reconstruction job of the tilt256 workload with its inputs
already resident in HBM: DFT gridding of 91 projections (256 x 256) onto the
512^3 oversampled grid, 200 Douglas-Rachford iterations, and the real-space
R-factor of all 91 projections every 10 iterations.  The metric is
reconstructions per hour; gridded k-points per second, p50/p99 job latency and
the final R-factor are reported beside it.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu-baseline]

Prints ONE JSON line.  --gpus > 1 is refused: the benchmark measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "recon/hour + gridded k-pts/s at 256³×91 proj; p50/p99 latency; R-factor Δ vs FP64"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the (box-pruned) y-forward pass of a 256^3 box in a
# 512^3 grid; None where not measured. rocprofv3 --pmc FETCH_SIZE (x2.00: the
# counter reads half of a known 1 GiB read on MI355X) + WRITE_SIZE (x1.00, exact
# on a known 1 GiB fill), in profiles/tilt256_hbm_traffic.md: 65853 KiB
# fetched, 263168 KiB written.
TRAFFIC = {512: 2 * 65853 * 1024 + 263168 * 1024}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(w, projs, ang, sup, mk_centred, model, n_cores):
    """The oracle (C + numpy restatement of the reference CPU path) on a bounded
    sample of the same workload, extrapolated to one whole job."""
    from oracle import oracle as O
    import scipy.fft

    n = int(round(w.dim * w.oversampling))
    vecs = w.vectors
    sample = {}
    # gridding: 2 projections through the full-grid index scan + FP64 direct DFT
    t0 = time.perf_counter()
    npts = 0
    for p in (0, w.num_projs // 2):
        Rf, _ = O.rotation(*ang[p], *vecs)
        ind, D, u, v, _ = O.select(n, n, True, Rf, w.cutoff)
        un = (u / np.float32(n)).astype(np.float32)
        vn = (v / np.float32(n)).astype(np.float32)
        O.dft64(projs[:, :, p], un, vn)
        npts += len(ind)
    t_grid_proj = (time.perf_counter() - t0) / 2
    # iteration: 2 DR iterations' worth of FFT + constraint + replacement on the n^3 grid
    shape = (n, n, n)
    x = np.zeros(shape, np.float32)
    S = np.fft.ifftshift(O.paddzero(sup.astype(np.float32), shape))
    M = np.fft.ifftshift(mk_centred)
    meas = M != 0
    t0 = time.perf_counter()
    for _ in range(2):
        r = np.where(x < 0, np.float32(0), x) * S
        y = 2 * r - x
        k = scipy.fft.fftn(y.astype(np.complex64), workers=n_cores)
        _ = np.sum(np.abs(np.abs(k[meas]) - np.abs(M[meas])))
        k[meas] = M[meas]
        pf = np.real(scipy.fft.ifftn(k, workers=n_cores)).astype(np.float32)
        x = x + pf - r
    t_iter = (time.perf_counter() - t0) / 2
    del x, S, M, meas, k, pf, r, y
    # R-factor: 2 projections of the Dim^3 model with the FP64 trilinear projector
    t0 = time.perf_counter()
    for p in (0, 1):
        O.project(model, *ang[p])
    t_proj = (time.perf_counter() - t0) / 2
    n_eval = w.iterations // w.monitor_every if w.monitor_R else 0
    total = w.num_projs * t_grid_proj + w.iterations * t_iter + n_eval * w.num_projs * t_proj
    return {
        "value": 3600.0 / total,
        "unit": "recon/hour",
        "cores": n_cores,
        "kind": "port",
        "sample": (f"oracle (C/OpenMP + numpy/pocketfft) on 2 of {w.num_projs} projections gridded "
                   f"({t_grid_proj:.2f} s each, full-grid FP32 index scan + FP64 direct DFT), "
                   f"2 DR iterations on the {n}^3 grid ({t_iter:.2f} s each), 2 of {w.num_projs} trilinear "
                   f"projections ({t_proj:.2f} s each); extrapolated linearly to {w.num_projs} projections, "
                   f"{w.iterations} iterations and {n_eval} R-factor evaluations = {total:.0f} s per job"),
    }


def rfactor_fp64_delta(ctx, model, projs, ang, idx):
    """|R_gpu - R_fp64| of the final reconstruction on a few projections: the GPU
    R-factor check (FP32 projector, FP64 sums) against the oracle's FP64 projector
    and sums, same model, same formula (Tian_calc_Rfactor_realspace)."""
    from oracle import oracle as O
    sub_p = np.ascontiguousarray(projs[:, :, idx])
    sub_a = ang[idx]
    r_gpu, _, _ = ctx.rfactor(model, sub_p, sub_a)
    r_64, _, _ = O.rfactor_posthoc(model, sub_p, sub_a)
    return abs(r_gpu - r_64), r_gpu


def box_rows(n: int, b: int, ts: int) -> int:
    """Rows of the ts-row tiles along an axis of length n that meet the centred
    box of extent b (FFT order: [0, b - b//2) and [n - b//2, n))."""
    nt, lo_end, hi_begin = n // ts, (b - b // 2 + ts - 1) // ts, (n - b // 2) // ts
    return n if lo_end >= hi_begin else ts * (lo_end + nt - hi_begin)


def grid_value_error(job, projs, idx, npts=2000, crop=0):
    """Norm-wise relative error of the job's gridded values against the oracle's
    FP64 direct DFT (fillInFourierGrid_DFT.m:243) at the same in-plane
    coordinates, on npts seeded points of each listed projection (with crop > 0,
    only points inside the central crop^3 of k-space)."""
    from oracle import oracle as O
    pts = job.points()
    keep = np.ones(len(pts["ind"]), bool)
    if crop:
        l1, l2, _ = job.half_shape
        i1 = pts["ind"] % l1
        i2 = (pts["ind"] // l1) % l2
        i3 = pts["ind"] // (l1 * l2)
        k = [i1 - job.n1 // 2, i2 - job.n2 // 2, i3 - job.n1 // 2]
        keep = (np.abs(k[0]) < crop // 2) & (np.abs(k[1]) < crop // 2) & (np.abs(k[2]) < crop // 2)
    num = den = 0.0
    for p in idx:
        g = np.flatnonzero((pts["proj"] == p) & keep)
        g = np.random.default_rng(p).choice(g, size=min(npts, len(g)), replace=False)
        F64 = O.dft64(projs[:, :, p], pts["un"][g], pts["vn"][g])
        num += float(np.sum(np.abs(pts["val"][g].astype(np.complex128) - F64) ** 2))
        den += float(np.sum(np.abs(F64) ** 2))
    return (num / den) ** 0.5


def serve_tilt(ctx, p, w, projs, ang, sup, njobs, precision, max_active):
    """The served form on the same workload: njobs identical requests submitted at
    once to the request server (continuous batching, same-shape iterations in
    shared launches); throughput is njobs over the time from the first
    submission to the last result."""
    from genfire_amd import _lib
    cfg = _lib.default_server_config()
    cfg.max_active = max_active
    srv = _lib.Server(ctx, cfg)
    p.precision = precision
    req = _lib.Server.request_inputs(projs, ang, sup)
    # one untimed request first (kernel and allocator warm-up)
    srv.wait(srv.submit(p, req))
    srv.close()
    srv = _lib.Server(ctx, cfg)
    t0 = time.perf_counter()
    ids = [srv.submit(p, req) for _ in range(njobs)]
    srv.drain()
    span = time.perf_counter() - t0
    st = srv.stats()
    res = [srv.result(i) for i in ids]
    errR = [r[2] for r in res]
    srv.close()
    return {
        "jobs": njobs, "max_active": max_active, "recon_per_hour": njobs / span * 3600.0,
        "makespan_s": span, "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
        "batched_launches": st["batched_launches"], "batched_jobs": st["batched_jobs"],
        "peak_footprint_gb": st["peak_footprint_bytes"] / 1e9,
        "rfactor_final": float(errR[0][-1]) if len(errR[0]) else None,
        "max_active_seen": st["max_active_seen"],
        "phases_ms": [{k: round(v, 1) for k, v in r[3].items()} for r in res],
    }


def serve_mix(ctx, nreq, rate, seed, max_active):
    """The serving mix: a seeded Poisson stream of requests of uneven size
    (phantom.serving_mix) submitted at their arrival times; reports jobs/hour
    and p50/p99 time to reconstruction."""
    from genfire_amd import _lib, phantom
    reqs = phantom.serving_mix(nreq, rate, seed)
    pm = phantom.gpu_project_many(ctx)
    t0 = time.perf_counter()
    vols = {}
    inputs = []
    for r in reqs:
        w = phantom.mix_workload(r)
        if r.dim not in vols:   # one phantom per size; tilts, noise and counts differ per request
            rng = np.random.Generator(np.random.PCG64(1000 + r.dim))
            vols[r.dim] = phantom.blob_phantom(r.dim, w.n_blobs, w.sigma, w.radius, rng)
        rng = np.random.Generator(np.random.PCG64(r.seed))
        a = phantom.tilt_angles(w, rng)
        pj = phantom.add_noise(np.asarray(pm(vols[r.dim], a), np.float32), w.noise, rng)
        prm = _lib.default_params()
        prm.griddingMethod = 2   # DFT gridding
        prm.numIterations = w.iterations
        prm.oversamplingRatio = w.oversampling
        prm.interpolationCutoffDistance = w.cutoff
        prm.constraintEnforcementMode = w.mode
        prm.algorithm = _lib.ALG_DR
        for i, v in enumerate(w.vectors):
            for k in range(3):
                getattr(prm, f"vector{i + 1}")[k] = v[k]
        inputs.append((prm, _lib.Server.request_inputs(pj, a, phantom.spherical_support(r.dim, w.support_radius))))
    log(f"serving mix: {nreq} requests prepared in {time.perf_counter() - t0:.1f} s")
    cfg = _lib.default_server_config()
    cfg.max_active = max_active
    srv = _lib.Server(ctx, cfg)
    t0 = time.perf_counter()
    for r, (prm, req) in zip(reqs, inputs):
        dt = r.arrival_s - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        srv.submit(prm, req)
    srv.drain()
    span = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    sizes = {}
    for r in reqs:
        sizes[r.dim] = sizes.get(r.dim, 0) + 1
    return {
        "metric": "serving mix: jobs/hour with p50/p99 time to reconstruction",
        "value": st["jobs_per_hour"], "unit": "jobs/hour", "n_gpus": 1, "higher_is_better": True,
        "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
        "requests": nreq, "arrival_rate_per_s": rate, "span_s": span, "completed": st["completed"],
        "failed": st["failed"], "max_active": max_active, "max_active_seen": st["max_active_seen"],
        "batched_launches": st["batched_launches"], "batched_jobs": st["batched_jobs"],
        "peak_footprint_gb": st["peak_footprint_bytes"] / 1e9, "sizes": sizes,
        "data": "synthetic (one seeded blob phantom per size, per-request tilts, counts and Poisson noise)",
        "config": {"workload": "serving mix", "sizes": list(phantom.MIX_SIZES), "projections": [31, 91],
                   "iterations": [50, 200], "algorithm": "dr", "oversampling": 2.0, "seed": seed},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="tilt256")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-kernels", type=int, default=1,
                    help="HIP-event timing of the FFT pass kernels for the roofline line")
    ap.add_argument("--iterations", type=int, default=None,
                    help="override the workload's iteration count (short profiling runs only)")
    ap.add_argument("--precision", choices=["fp32", "split_bf16", "fp16"], default="fp32",
                    help="gridding GEMM precision (FFT and constraint steps stay FP32)")
    ap.add_argument("--gridding", choices=["dft", "fft"], default="dft",
                    help="gridding method: DFT (griddingMethod 2, the headline path) or FFT (griddingMethod 1)")
    ap.add_argument("--served-jobs", type=int, default=6,
                    help="requests of the served form on the same workload (0: skip)")
    ap.add_argument("--serve-mix", type=int, default=0,
                    help="run only the serving mix with this many requests and print its line")
    ap.add_argument("--mix-rate", type=float, default=4.0)
    ap.add_argument("--max-active", type=int, default=4)
    ap.add_argument("--fp16-inputs", action="store_true",
                    help="store the projections as FP16 with one max-abs scale per projection")
    ap.add_argument("--kspace-crop", type=int, default=0,
                    help="report the FP64 gridded-value error on the central crop^3 of k-space only")
    ap.add_argument("--single-axis", action="store_true",
                    help="tilt series about y only (phi = psi = 0) instead of the workload's three Euler angles")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a 1 GiB device-to-device copy (the HBM-counter calibration kernel)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the library runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    from genfire_amd import _lib, phantom

    if args.calibrate_copy:
        import torch
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 3 x 1 GiB read + 1 GiB write")
        del src, dst

    if args.serve_mix:
        ctx = _lib.Context(0)
        print(json.dumps(serve_mix(ctx, args.serve_mix, args.mix_rate, 7, args.max_active)))
        return

    w = phantom.WORKLOADS[args.workload]
    import dataclasses
    if args.iterations:
        w = dataclasses.replace(w, iterations=args.iterations)
    if args.single_axis:
        w = dataclasses.replace(w, three_euler=False)
    ctx = _lib.Context(0)
    t0 = time.perf_counter()
    vol, projs, ang, sup = phantom.make_inputs(w, phantom.gpu_project_many(ctx))
    log(f"inputs {w.name}: phantom {vol.shape}, projections {projs.shape} in {time.perf_counter() - t0:.1f} s")

    p = _lib.default_params()
    p.griddingMethod = 1 if args.gridding == "fft" else 2
    p.numIterations = w.iterations
    p.oversamplingRatio = w.oversampling
    p.interpolationCutoffDistance = w.cutoff
    p.constraintEnforcementMode = w.mode
    p.algorithm = _lib.ALG_DR if w.algorithm == "dr" else _lib.ALG_GENFIRE
    p.monitor_R = w.monitor_R
    p.monitorR_loopLength = w.monitor_every
    for i, v in enumerate(w.vectors):
        for j in range(3):
            getattr(p, f"vector{i + 1}")[j] = v[j]
    prec = {"fp32": _lib.PREC_FP32, "split_bf16": _lib.PREC_SPLIT_BF16, "fp16": _lib.PREC_FP16}[args.precision]
    p.precision = prec
    job = _lib.Job(ctx, p, w.dim, w.dim, w.num_projs)
    if args.fp16_inputs:
        # FP16 storage with a max-abs scale per projection; the FP32 values the
        # job then sees replace projs everywhere below
        scale = np.maximum(np.abs(projs).reshape(-1, w.num_projs, order="F").max(axis=0), 1e-30).astype(np.float32)
        projs16 = (projs / scale[None, None, :]).astype(np.float16)
        job.set_inputs_f16(projs16, scale, ang, sup)   # host -> HBM, outside the timed region
        projs = (projs16.astype(np.float32) * scale[None, None, :]).astype(np.float32)
    else:
        job.set_inputs(projs, ang, sup)          # host -> HBM, outside the timed region

    lat = []
    stats = None
    for s in range(args.warmup + args.steps):
        timed = s >= args.warmup
        if timed and args.profile_kernels:
            job.set_profiling(True)
        t = time.perf_counter()
        job.grid()
        job.reconstruct()                    # returns after the job's stream drained
        dt = time.perf_counter() - t
        if timed:
            lat.append(dt)
            stats = stats or {}
            for k in range(5):
                n_, ms_ = job.kernel_stats(k)
                a, b = stats.get(k, (0, 0.0))
                stats[k] = (a + n_, b + ms_)
            job.set_profiling(False)
        log(f"step {s} {'timed' if timed else 'warmup'}: {dt * 1e3:.1f} ms {job.timing()}")

    tm = job.timing()
    # DFT gridding: points in the master list; FFT gridding: lattice points of all spectra
    npts = job.num_points() if args.gridding == "dft" else w.num_projs * job.n1 * job.n1
    errR, R1, R2 = job.errR()
    rec = job.reconstruction()
    lat_ms = np.array(lat) * 1e3
    ms_step = float(np.mean(lat_ms))
    n = job.n1
    h = n // 2 + 1
    # P2 (y-forward pass of the half-spectrum iteration) with the support box
    # pruning: it reads the rows with y and z in the Dim box and writes every ky
    # of the z rows its tiles cover (complex64)
    zrows = box_rows(n, w.dim, 16)
    pass_bytes = 8.0 * h * (w.dim * w.dim + n * zrows)
    nfp, msfp = stats[_lib.KSTAT_FFT_PASS]
    nit, msit = stats[_lib.KSTAT_ITERATION]
    pass_ms = msfp / max(nfp, 1)
    achieved = pass_bytes / (pass_ms * 1e-3) / 1e9 if nfp else float("nan")
    it_ms = msit / max(nit, 1) if nit else tm["iterate_ms"] / w.iterations
    # algorithmic bytes per iteration with half-spectrum storage:
    # 16 n^3 + 12 N_meas,half + n^3/8, N_meas,half ~ N_meas / 2
    nmeas = int(np.count_nonzero(job.measuredK()))
    b_it = 16.0 * n ** 3 + 12.0 * (nmeas / 2) + n ** 3 / 8.0

    out = {
        "metric": METRIC,
        "value": 3600.0 / (ms_step * 1e-3),
        "unit": "recon/hour",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded Gaussian-blob phantom, trilinear projections, Poisson 1e4 counts)",
        "config": {"workload": w.name, "volume": w.dim, "projections": w.num_projs, "projection_size": w.dim,
                   "grid": n, "iterations": w.iterations, "algorithm": w.algorithm,
                   "rfactor_every": w.monitor_every, "gridding": args.gridding, "gridding_precision": args.precision,
                   "cutoff": w.cutoff, "oversampling": w.oversampling,
                   "tilt_series": "three Euler angles" if w.three_euler else "single axis (y)",
                   "projection_storage": "fp16 + per-projection scale" if args.fp16_inputs else "fp32"},
        "gridded_kpts_per_s": npts / (tm["grid_ms"] * 1e-3),
        "gridded_points": npts,
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
        "stage_ms": {"grid": tm["grid_ms"], "iterations": tm["iterate_ms"], "rfactor": tm["rfactor_ms"],
                     "iteration_avg": it_ms},
        "rfactor_final": float(errR[-1]) if len(errR) else None,
        "iteration_roofline_frac": (b_it / (it_ms * 1e-3) / 1e9) / HBM_PEAK_GBS,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": TRAFFIC.get(n),
                     "kernel": f"k_half_y_fwd<{n}> (y-forward pass, box-pruned), {pass_bytes / 1e9:.3f} GB algorithmic "
                               f"per launch, {pass_ms:.3f} ms mean over {nfp} launches"},
    }
    if args.served_jobs > 0:
        out["served"] = serve_tilt(ctx, p, w, projs, ang, sup, args.served_jobs, _lib.PREC_FP32,
                                   args.max_active)
        out["served"]["precision"] = "fp32"
        out["served"]["vs_this_line"] = out["served"]["recon_per_hour"] / out["value"]
        p.precision = prec
    if prec != _lib.PREC_FP32:
        # the same job with FP32 gridding, untimed: volume and R-factor deltas
        p.precision = _lib.PREC_FP32
        ref = _lib.Job(ctx, p, w.dim, w.dim, w.num_projs)
        ref.set_inputs(projs, ang, sup)
        ref.grid()
        ref.reconstruct()
        rec32 = ref.reconstruction()
        errR32 = ref.errR()[0]
        mk32 = ref.measuredK()
        mk = job.measuredK()
        out["precision_check"] = {
            "reference": "same job with FP32 gridding on the GPU",
            "measuredK_rel_err": float(np.linalg.norm((mk - mk32).astype(np.complex128))
                                       / np.linalg.norm(mk32.astype(np.complex128))),
            "volume_rel_err": float(np.linalg.norm(rec - rec32) / np.linalg.norm(rec32)),
            "rfactor_delta": float(abs(errR[-1] - errR32[-1])) if len(errR) else None,
        }
        del ref
    if not args.no_cpu_baseline:
        ncores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log("cpu baseline (oracle sample)")
        try:
            out["cpu_baseline"] = cpu_baseline(w, projs, ang, sup, job.measuredK(), rec, ncores)
            d, r = rfactor_fp64_delta(ctx, rec, projs, ang, [0, w.num_projs // 2, w.num_projs - 1])
            out["rfactor_delta_vs_fp64"] = d
            out["rfactor_posthoc_sample"] = r
            if args.gridding == "dft":
                out["grid_value_rel_err_vs_fp64"] = grid_value_error(
                    job, projs, [0, w.num_projs // 2, w.num_projs - 1], crop=args.kspace_crop)
            if args.kspace_crop:
                out["grid_value_error_region"] = f"central {args.kspace_crop}^3 of k-space"
        except Exception as e:  # the baseline is reported, never the target
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
