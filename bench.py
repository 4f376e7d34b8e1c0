#!/usr/bin/env python3
"""Benchmark of the GENFIRE reconstruction hot path on one MI355X.

A step is one whole reconstruction job of the tilt256 workload with its inputs
already resident in HBM: DFT gridding of 91 projections (256 x 256) onto the
512^3 oversampled grid, 200 Douglas-Rachford iterations, and the real-space
R-factor of all 91 projections every 10 iterations.  The timed steps run the
library's default path (steady-state iterations replay a captured HIP graph);
per-kernel HIP-event timings come from one more, untimed step.  The metric is
reconstructions per hour; gridded k-points per second, p50/p99 job latency and
the final R-factor are reported beside it, with the served forms (FP32 and
reduced-precision gridding), a serving-mix summary, the CPU baseline and the
FP64-oracle checks.

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
# HBM/MALL bytes per evaluation of the R-factor projector (same counters and
# factors; an evaluation is one launch per z slab), keyed by (workload,
# three-Euler series); None where not measured.  tilt256 (three Euler angles,
# two slabs): FETCH_SIZE 3,657,621 KiB (x2.00) + WRITE_SIZE 43,669 KiB per
# evaluation, profiles/project_tilt256_traffic.md.
PROJ_TRAFFIC = {("tilt256", True): 2 * 3657621 * 1024 + 43669 * 1024}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_sample(w, projs, ang, sup, mk_centred, model, threads, n_grid, n_iter, n_proj):
    """Seconds per unit of the oracle's restatement of the reference CPU path on
    `threads` cores: per gridded projection (full-grid FP32 index scan, then the
    values by the CPU NUFFT and, reported beside it, by the FP64 direct DFT),
    per DR iteration on the n^3 grid, per projected R-factor projection."""
    from oracle import oracle as O
    import scipy.fft

    O.set_threads(threads)
    n = int(round(w.dim * w.oversampling))
    t_scan = t_nufft = t_dft = 0.0
    for p in np.linspace(0, w.num_projs - 1, n_grid).astype(int):
        t0 = time.perf_counter()
        Rf, _ = O.rotation(*ang[p], *w.vectors)
        ind, D, u, v, _ = O.select(n, n, True, Rf, w.cutoff)
        un = (u / np.float32(n)).astype(np.float32)
        vn = (v / np.float32(n)).astype(np.float32)
        t1 = time.perf_counter()
        O.nufft2d2(projs[:, :, p], un, vn, workers=threads)
        t2 = time.perf_counter()
        O.dft64(projs[:, :, p], un, vn)
        t3 = time.perf_counter()
        t_scan += t1 - t0
        t_nufft += t2 - t1
        t_dft += t3 - t2
    shape = (n, n, n)
    x = np.zeros(shape, np.float32)
    S = np.fft.ifftshift(O.paddzero(sup.astype(np.float32), shape))
    M = np.fft.ifftshift(mk_centred)
    meas = M != 0
    Mm = M[meas]
    t0 = time.perf_counter()
    for _ in range(n_iter):
        r = np.where(x < 0, np.float32(0), x) * S
        y = 2 * r - x
        k = scipy.fft.fftn(y.astype(np.complex64), workers=threads)
        _ = np.sum(np.abs(np.abs(k[meas]) - np.abs(Mm)))
        k[meas] = Mm
        pf = np.real(scipy.fft.ifftn(k, workers=threads)).astype(np.float32)
        x = x + pf - r
    t_iter = (time.perf_counter() - t0) / n_iter
    del x, S, M, meas, Mm, k, pf, r, y
    t0 = time.perf_counter()
    for p in range(n_proj):
        O.project(model, *ang[p])
    t_proj = (time.perf_counter() - t0) / n_proj
    return {"scan": t_scan / n_grid, "nufft": t_nufft / n_grid, "dft": t_dft / n_grid, "iter": t_iter,
            "proj": t_proj}


def cpu_baseline(w, projs, ang, sup, mk_centred, model, n_cores):
    """The oracle (C/OpenMP + numpy/pocketfft restatement of the reference CPU
    path) on a bounded sample of the same workload, extrapolated to one whole
    job: on n_cores threads (3 projections gridded, 2 iterations, 3 projector
    calls) and on one core (1, 1, 1)."""
    n = int(round(w.dim * w.oversampling))
    n_eval = w.iterations // w.monitor_every if w.monitor_R else 0

    def job_s(t, values):
        return (w.num_projs * (t["scan"] + t[values]) + w.iterations * t["iter"] +
                n_eval * w.num_projs * t["proj"])

    tm = _cpu_sample(w, projs, ang, sup, mk_centred, model, n_cores, 3, 2, 3)
    t1 = _cpu_sample(w, projs, ang, sup, mk_centred, model, 1, 1, 1, 1)
    from oracle import oracle as O
    O.set_threads(n_cores)
    total = job_s(tm, "nufft")
    return {
        "value": 3600.0 / total,
        "unit": "recon/hour",
        "cores": n_cores,
        "kind": "port",
        "cpu_model": _cpu_model(),
        "nproc": os.cpu_count(),
        "value_1core": 3600.0 / job_s(t1, "nufft"),
        "value_direct_dft": 3600.0 / job_s(tm, "dft"),
        "sample": (f"oracle (C/OpenMP + numpy/pocketfft restatement of the reference CPU path) on {n_cores} threads: "
                   f"3 of {w.num_projs} projections gridded (full-grid FP32 index scan {tm['scan']:.2f} s, "
                   f"values by the CPU type-2 NUFFT {tm['nufft']:.2f} s, by the FP64 direct DFT {tm['dft']:.2f} s "
                   f"per projection), 2 DR iterations on the {n}^3 grid ({tm['iter']:.2f} s each), 3 trilinear "
                   f"projections ({tm['proj']:.3f} s each); extrapolated linearly to {w.num_projs} projections, "
                   f"{w.iterations} iterations and {n_eval} R-factor evaluations = {total:.0f} s per job with NUFFT "
                   f"values ({job_s(tm, 'dft'):.0f} s with the direct DFT); one core: 1 projection, 1 iteration, "
                   f"1 projector call = {job_s(t1, 'nufft'):.0f} s per job"),
    }


def fp64_iteration_check(ctx, p, w, projs, ang, sup, mk_centred, iters):
    """The timed workload's first `iters` DR iterations with the R-factor monitor
    at every iteration, on the GPU and in the oracle's FP64 restatement of
    reconstruct.m:117-185 from the same measuredK: relative error of the
    returned volume and the largest |errR| and |errK| differences."""
    from genfire_amd import _lib
    from oracle import oracle as O
    import ctypes
    q = _lib.GenfireParams()
    ctypes.memmove(ctypes.addressof(q), ctypes.addressof(p), ctypes.sizeof(q))
    q.numIterations = iters
    q.monitor_R = 1
    q.monitorR_loopLength = 1
    job = _lib.Job(ctx, q, w.dim, w.dim, w.num_projs)
    job.set_inputs(projs, ang, sup)
    job.set_measuredK(mk_centred)
    job.reconstruct()
    mode = int(p.constraintEnforcementMode)
    ref = O.reconstruct(mk_centred, mk_centred != 0, sup, w.dim, w.dim, iters, mode=mode, algorithm=w.algorithm,
                        precision="fp64", monitor_R=True, loop_length=1, projections=projs, angles=ang)
    rec = job.reconstruction()
    errR = job.errR()[0]
    out = {
        "iterations": iters,
        "volume_rel_err": float(np.linalg.norm(rec - ref.reconstruction) / np.linalg.norm(ref.reconstruction)),
        "rfactor_max_abs_delta": float(np.abs(errR - ref.errR).max()),
        "errK_max_abs_delta": float(np.abs(job.errK() - ref.errK).max()),
        "reference": "oracle FP64 restatement (numpy/pocketfft), same measuredK",
    }
    job.close()
    return out


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


def fp64_regrid(job, projs, width=16):
    """The gridding rerun on the CPU in FP64 at every point of the job's master
    list (fillInFourierGrid_DFT.m:236-291): values by the oracle's type-2 NUFFT
    with a width-16 kernel (agrees with the FP64 direct sum to 1e-12; the direct
    sum is out of reach at these sizes), then the weighted average and
    CentroSymmetricity in FP64.  Returns (points, FP64 values, measuredK as the
    complex64 grid a job takes)."""
    from oracle import oracle as O
    pts = job.points()
    proj = pts["proj"]
    order = np.argsort(proj, kind="stable")
    bounds = np.searchsorted(proj[order], np.arange(projs.shape[2] + 1))
    v64 = np.empty(len(proj), np.complex128)
    for p in range(projs.shape[2]):
        g = order[bounds[p]:bounds[p + 1]]
        if len(g):
            v64[g] = O.nufft2d2(projs[:, :, p], pts["un"][g], pts["vn"][g], w=width)
    total = int(np.prod(job.half_shape))
    FS = O.weighted_average(pts["ind"], pts["dist"], v64, total)
    mk = O.centro_symmetricity(FS, job.n1, job.n1 // 2, job.n2, dtype=np.complex64)
    return pts, v64, mk


def crop_mask(pts, job, crop):
    """Points of the master list inside the central crop^3 of k-space."""
    l1, l2, _ = job.half_shape
    i1 = pts["ind"] % l1
    i2 = (pts["ind"] // l1) % l2
    i3 = pts["ind"] // (l1 * l2)
    k = [i1 - job.n1 // 2, i2 - job.n2 // 2, i3 - job.n1 // 2]
    return (np.abs(k[0]) < crop // 2) & (np.abs(k[1]) < crop // 2) & (np.abs(k[2]) < crop // 2)


def rel(a, b):
    return float(np.linalg.norm((a - b).astype(np.complex128)) / np.linalg.norm(b.astype(np.complex128)))


# The larger configurations, one reconstruction each beside the headline line:
# (workload, gridding precisions timed, FP16 projection storage, k-space crop of
# the value error)
EXTRA_CONFIGS = {
    "tilt512": (("split_bf16", "fp16", "fp32"), False, 0),
    "tilt512x121": (("split_bf16",), True, 128),
}


def extra_config(ctx, name, steps=1):
    """One configuration end to end: each listed gridding precision timed over
    `steps` reconstructions after one warm-up (grid + 200/300 DR iterations +
    R-factor monitor, inputs resident in HBM), then the FP64 regridding check:
    relative error of the gridded values over every point (and in the k-space
    crop), and of the final volume and R-factor against the same FP32
    iterations started from the FP64-gridded measuredK."""
    from genfire_amd import _lib, phantom
    precs_named, fp16_inputs, crop = EXTRA_CONFIGS[name]
    precs = {"fp32": _lib.PREC_FP32, "split_bf16": _lib.PREC_SPLIT_BF16, "fp16": _lib.PREC_FP16}
    w = phantom.WORKLOADS[name]
    t0 = time.perf_counter()
    vol, projs, ang, sup = phantom.make_inputs(w, phantom.gpu_project_many(ctx))
    del vol
    log(f"{name}: inputs in {time.perf_counter() - t0:.1f} s")
    p = _lib.default_params()
    p.griddingMethod = 2
    p.numIterations = w.iterations
    p.oversamplingRatio = w.oversampling
    p.interpolationCutoffDistance = w.cutoff
    p.constraintEnforcementMode = w.mode
    p.algorithm = _lib.ALG_DR if w.algorithm == "dr" else _lib.ALG_GENFIRE
    p.monitor_R = w.monitor_R
    p.monitorR_loopLength = w.monitor_every
    for i, vv in enumerate(w.vectors):
        for j in range(3):
            getattr(p, f"vector{i + 1}")[j] = vv[j]
    if fp16_inputs:
        scale = np.maximum(np.abs(projs).reshape(-1, w.num_projs, order="F").max(axis=0), 1e-30).astype(np.float32)
        projs16 = (projs / scale[None, None, :]).astype(np.float16)
        projs = (projs16.astype(np.float32) * scale[None, None, :]).astype(np.float32)
    out = {"workload": name, "volume": w.dim, "projections": w.num_projs, "grid": None, "iterations": w.iterations,
           "algorithm": w.algorithm, "rfactor_every": w.monitor_every,
           "projection_storage": "fp16 + per-projection scale" if fp16_inputs else "fp32", "forms": {}}
    recs = {}
    vals = {}
    first = None
    for sp in precs_named:
        p.precision = precs[sp]
        job = _lib.Job(ctx, p, w.dim, w.dim, w.num_projs)
        if fp16_inputs:
            job.set_inputs_f16(projs16, scale, ang, sup)
        else:
            job.set_inputs(projs, ang, sup)
        lat = []
        for s in range(1 + steps):
            t = time.perf_counter()
            job.grid()
            job.reconstruct()
            if s:
                lat.append(time.perf_counter() - t)
        tm = job.timing()
        errR = job.errR()[0]
        recs[sp] = (job.reconstruction(), float(errR[-1]) if len(errR) else None)
        ms = float(np.mean(lat)) * 1e3
        out["grid"] = job.n1
        out["forms"][sp] = {"recon_per_hour": 3600.0 / (ms * 1e-3), "ms_per_recon": ms,
                            "gridded_kpts_per_s": job.num_points() / (tm["grid_ms"] * 1e-3),
                            "stage_ms": {"grid": tm["grid_ms"], "iterations": tm["iterate_ms"],
                                         "rfactor": tm["rfactor_ms"]},
                            "values": values_route(argparse.Namespace(gridding="dft"), sp),
                            "graph_replays": job.graph_replays(), "rfactor_final": recs[sp][1]}
        log(f"{name} {sp}: {ms:.0f} ms per reconstruction, {tm}")
        vals[sp] = job.points()["val"]
        if first is None:
            first = job          # the FP64 regridding reads its master list
        else:
            job.close()
    try:
        t = time.perf_counter()
        pts, v64, mk = fp64_regrid(first, projs)
        log(f"{name}: FP64 regridding of {len(v64)} points in {time.perf_counter() - t:.0f} s")
        keep = crop_mask(pts, first, crop) if crop else None
        first.close()
        p.precision = _lib.PREC_FP32
        ref = _lib.Job(ctx, p, w.dim, w.dim, w.num_projs)
        ref.set_inputs(projs, ang, sup)
        ref.set_measuredK(mk)
        del mk
        ref.reconstruct()
        rec64 = ref.reconstruction()
        errR64 = ref.errR()[0]
        ref.close()
        chk = {"reference": "gridding rerun on the CPU in FP64 at every master-list point (type-2 NUFFT, "
                            "width-16 kernel, within 1e-12 of the direct sum; FP64 weighted average and "
                            "centro-symmetry), then the same FP32 GPU iterations from that measuredK",
               "points": int(len(v64))}
        for sp, v32 in vals.items():
            chk[f"grid_value_rel_err_{sp}"] = rel(v32, v64)
            if crop:
                chk[f"grid_value_rel_err_{sp}_crop{crop}"] = rel(v32[keep], v64[keep])
        for sp, (rec, rf) in recs.items():
            chk[f"volume_rel_err_{sp}"] = float(np.linalg.norm(rec - rec64) / np.linalg.norm(rec64))
            if rf is not None and len(errR64):
                chk[f"rfactor_delta_{sp}"] = abs(rf - float(errR64[-1]))
        out["fp64_check"] = chk
    except Exception as e:  # reported, never the target
        out["fp64_check"] = {"error": str(e)}
    return out


def proj_slabs(d: int) -> int:
    """Launches of the trilinear projector per evaluation of a d^3 model: its
    z-duplicated form (8 (d+1)^2 d bytes) is swept in z slabs of at most 128 MiB
    once it passes 128 MiB (the library's project_slabs, without the
    GENFIRE_PROJECTOR_SLABS override)."""
    form, cap = 8 * (d + 1) * (d + 1) * d, 128 << 20
    if form <= cap:
        return 1
    n = min(d, -(-form // cap))
    t = -(-d // n)
    return -(-d // t)


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
        "graph_replays": st["graph_replays"], "peak_footprint_gb": st["peak_footprint_bytes"] / 1e9,
        "rfactor_final": float(errR[0][-1]) if len(errR[0]) else None,
        "max_active_seen": st["max_active_seen"],
        "phases_ms": [{k: round(v, 1) for k, v in r[3].items()} for r in res],
    }


def mix_requests(ctx, nreq, rate, seed):
    """The requests of the serving mix (phantom.serving_mix): a seeded Poisson
    stream of uneven sizes, one blob phantom per size, per-request tilts, counts
    and noise; returns (requests, [(params, request inputs)])."""
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
        for i, vv in enumerate(w.vectors):
            for k in range(3):
                getattr(prm, f"vector{i + 1}")[k] = vv[k]
        inputs.append((prm, _lib.Server.request_inputs(pj, a, phantom.spherical_support(r.dim, w.support_radius))))
    log(f"serving mix: {nreq} requests prepared in {time.perf_counter() - t0:.1f} s")
    return reqs, inputs


def serve_mix(ctx, nreq, rate, seed, max_active, saturate=False, prepared=None):
    """The serving mix submitted at its arrival times (saturate: all at once,
    the server's capacity on the mix); reports jobs/hour and p50/p99 time to
    reconstruction."""
    from genfire_amd import _lib, phantom
    reqs, inputs = prepared if prepared is not None else mix_requests(ctx, nreq, rate, seed)
    cfg = _lib.default_server_config()
    cfg.max_active = max_active
    srv = _lib.Server(ctx, cfg)
    t0 = time.perf_counter()
    for r, (prm, req) in zip(reqs, inputs):
        dt = r.arrival_s - (time.perf_counter() - t0)
        if dt > 0 and not saturate:
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
        "requests": nreq, "arrival_rate_per_s": None if saturate else rate, "span_s": span,
        "completed": st["completed"], "failed": st["failed"], "max_active": max_active,
        "max_active_seen": st["max_active_seen"],
        "batched_launches": st["batched_launches"], "batched_jobs": st["batched_jobs"],
        "graph_replays": st["graph_replays"], "peak_footprint_gb": st["peak_footprint_bytes"] / 1e9,
        "sizes": sizes,
        "data": "synthetic (one seeded blob phantom per size, per-request tilts, counts and Poisson noise)",
        "config": {"workload": "serving mix", "sizes": list(phantom.MIX_SIZES), "projections": [31, 91],
                   "iterations": [50, 200], "algorithm": "dr", "oversampling": 2.0, "seed": seed},
    }


def values_route(args, prec):
    """How the job's gridded values are computed (the library's choice)."""
    if args.gridding == "fft":
        return "fft gridding (no per-point values)"
    forced = os.environ.get("GENFIRE_GRID_KERNEL")
    if prec == "fp32":
        return {"mfma": "direct sum on MFMA (FP32)", "valu": "direct sum on VALU (FP32)"}.get(
            forced, "type-2 NUFFT (FP32, exponential-of-semicircle kernel, width 8)")
    return f"direct sum on MFMA ({prec})"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="tilt256")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--iterations", type=int, default=None,
                    help="override the workload's iteration count (short profiling runs only)")
    ap.add_argument("--precision", choices=["fp32", "split_bf16", "fp16"], default="fp32",
                    help="gridding GEMM precision (FFT and constraint steps stay FP32)")
    ap.add_argument("--gridding", choices=["dft", "fft"], default="dft",
                    help="gridding method: DFT (griddingMethod 2, the headline path) or FFT (griddingMethod 1)")
    ap.add_argument("--served-jobs", type=int, default=24,
                    help="requests of the served form on the same workload, submitted at once (0: skip)")
    ap.add_argument("--served-precision", choices=["fp32", "split_bf16", "fp16"], default="split_bf16",
                    help="gridding precision of the second served line (the mixed-precision served form)")
    ap.add_argument("--serve-mix", type=int, default=0,
                    help="run only the serving mix with this many requests and print its line")
    ap.add_argument("--mix-summary", type=int, default=48,
                    help="requests of the serving-mix summary appended to the line (0: skip)")
    ap.add_argument("--mix-rate", type=float, default=4.0)
    ap.add_argument("--max-active", type=int, default=4)
    ap.add_argument("--fp64-iterations", type=int, default=3,
                    help="iterations of the GPU-vs-FP64-oracle reconstruction check (0: skip)")
    ap.add_argument("--fp16-inputs", action="store_true",
                    help="store the projections as FP16 with one max-abs scale per projection")
    ap.add_argument("--kspace-crop", type=int, default=0,
                    help="report the FP64 gridded-value error on the central crop^3 of k-space only")
    ap.add_argument("--single-axis", action="store_true",
                    help="tilt series about y only (phi = psi = 0) instead of the workload's three Euler angles")
    ap.add_argument("--extra-configs", default="tilt512,tilt512x121",
                    help="comma-separated larger configurations run once each beside the line ('' for none)")
    ap.add_argument("--projector", choices=["linear", "cubic"], default="linear",
                    help="the R-factor monitor's interpolant: trilinear, or interp3 'cubic' (cubic convolution)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a 1 GiB device-to-device copy (the HBM-counter calibration kernel)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU: --gpus must be 1"}))
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
    precs = {"fp32": _lib.PREC_FP32, "split_bf16": _lib.PREC_SPLIT_BF16, "fp16": _lib.PREC_FP16}
    prec = precs[args.precision]
    p.precision = prec
    p.projectorInterpolation = _lib.INTERP_CUBIC if args.projector == "cubic" else _lib.INTERP_LINEAR
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

    # timed steps: the library's default path (steady-state iterations replay a
    # captured HIP graph), inputs resident in HBM
    lat = []
    for s in range(args.warmup + args.steps):
        timed = s >= args.warmup
        t = time.perf_counter()
        job.grid()
        job.reconstruct()                    # returns after the job's stream drained
        dt = time.perf_counter() - t
        if timed:
            lat.append(dt)
        log(f"step {s} {'timed' if timed else 'warmup'}: {dt * 1e3:.1f} ms {job.timing()}")
    tm = job.timing()
    replays = job.graph_replays()
    # per-kernel HIP-event timing in one more, untimed step (profiling launches
    # every iteration directly)
    job.set_profiling(True)
    job.grid()
    job.reconstruct()
    stats = {k: job.kernel_stats(k) for k in range(5)}
    job.set_profiling(False)
    log(f"profiled step: {job.timing()}")

    # DFT gridding: points in the master list; FFT gridding: lattice points of all spectra
    npts = job.num_points() if args.gridding == "dft" else w.num_projs * job.n1 * job.n1
    errR, R1, R2 = job.errR()
    rec = job.reconstruction()
    lat_ms = np.array(lat) * 1e3
    ms_step = float(np.mean(lat_ms))
    n = job.n1
    h = n // 2 + 1
    # dominant kernel: the projector of the R-factor monitor, P (4 Dim^3 + 8 Dim1 Dim2)
    # algorithmic bytes per launch (every projection reads the model once and its
    # measured projection once, writing nothing back but R sums)
    npj, mspj = stats[_lib.KSTAT_PROJECT]
    proj_bytes = float(w.num_projs) * (4.0 * w.dim ** 3 + 8.0 * w.dim * w.dim)
    proj_ms = mspj / max(npj, 1)
    proj_achieved = proj_bytes / (proj_ms * 1e-3) / 1e9 if npj else float("nan")
    # P2 (y-forward pass of the half-spectrum iteration) with the support box
    # pruning: it reads the rows with y and z in the Dim box and writes every ky
    # of the z rows its tiles cover (complex64)
    zrows = box_rows(n, w.dim, 16)
    pass_bytes = 8.0 * h * (w.dim * w.dim + n * zrows)
    nfp, msfp = stats[_lib.KSTAT_FFT_PASS]
    pass_ms = msfp / max(nfp, 1)
    pass_achieved = pass_bytes / (pass_ms * 1e-3) / 1e9 if nfp else float("nan")
    nit, msit = stats[_lib.KSTAT_ITERATION]
    it_ms = msit / max(nit, 1) if nit else tm["iterate_ms"] / w.iterations
    # algorithmic bytes per iteration with half-spectrum storage:
    # 16 n^3 + 12 N_meas,half + n^3/8, N_meas,half ~ N_meas / 2
    nmeas = int(np.count_nonzero(job.measuredK()))
    b_it = 16.0 * n ** 3 + 12.0 * (nmeas / 2) + n ** 3 / 8.0
    proj_traffic = PROJ_TRAFFIC.get((w.name, w.three_euler))

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
                   "rfactor_every": w.monitor_every, "projector": args.projector,
                   "gridding": args.gridding, "gridding_precision": args.precision,
                   "values": values_route(args, args.precision),
                   "cutoff": w.cutoff, "oversampling": w.oversampling,
                   "tilt_series": "three Euler angles" if w.three_euler else "single axis (y)",
                   "projection_storage": "fp16 + per-projection scale" if args.fp16_inputs else "fp32",
                   "iterations_path": f"HIP graph replay for {replays} of {w.iterations} iterations"},
        "gridded_kpts_per_s": npts / (tm["grid_ms"] * 1e-3),
        "gridded_points": npts,
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
        "stage_ms": {"grid": tm["grid_ms"], "iterations": tm["iterate_ms"], "rfactor": tm["rfactor_ms"],
                     "iteration_avg": tm["iterate_ms"] / max(w.iterations, 1),
                     "iteration_avg_profiled": it_ms},
        "rfactor_final": float(errR[-1]) if len(errR) else None,
        "iteration_roofline_frac": (b_it / (it_ms * 1e-3) / 1e9) / HBM_PEAK_GBS,
        "roofline": {"bound": "hbm", "achieved": proj_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": proj_achieved / HBM_PEAK_GBS, "traffic": proj_traffic,
                     "kernel": f"k_project (R-factor projector, the largest kernel; {proj_slabs(w.dim)} z-slab "
                               f"launch(es) per evaluation), {proj_bytes / 1e9:.3f} GB algorithmic per evaluation "
                               f"(P (4 Dim^3 + 8 Dim^2)), {proj_ms:.3f} ms mean per evaluation over {npj} "
                               f"evaluations"},
        "roofline_passes": {"bound": "hbm", "achieved": pass_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": pass_achieved / HBM_PEAK_GBS, "traffic": TRAFFIC.get(n),
                            "kernel": f"k_half_y_fwd<{n}> (y-forward pass, box-pruned), {pass_bytes / 1e9:.3f} GB "
                                      f"algorithmic per launch, {pass_ms:.3f} ms mean over {nfp} launches"},
    }
    if args.served_jobs > 0:
        served = []
        for sp in ("fp32", args.served_precision):
            r = serve_tilt(ctx, p, w, projs, ang, sup, args.served_jobs, precs[sp], args.max_active)
            r["precision"] = sp
            r["vs_this_line"] = r["recon_per_hour"] / out["value"]
            served.append(r)
        out["served"] = served[0]
        out["served_mixed"] = served[1]
        if served[0]["rfactor_final"] is not None and served[1]["rfactor_final"] is not None:
            out["served_mixed"]["rfactor_delta_vs_fp32"] = abs(served[1]["rfactor_final"] - served[0]["rfactor_final"])
        p.precision = prec
    if args.mix_summary > 0:
        keys = ("value", "unit", "p50_ms", "p99_ms", "mean_ms", "requests", "arrival_rate_per_s", "completed",
                "failed", "max_active_seen", "batched_launches", "batched_jobs", "graph_replays", "sizes")
        prepared = mix_requests(ctx, args.mix_summary, args.mix_rate, 7)
        m = serve_mix(ctx, args.mix_summary, args.mix_rate, 7, args.max_active, prepared=prepared)
        out["serving_mix"] = {k: m[k] for k in keys}
        # the same requests submitted at once: the server's capacity on the mix
        m = serve_mix(ctx, args.mix_summary, args.mix_rate, 7, args.max_active, saturate=True, prepared=prepared)
        out["serving_mix_saturated"] = {k: m[k] for k in keys}
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
        p.precision = prec
        del ref
    if not args.no_cpu_baseline:
        ncores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log("cpu baseline (oracle sample)")
        try:
            mk = job.measuredK()
            out["cpu_baseline"] = cpu_baseline(w, projs, ang, sup, mk, rec, ncores)
            d, r = rfactor_fp64_delta(ctx, rec, projs, ang, [0, w.num_projs // 2, w.num_projs - 1])
            out["rfactor_delta_vs_fp64"] = d
            out["rfactor_posthoc_sample"] = r
            if args.gridding == "dft":
                out["grid_value_rel_err_vs_fp64"] = grid_value_error(
                    job, projs, [0, w.num_projs // 2, w.num_projs - 1], crop=args.kspace_crop)
            if args.kspace_crop:
                out["grid_value_error_region"] = f"central {args.kspace_crop}^3 of k-space"
            if args.fp64_iterations > 0:
                log(f"FP64 oracle check: {args.fp64_iterations} iterations")
                out["fp64_iteration_check"] = fp64_iteration_check(ctx, p, w, projs, ang, sup, mk,
                                                                   args.fp64_iterations)
        except Exception as e:  # the baseline is reported, never the target
            out["cpu_baseline"] = out.get("cpu_baseline") or {"error": str(e)}
            out["fp64_check_error"] = str(e)
    extras = [c for c in args.extra_configs.split(",") if c]
    if extras:
        job.close()
        out["configs_extra"] = {c: extra_config(ctx, c) for c in extras}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
