"""Parity at the benchmarked sizes, through the C ABI.

* tilt256 (256^3 phantom of 20000 blobs, 91 three-Euler projections of 256^2,
  512^3 grid; the timed workload): the index pass of every projection and the
  per-voxel hit counts bit-exact against the oracle's full-grid FP32 scan
  (fillInFourierGrid_DFT.m:175-192), values against the FP64 direct DFT on 2000
  points of every projection (:236-248), and 10 DR iterations with enforcement
  mode 1 (a schedule stage at every iteration) and the R-factor monitor every 5
  iterations on the 256^3 model against the oracle's FP64 restatement of
  reconstruct.m:117-185.
* 1024^3 grids (tilt512, tilt512x121): indices and hits bit-exact on three
  projections, values against FP64 in FP32 and in split-BF16 from FP16-stored
  projections; the box-pruned GENFIRE passes bit-identical to the unpruned ones;
  the half-spectrum passes against the full-spectrum (complex) passes; and a
  GENFIRE fixed point at 1024^3 with M = fftn(x*) formed on the host.
* serving mix: requests of 192^3 to 320^3 with 31 to 91 projections, each
  result bit-identical alone and in the mix.
"""
import numpy as np
import pytest
import scipy.fft

from genfire_amd import _lib, phantom

pytestmark = pytest.mark.gpu


def _params(w, iters, mode, alg, monitor=0, L=10, prec=_lib.PREC_FP32):
    p = _lib.default_params()
    p.griddingMethod = 2   # DFT gridding
    p.numIterations = iters
    p.oversamplingRatio = w.oversampling
    p.interpolationCutoffDistance = w.cutoff
    p.constraintEnforcementMode = mode
    p.algorithm = _lib.ALG_DR if alg == "dr" else _lib.ALG_GENFIRE
    p.monitor_R = monitor
    p.monitorR_loopLength = L
    p.precision = prec
    for i, v in enumerate(w.vectors):
        for k in range(3):
            getattr(p, f"vector{i + 1}")[k] = v[k]
    return p


def _check_indices(O, job, ang, vectors, cutoff, projs_checked):
    """Per projection: sorted half-grid indices, D, u/n1 and v/n2 bit-exact
    against the oracle's whole-grid scan; returns the oracle's hit counts over
    the projections checked."""
    n1, n2 = job.n1, job.n2
    pts = job.points()
    hits = np.zeros(int(np.prod(job.half_shape)), np.int32)
    for p in projs_checked:
        R, _ = O.rotation(*ang[p], *vectors)
        ind, D, u, v, _ = O.select(n1, n2, True, R, cutoff)
        un = (u / np.float32(n1)).astype(np.float32)
        vn = (v / np.float32(n2)).astype(np.float32)
        g = np.flatnonzero(pts["proj"] == p)
        go = g[np.argsort(pts["ind"][g])]
        assert len(go) == len(ind), (p, len(go), len(ind))
        np.testing.assert_array_equal(pts["ind"][go], ind)
        np.testing.assert_array_equal(pts["dist"][go].view(np.uint32), D.view(np.uint32))
        np.testing.assert_array_equal(pts["un"][go].view(np.uint32), un.view(np.uint32))
        np.testing.assert_array_equal(pts["vn"][go].view(np.uint32), vn.view(np.uint32))
        hits[ind] += 1          # find() gives each index once per projection
    return pts, hits


def _value_error(O, pts, projs, projs_checked, npts=2000):
    num = den = 0.0
    worst = 0.0
    for p in projs_checked:
        g = np.flatnonzero(pts["proj"] == p)
        g = np.random.default_rng(p).choice(g, size=min(npts, len(g)), replace=False)
        F64 = O.dft64(projs[:, :, p], pts["un"][g], pts["vn"][g])
        e2 = float(np.sum(np.abs(pts["val"][g].astype(np.complex128) - F64) ** 2))
        d2 = float(np.sum(np.abs(F64) ** 2))
        worst = max(worst, (e2 / d2) ** 0.5)
        num += e2
        den += d2
    return (num / den) ** 0.5, worst


# ---------------------------------------------------------------------------
# tilt256: the timed workload
# ---------------------------------------------------------------------------

TILT256_ITERS = 10


@pytest.fixture(scope="module")
def tilt256(ctx):
    w = phantom.WORKLOADS["tilt256"]
    print("[headline] tilt256 inputs", flush=True)
    vol, projs, ang, sup = phantom.make_inputs(w, phantom.gpu_project_many(ctx))
    job = _lib.Job(ctx, _params(w, TILT256_ITERS, 1, "dr", monitor=1, L=5), w.dim, w.dim, w.num_projs)
    job.set_inputs(projs, ang, sup)
    job.grid()
    print(f"[headline] tilt256 gridded: {job.num_points()} points on the {job.n1}^3 grid", flush=True)
    return w, projs, ang, sup, job


def test_tilt256_indices_and_hits_bitexact(O, tilt256):
    w, projs, ang, sup, job = tilt256
    print(f"[headline] tilt256 index pass, {w.num_projs} projections", flush=True)
    assert job.n1 == 512 and job.n2 == 512
    pts, hits = _check_indices(O, job, ang, w.vectors, w.cutoff, range(w.num_projs))
    assert len(pts["ind"]) == int(hits.sum())
    np.testing.assert_array_equal(job.hits(), hits)


def test_tilt256_values_vs_fp64(O, tilt256):
    w, projs, ang, sup, job = tilt256
    print("[headline] tilt256 values, 2000 points of every projection", flush=True)
    err, worst = _value_error(O, job.points(), projs, range(w.num_projs))
    print(f"[headline] tilt256 value error {err:.2e} (worst projection {worst:.2e})", flush=True)
    assert worst < 1e-5, worst


def test_tilt256_iterations_vs_fp64(O, tilt256):
    """10 DR iterations on the 512^3 grid (half-box z pass at 512, box-pruned
    passes), enforcement mode 1 and the monitor every 5 iterations."""
    w, projs, ang, sup, job = tilt256
    print(f"[headline] tilt256 {TILT256_ITERS} DR iterations", flush=True)
    job.reconstruct()
    mk = job.measuredK()
    ref = O.reconstruct(mk, mk != 0, sup, w.dim, w.dim, TILT256_ITERS, mode=1, algorithm="dr",
                        precision="fp64", monitor_R=True, loop_length=5, projections=projs, angles=ang)
    del mk
    rec = job.reconstruction()
    err = np.linalg.norm(rec - ref.reconstruction) / np.linalg.norm(ref.reconstruction)
    errR, R1, R2 = job.errR()
    print(f"[headline] tilt256 volume error {err:.2e}, errK max diff "
          f"{np.abs(job.errK() - ref.errK).max():.2e}, errR {errR} vs {ref.errR}", flush=True)
    assert err < 1e-4, err
    np.testing.assert_allclose(job.errK(), ref.errK, atol=2e-5, rtol=0)
    np.testing.assert_allclose(errR, ref.errR, atol=1e-4, rtol=0)
    np.testing.assert_allclose(R1, ref.Rarr_record, atol=1e-4, rtol=0)
    np.testing.assert_allclose(R2, ref.Rarr2_record, atol=1e-4, rtol=0)


# ---------------------------------------------------------------------------
# 1024^3 grids
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("name,prec,f16", [("tilt512", _lib.PREC_FP32, False),
                                           ("tilt512x121", _lib.PREC_SPLIT_BF16, True)])
def test_n1024_indices_hits_values(ctx, O, name, prec, f16):
    """Three projections of the workload's tilt series (first, middle, last) on
    the 1024^3 grid.  tilt512x121 stores them as FP16 with one scale per
    projection and grids in split BF16; its values are checked against the FP64
    DFT of the expanded FP32 values."""
    w = phantom.WORKLOADS[name]
    print(f"[headline] {name} indices on the 1024^3 grid", flush=True)
    rng = np.random.Generator(np.random.PCG64(w.seed))
    ang_all = phantom.tilt_angles(w, rng)
    sel = [0, w.num_projs // 2, w.num_projs - 1]
    ang = ang_all[sel]
    projs = np.random.default_rng(31).uniform(0.0, 1.0, (w.dim, w.dim, len(sel))).astype(np.float32)
    job = _lib.Job(ctx, _params(w, 1, 3, "dr", prec=prec), w.dim, w.dim, len(sel))
    if f16:
        scale = np.abs(projs).reshape(-1, len(sel), order="F").max(axis=0).astype(np.float32)
        h = (projs / scale[None, None, :]).astype(np.float16)
        job.set_inputs_f16(h, scale, ang)
        projs = (h.astype(np.float32) * scale[None, None, :]).astype(np.float32)
    else:
        job.set_inputs(projs, ang)
    job.grid()
    assert job.n1 == 1024
    pts, hits = _check_indices(O, job, ang, w.vectors, w.cutoff, range(len(sel)))
    np.testing.assert_array_equal(job.hits(), hits)
    err, worst = _value_error(O, pts, projs, range(len(sel)))
    tol = 1e-5 if prec == _lib.PREC_FP32 else 1.5e-5
    print(f"[headline] {name} value error {err:.2e} (worst {worst:.2e})", flush=True)
    assert worst < tol, worst


@pytest.fixture(scope="module")
def n1024_measured(ctx):
    """measuredK of 21 tilt512 projections of a 512^3 blob phantom on the 1024^3
    grid, gridded by the library (its indices and values are checked above)."""
    w = phantom.WORKLOADS["tilt512"]
    print("[headline] 1024^3 measuredK from 21 projections", flush=True)
    rng = np.random.Generator(np.random.PCG64(5))
    vol = phantom.blob_phantom(w.dim, 4000, w.sigma, w.radius, rng)
    ang = phantom.tilt_angles(w, rng)[::max(1, w.num_projs // 21)][:21]
    projs = np.asarray(phantom.gpu_project_many(ctx)(vol, ang), np.float32)
    sup = phantom.spherical_support(w.dim, w.support_radius)
    job = _lib.Job(ctx, _params(w, 1, 3, "dr"), w.dim, w.dim, len(ang))
    job.set_inputs(projs, ang, sup)
    job.grid()
    mk = job.measuredK()
    job.close()
    return w, projs, ang, sup, mk


def _run_1024(ctx, w, projs, ang, sup, mk, iters, alg, monitor=0, L=5):
    job = _lib.Job(ctx, _params(w, iters, 1, alg, monitor=monitor, L=L), w.dim, w.dim, projs.shape[2])
    job.set_inputs(projs, ang, sup)
    job.set_measuredK(mk)
    job.reconstruct()
    out = (job.reconstruction(), job.errK(), job.errR()[0])
    job.close()
    return out


def test_n1024_pruned_passes_bit_identical(ctx, n1024_measured, monkeypatch):
    """GENFIRE on the 1024^3 grid: the box-pruned half-spectrum passes (half-box
    z pass at 1024) compute the same numbers as the passes over every row."""
    w, projs, ang, sup, mk = n1024_measured
    print("[headline] 1024^3 pruned vs unpruned GENFIRE", flush=True)
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("GENFIRE_PRUNE", flag)
        out[flag] = _run_1024(ctx, w, projs, ang, sup, mk, 6, "genfire", monitor=1, L=3)
    for a, b in zip(out["1"], out["0"]):
        np.testing.assert_array_equal(a, b)


def test_n1024_half_matches_full_spectrum(ctx, n1024_measured, monkeypatch):
    """DR on the 1024^3 grid: the half-spectrum passes against the full-spectrum
    complex passes (a different transform and replacement, so rounding
    differs)."""
    w, projs, ang, sup, mk = n1024_measured
    print("[headline] 1024^3 half vs full spectrum DR", flush=True)
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("GENFIRE_HALF", flag)
        out[flag] = _run_1024(ctx, w, projs, ang, sup, mk, 6, "dr", monitor=1, L=3)
    rec_h, rec_f = out["1"][0], out["0"][0]
    err = np.linalg.norm(rec_h - rec_f) / np.linalg.norm(rec_f)
    print(f"[headline] 1024^3 half vs full: volume {err:.2e}", flush=True)
    assert err < 1e-5, err
    np.testing.assert_allclose(out["1"][1], out["0"][1], atol=2e-6, rtol=0)
    np.testing.assert_allclose(out["1"][2], out["0"][2], atol=1e-5, rtol=0)


def test_n1024_fixed_point(ctx):
    """M = fftn(x*) of a positive object in a 512^3 support, every voxel of the
    1024^3 grid measured and enforced: GENFIRE returns x* and errK falls to
    rounding.  The spectrum is formed on the host in C order with the axes
    reversed, which is the column-major layout of the library."""
    print("[headline] 1024^3 fixed point", flush=True)
    dim, n = 512, 1024
    rng = np.random.default_rng(17)
    c0, c1 = dim // 2 - 96, dim // 2 + 96
    obj_c = np.zeros((dim, dim, dim), np.float32)          # C order: [z, y, x]
    obj_c[c0:c1, c0:c1, c0:c1] = rng.uniform(0.1, 1.0, (c1 - c0,) * 3).astype(np.float32)
    big = np.zeros((n, n, n), np.complex64)
    s = n // 2 - dim // 2
    big[s:s + dim, s:s + dim, s:s + dim] = obj_c
    big = np.fft.ifftshift(big)
    big = scipy.fft.fftn(big, workers=-1, overwrite_x=True)
    big = np.fft.fftshift(big)
    mk = big.T                                             # column-major (x, y, z) view
    del big
    w = phantom.WORKLOADS["tilt512"]
    p = _params(w, 3, 3, "genfire")
    job = _lib.Job(ctx, p, dim, dim, 3)
    job.set_inputs(np.ones((dim, dim, 3), np.float32), np.zeros((3, 3)), (obj_c > 0).T)
    job.set_measuredK(mk)
    del mk
    job.reconstruct()
    rec = job.reconstruction()
    obj = obj_c.T
    d = float(np.abs(rec - obj).max())
    print(f"[headline] 1024^3 fixed point: max |x - x*| {d:.2e}, errK {job.errK()}", flush=True)
    assert d < 1e-4 * float(obj.max()), d
    assert job.errK()[-1] < 1e-5
    job.close()


# ---------------------------------------------------------------------------
# serving mix at its larger sizes
# ---------------------------------------------------------------------------

# (dim, projections, three Euler, iterations, seed); iterations cut from the
# mix's 50-200 to keep the test short (the arithmetic per iteration is the
# same at any count)
MIX_LARGE = [
    (192, 31, True, 24, 101),
    (256, 91, True, 20, 102),
    (320, 61, False, 16, 103),
    (256, 45, False, 28, 104),
    (192, 77, True, 18, 105),
    (320, 38, True, 22, 106),
]


def test_serving_mix_large_sizes_alone_and_in_mix(ctx):
    print("[headline] serving mix 192-320, 31-91 projections", flush=True)
    pm = phantom.gpu_project_many(ctx)
    reqs = []
    for dim, P, euler, iters, seed in MIX_LARGE:
        w = phantom.mix_workload(phantom.MixRequest(0.0, dim, P, euler, iters, seed))
        rng = np.random.Generator(np.random.PCG64(seed))
        vol = phantom.blob_phantom(dim, w.n_blobs, w.sigma, w.radius, rng)
        a = phantom.tilt_angles(w, rng)
        pj = phantom.add_noise(np.asarray(pm(vol, a), np.float32), w.noise, rng)
        prm = _params(w, iters, w.mode, "dr", monitor=1, L=4)
        reqs.append((pj, a, phantom.spherical_support(dim, w.support_radius), prm))
    alone = []
    for pj, a, sup, prm in reqs:
        d1, d2, P = pj.shape
        job = _lib.Job(ctx, prm, d1, d2, P)
        job.set_inputs(pj, a, sup)
        job.grid()
        job.reconstruct()
        alone.append((job.reconstruction(), job.errK(), job.errR()[0]))
        job.close()
    print("[headline] alone runs done", flush=True)
    cfg = _lib.default_server_config()
    cfg.max_active = 4
    srv = _lib.Server(ctx, cfg)
    ids = [srv.submit(prm, pj, a, sup) for pj, a, sup, prm in reqs]
    srv.drain()
    for rid, (rec0, ek0, er0), m in zip(ids, alone, MIX_LARGE):
        assert srv.status(rid) == _lib.REQ_DONE, m
        rec, ek, er, lat = srv.result(rid)
        print(f"[headline] request {m}: {lat['total']:.0f} ms", flush=True)
        np.testing.assert_array_equal(rec, rec0)
        np.testing.assert_array_equal(ek, ek0)
        np.testing.assert_array_equal(er, er0)
    st = srv.stats()
    print(f"[headline] serving stats {st}", flush=True)
    assert st["completed"] == len(MIX_LARGE) and st["failed"] == 0
    srv.close()
