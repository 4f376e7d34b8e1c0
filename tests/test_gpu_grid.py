"""GPU parity of the DFT gridding (fillInFourierGrid_DFT.m:85-310 + CentroSymmetricity.m)
against the CPU oracle, through the C ABI.

Bar: grid indices, per-point D / u/n1 / v/n2 and per-voxel hit counts bit-exact;
gridded values within 1e-5 norm-wise relative error of the FP64 DFT."""
import numpy as np
import pytest

from genfire_amd import _lib

pytestmark = pytest.mark.gpu

VALUE_TOL = 1e-5


def _params(iters=5, ratio=2.0, cutoff=0.3, centro=1, v3=(1, 0, 0)):
    p = _lib.default_params()
    p.griddingMethod = 2   # DFT gridding
    p.numIterations = iters
    p.oversamplingRatio = ratio
    p.interpolationCutoffDistance = cutoff
    p.DFT_CentroSymmetricity = centro
    p.constraintEnforcementMode = 3
    for i in range(3):
        p.vector3[i] = float(v3[i])
    return p


CASES = {
    # name: (dim1, dim2, ratio, cutoff, centro, vector3, angles)
    "tilt64_single_axis": (64, 64, 2.0, 0.3, 1, (1, 0, 0),
                           np.stack([np.zeros(41), np.arange(-60, 61, 3.0), np.zeros(41)], 1)),
    "three_euler_32": (32, 32, 2.0, 0.3, 1, (0, 0, 1),
                       np.random.default_rng(5).uniform([-180, -70, -180], [180, 70, 180], (9, 3))),
    "odd_15x15_ratio3": (15, 15, 3.0, 0.7, 1, (1, 0, 0),
                         np.random.default_rng(6).uniform([-20, -70, -20], [20, 70, 20], (5, 3))),
    "mixed_16x15_ratio3": (16, 15, 3.0, 0.5, 1, (1, 0, 0),
                           np.random.default_rng(7).uniform([-20, -70, -20], [20, 70, 20], (5, 3))),
    "odd_15x16_ratio3": (15, 16, 3.0, 0.5, 1, (0, 0, 1),
                         np.random.default_rng(8).uniform([-20, -70, -20], [20, 70, 20], (5, 3))),
    "noncentro_16_ratio3": (16, 16, 3.0, 0.4, 0, (1, 0, 0),
                            np.random.default_rng(9).uniform([-20, -70, -20], [20, 70, 20], (4, 3))),
    "exact_planes_45deg": (32, 32, 2.0, 0.3, 1, (1, 0, 0),
                           np.array([[0, 0, 0], [0, 45, 0], [0, -45, 0], [0, 90, 0], [90, 45, 0]], float)),
}


def _run(ctx, name, with_values=True):
    d1, d2, ratio, cutoff, centro, v3, ang = CASES[name]
    P = len(ang)
    rng = np.random.default_rng(11)
    projs = rng.uniform(0.0, 1.0, (d1, d2, P)).astype(np.float32)
    job = _lib.Job(ctx, _params(ratio=ratio, cutoff=cutoff, centro=centro, v3=v3), d1, d2, P)
    job.set_inputs(projs, ang)
    job.grid()
    return job, projs, ang, (d1, d2, ratio, cutoff, centro, v3)


@pytest.mark.parametrize("name", list(CASES))
def test_indices_and_hits_bitexact(ctx, O, name):
    print(f"[grid] indices {name}", flush=True)
    job, projs, ang, (d1, d2, ratio, cutoff, centro, v3) = _run(ctx, name)
    n1, n2 = job.n1, job.n2
    pts = job.points()
    vecs = ((0, 0, 1), (0, 1, 0), v3)
    ref = O.grid_dft(projs, ang, n1, n2, cutoff, centro=bool(centro), vectors=vecs, values=False)
    assert job.num_points() == len(ref.master_ind)
    for p in range(len(ang)):
        g = pts["proj"] == p
        r = ref.master_proj == p
        go = np.argsort(pts["ind"][g])
        np.testing.assert_array_equal(pts["ind"][g][go], ref.master_ind[r])
        np.testing.assert_array_equal(pts["dist"][g][go].view(np.uint32), ref.master_dist[r].view(np.uint32))
        np.testing.assert_array_equal(pts["un"][g][go].view(np.uint32), ref.master_un[r].view(np.uint32))
        np.testing.assert_array_equal(pts["vn"][g][go].view(np.uint32), ref.master_vn[r].view(np.uint32))
    np.testing.assert_array_equal(job.hits(), ref.hits)


@pytest.mark.parametrize("name", ["tilt64_single_axis", "three_euler_32", "odd_15x15_ratio3",
                                  "mixed_16x15_ratio3", "odd_15x16_ratio3", "noncentro_16_ratio3",
                                  "exact_planes_45deg"])
def test_values_and_measuredK(ctx, O, name):
    print(f"[grid] values {name}", flush=True)
    job, projs, ang, (d1, d2, ratio, cutoff, centro, v3) = _run(ctx, name)
    n1, n2 = job.n1, job.n2
    pts = job.points()
    vecs = ((0, 0, 1), (0, 1, 0), v3)
    # per-point values against the FP64 DFT at the same (FP32) in-plane coordinates
    for p in range(len(ang)):
        g = pts["proj"] == p
        F64 = O.dft64(projs[:, :, p], pts["un"][g], pts["vn"][g])
        err = np.linalg.norm(pts["val"][g].astype(np.complex128) - F64) / np.linalg.norm(F64)
        assert err < VALUE_TOL, (p, err)
    ref = O.grid_dft(projs, ang, n1, n2, cutoff, centro=bool(centro), vectors=vecs)
    mk = job.measuredK()
    # same measured set (nonzero pattern) and values within tolerance
    np.testing.assert_array_equal(mk != 0, ref.measuredK != 0)
    err = np.linalg.norm((mk - ref.measuredK).astype(np.complex128)) / np.linalg.norm(ref.measuredK.astype(np.complex128))
    assert err < VALUE_TOL, err
    # the average of the GPU's own values through the oracle's averaging and
    # CentroSymmetricity restatement must agree to FP32 rounding of the cast
    own = O.grid_dft(projs, ang, n1, n2, cutoff, centro=bool(centro), vectors=vecs,
                     value_fn=lambda p, pr, un, vn: _gpu_vals_sorted(pts, p, ref))
    np.testing.assert_allclose(mk, own.measuredK, rtol=0, atol=4e-6 * np.abs(own.measuredK).max())


def _gpu_vals_sorted(pts, p, ref):
    g = pts["proj"] == p
    go = np.argsort(pts["ind"][g])
    return pts["val"][g][go]


@pytest.fixture
def grid_kernel(request, monkeypatch):
    """The FP32 value kernel: the default (type-2 non-uniform FFT where the
    padded grid has a supported FFT length) or the direct sum on MFMA / VALU."""
    if request.param:
        monkeypatch.setenv("GENFIRE_GRID_KERNEL", request.param)
    else:
        monkeypatch.delenv("GENFIRE_GRID_KERNEL", raising=False)
    return request.param or "default"


@pytest.mark.parametrize("grid_kernel", [None, "mfma", "valu"], indirect=True)
@pytest.mark.parametrize("name", ["three_euler_32", "odd_15x15_ratio3", "mixed_16x15_ratio3", "noncentro_16_ratio3"])
def test_values_by_kernel(ctx, O, name, grid_kernel):
    print(f"[grid] values {name} kernel={grid_kernel}", flush=True)
    job, projs, ang, (d1, d2, ratio, cutoff, centro, v3) = _run(ctx, name)
    pts = job.points()
    for p in range(len(ang)):
        g = pts["proj"] == p
        F64 = O.dft64(projs[:, :, p], pts["un"][g], pts["vn"][g])
        err = np.linalg.norm(pts["val"][g].astype(np.complex128) - F64) / np.linalg.norm(F64)
        assert err < VALUE_TOL, (p, err)


@pytest.mark.parametrize("d1,d2", [(256, 256), (96, 48)])
def test_nufft_agrees_with_direct_sum(ctx, monkeypatch, d1, d2):
    # the two FP32 forms of the same values: every point, to 1e-6 relative
    # (each is within about 5e-7 of the FP64 sum)
    print(f"[grid] NUFFT vs MFMA {d1}x{d2}", flush=True)
    ang = np.random.default_rng(24).uniform([-5, -70, -5], [5, 70, 5], (3, 3))
    projs = np.random.default_rng(14).uniform(0.0, 1.0, (d1, d2, len(ang))).astype(np.float32)
    vals = {}
    for kern in ("nufft", "mfma"):
        if kern == "mfma":
            monkeypatch.setenv("GENFIRE_GRID_KERNEL", "mfma")
        else:
            monkeypatch.delenv("GENFIRE_GRID_KERNEL", raising=False)
        job = _lib.Job(ctx, _params(ratio=2.0, cutoff=0.3, v3=(0, 0, 1)), d1, d2, len(ang))
        job.set_inputs(projs, ang)
        job.grid()
        vals[kern] = job.points()["val"].astype(np.complex128)
    a, b = vals["nufft"], vals["mfma"]
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-6


@pytest.mark.parametrize("grid_kernel", [None, "mfma"], indirect=True)
@pytest.mark.parametrize("d1,d2", [(256, 256), (320, 192)])
def test_fp32_values_large_projections(ctx, O, d1, d2, grid_kernel):
    # Projections past 128 rows and columns: the NUFFT's 512 and 640 x 384
    # padded grids, and for the direct sum more than one entry of the coarsest
    # phase tables (rows a = 128i + 16j + k), 320 x 192 ending in a partial
    # 128-column block.  3000 points per projection against the FP64 DFT.
    print(f"[grid] FP32 values {d1}x{d2} kernel={grid_kernel}", flush=True)
    ang = np.random.default_rng(23).uniform([-5, -70, -5], [5, 70, 5], (2, 3))
    projs = np.random.default_rng(13).uniform(0.0, 1.0, (d1, d2, len(ang))).astype(np.float32)
    job = _lib.Job(ctx, _params(ratio=2.0, cutoff=0.3, v3=(0, 0, 1)), d1, d2, len(ang))
    job.set_inputs(projs, ang)
    job.grid()
    pts = job.points()
    for p in range(len(ang)):
        g = np.flatnonzero(pts["proj"] == p)
        g = np.random.default_rng(p).choice(g, size=min(3000, len(g)), replace=False)
        F64 = O.dft64(projs[:, :, p], pts["un"][g], pts["vn"][g])
        err = np.linalg.norm(pts["val"][g].astype(np.complex128) - F64) / np.linalg.norm(F64)
        assert err < VALUE_TOL, (p, err)


def test_recIFFT(ctx, O):
    print("[grid] recIFFT", flush=True)
    job, projs, ang, (d1, d2, *_rest) = _run(ctx, "tilt64_single_axis")
    mk = job.measuredK()
    ref = O.rec_ifft(mk.astype(np.complex128), d1, d2)
    got = job.recIFFT()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 1e-5, err


# Reduced-precision gridding GEMMs (split-BF16 hi+lo, FP16 with per-projection
# max-abs scaling), FP32 accumulation.  The index pass is shared, so indices and
# hits stay bit-exact; the values are held to the stated bounds against the FP64
# DFT (norm-wise relative error over every point of each projection checked).
LP_TOL = {_lib.PREC_SPLIT_BF16: 1.5e-5, _lib.PREC_FP16: 1.5e-3}
LP_CASES = {
    # name: (dim1, dim2, angles, points checked per projection or None for all)
    "tilt64": (64, 64, np.stack([np.zeros(5), np.linspace(-60, 60, 5), np.zeros(5)], 1), None),
    "euler_48x24": (48, 24, np.random.default_rng(21).uniform([-10, -70, -10], [10, 70, 10], (4, 3)), None),
    "euler_256": (256, 256, np.random.default_rng(22).uniform([-5, -70, -5], [5, 70, 5], (2, 3)), 3000),
}


@pytest.mark.parametrize("prec", [_lib.PREC_SPLIT_BF16, _lib.PREC_FP16])
@pytest.mark.parametrize("name", list(LP_CASES))
def test_reduced_precision_values(ctx, O, name, prec):
    print(f"[grid] precision {prec} {name}", flush=True)
    d1, d2, ang, nchk = LP_CASES[name]
    P = len(ang)
    rng = np.random.default_rng(12)
    projs = rng.uniform(0.0, 1.0, (d1, d2, P)).astype(np.float32)
    jobs = {}
    for pr in (_lib.PREC_FP32, prec):
        p = _params(ratio=2.0, cutoff=0.3, v3=(0, 0, 1))
        p.precision = pr
        jobs[pr] = _lib.Job(ctx, p, d1, d2, P)
        jobs[pr].set_inputs(projs, ang)
        jobs[pr].grid()
    a, b = jobs[_lib.PREC_FP32], jobs[prec]
    np.testing.assert_array_equal(a.hits(), b.hits())
    pa, pb = a.points(), b.points()
    np.testing.assert_array_equal(pa["ind"], pb["ind"])
    worst = 0.0
    for p in range(P):
        g = np.flatnonzero(pb["proj"] == p)
        if nchk is not None:
            g = np.random.default_rng(p).choice(g, size=min(nchk, len(g)), replace=False)
        F64 = O.dft64(projs[:, :, p], pb["un"][g], pb["vn"][g])
        err = np.linalg.norm(pb["val"][g].astype(np.complex128) - F64) / np.linalg.norm(F64)
        worst = max(worst, err)
    print(f"[grid] precision {prec} {name}: value error {worst:.2e}", flush=True)
    assert worst < LP_TOL[prec], worst
    mka, mkb = a.measuredK(), b.measuredK()
    np.testing.assert_array_equal(mka != 0, mkb != 0)
    err = np.linalg.norm((mkb - mka).astype(np.complex128)) / np.linalg.norm(mka.astype(np.complex128))
    assert err < LP_TOL[prec], err


def test_fp16_projection_storage(ctx):
    """genfire_job_set_inputs_f16 grids exactly what FP32 inputs of the expanded
    values (half * scale, exact in FP32) grid."""
    print("[grid] fp16 projection storage", flush=True)
    d1, d2, P = 48, 48, 5
    rng = np.random.default_rng(13)
    projs = rng.uniform(0.0, 50.0, (d1, d2, P)).astype(np.float32)
    scale = np.abs(projs).reshape(-1, P, order="F").max(axis=0).astype(np.float32)
    h = (projs / scale[None, None, :]).astype(np.float16)
    expanded = (h.astype(np.float32) * scale[None, None, :]).astype(np.float32)
    ang = np.stack([np.zeros(P), np.linspace(-60, 60, P), np.zeros(P)], 1)
    out = []
    for f16 in (True, False):
        job = _lib.Job(ctx, _params(ratio=2.0, cutoff=0.3), d1, d2, P)
        if f16:
            job.set_inputs_f16(h, scale, ang)
        else:
            job.set_inputs(expanded, ang)
        job.grid()
        out.append(job.measuredK())
    np.testing.assert_array_equal(out[0], out[1])
