"""GPU GENFIRE / DR iteration (reconstruct.m:28-194) against the oracle's numpy
restatement on the same measuredK, through the C ABI."""
import numpy as np
import pytest

from genfire_amd import _lib

pytestmark = pytest.mark.gpu


def _setup(O, dim=32, ratio=2.0, P=15, seed=0, dim2=None, euler=False, phi_psi=0.0):
    """Blob phantom, trilinear tilt series and the oracle's measuredK.  dim2 != dim
    makes the projections (and the grid) non-square: vol is dim x dim2 x dim.
    euler adds phi and psi in [-5, 5] degrees to the tilts about y."""
    dim2 = dim if dim2 is None else dim2
    rng = np.random.default_rng(seed)
    vol = np.zeros((dim, dim2, dim), np.float32)
    k1 = np.arange(dim) - dim // 2
    k2 = np.arange(dim2) - dim2 // 2
    for _ in range(12):
        x0 = rng.uniform(-min(dim, dim2) / 5, min(dim, dim2) / 5, 3)
        s = rng.uniform(1.0, 2.0)
        g = [np.exp(-((kk - q) ** 2) / (2 * s * s)) for kk, q in zip((k1, k2, k1), x0)]
        vol += (rng.uniform(0.5, 1) * g[0][:, None, None] * g[1][None, :, None] * g[2][None, None, :]).astype(np.float32)
    ang = np.stack([np.zeros(P), np.linspace(-60, 60, P), np.zeros(P)], 1)
    if euler:
        ang[:, 0] = phi_psi + rng.uniform(-5, 5, P)
        ang[:, 2] = phi_psi + rng.uniform(-5, 5, P)
    projs = np.stack([O.project(vol, *a) for a in ang], 2).astype(np.float32)
    n1, n2 = int(round(dim * ratio)), int(round(dim2 * ratio))
    g = O.grid_dft(projs, ang, n1, n2, 0.3)
    r = min(dim, dim2) * 0.45
    sup = (k1[:, None, None] ** 2 + k2[None, :, None] ** 2 + k1[None, None, :] ** 2) <= r * r
    return vol, projs, ang, g.measuredK, sup


def _job(ctx, mk, projs, ang, sup, iters, mode=3, alg="genfire", monitor=0, L=5, support=1, init=None):
    d1, d2, P = projs.shape
    p = _lib.default_params()
    p.griddingMethod = 2   # DFT gridding
    p.numIterations = iters
    p.oversamplingRatio = mk.shape[0] / d1
    p.constraintSupport = support
    p.interpolationCutoffDistance = 0.3
    p.constraintEnforcementMode = mode
    p.algorithm = _lib.ALG_DR if alg == "dr" else _lib.ALG_GENFIRE
    p.monitor_R = monitor
    p.monitorR_loopLength = L
    job = _lib.Job(ctx, p, d1, d2, P)
    job.set_inputs(projs, ang, sup, init)
    job.set_measuredK(mk)
    job.reconstruct()
    return job


@pytest.mark.parametrize("alg", ["genfire", "dr"])
@pytest.mark.parametrize("mode", [3, 1])
def test_iteration_matches_oracle(ctx, O, alg, mode):
    print(f"[iter] alg={alg} mode={mode}", flush=True)
    vol, projs, ang, mk, sup = _setup(O)
    iters = 30
    job = _job(ctx, mk, projs, ang, sup, iters, mode=mode, alg=alg, monitor=1, L=10)
    ref = O.reconstruct(mk, mk != 0, sup, 32, 32, iters, mode=mode, algorithm=alg, precision="fp64",
                        monitor_R=True, loop_length=10, projections=projs, angles=ang)
    rec = job.reconstruction()
    err = np.linalg.norm(rec - ref.reconstruction) / np.linalg.norm(ref.reconstruction)
    assert err < 1e-4, err
    np.testing.assert_allclose(job.errK(), ref.errK, atol=2e-5, rtol=0)
    errR, R1, R2 = job.errR()
    np.testing.assert_allclose(errR, ref.errR, atol=1e-4, rtol=0)
    np.testing.assert_allclose(R1, ref.Rarr_record, atol=1e-4, rtol=0)
    np.testing.assert_allclose(R2, ref.Rarr2_record, atol=1e-4, rtol=0)


@pytest.mark.parametrize("alg", ["genfire", "dr"])
@pytest.mark.parametrize("shape", [(64, 64, 20), (64, 32, 20), (96, 80, 12), (160, 32, 10), (192, 32, 8),
                                   (320, 32, 5)])
def test_half_spectrum_iteration_matches_oracle(ctx, O, alg, shape):
    """Grids n1 x n2 x n1 with n = 2 dim run the half-spectrum (R2C/C2R) passes:
    128^3 and 128 x 64 x 128 (radix 8), 192 x 160 x 192 (x and z with radix 6
    and 3, y with radix 10), 320 x 64 x 320 (radix 10 / 5), 384 x 64 x 384
    (radix 3 / 6), 640 x 64 x 640 (radix 5 / 10)."""
    d1, d2, iters = shape
    print(f"[iter] half spectrum alg={alg} dims={shape}", flush=True)
    vol, projs, ang, mk, sup = _setup(O, dim=d1, dim2=d2, P=13, seed=3)
    job = _job(ctx, mk, projs, ang, sup, iters, mode=1, alg=alg, monitor=1, L=10)
    ref = O.reconstruct(mk, mk != 0, sup, d1, d2, iters, mode=1, algorithm=alg, precision="fp64",
                        monitor_R=True, loop_length=10, projections=projs, angles=ang)
    rec = job.reconstruction()
    err = np.linalg.norm(rec - ref.reconstruction) / np.linalg.norm(ref.reconstruction)
    assert err < 1e-4, err
    np.testing.assert_allclose(job.errK(), ref.errK, atol=2e-5, rtol=0)
    errR, R1, R2 = job.errR()
    np.testing.assert_allclose(errR, ref.errR, atol=1e-4, rtol=0)


@pytest.mark.parametrize("alg,shape", [("genfire", (256, 32, 10)), ("dr", (256, 32, 10)),
                                       ("genfire", (64, 256, 10)), ("dr", (64, 256, 10)),
                                       ("dr", (512, 32, 6)), ("genfire", (64, 512, 6)), ("dr", (64, 512, 6))])
def test_half_spectrum_large_axes_match_oracle(ctx, O, alg, shape):
    """The pass lengths of the 512^3 and 1024^3 grids on grids the FP64 oracle
    finishes in seconds: 512 x 64 x 512 and 1024 x 64 x 1024 (x pass at 256 and
    512 complex points, the half-box z pass at 512 and 1024, whose tiles need
    more than 64 KiB of LDS at 1024), 128 x 512 x 128 and 128 x 1024 x 128 (the
    y passes at 512 and 1024).  The box is Dim = n/2 along every axis, as in the
    benchmarked workloads; R-factor monitor every 5 or 3 iterations."""
    d1, d2, iters = shape
    L = 5 if iters >= 10 else 3
    print(f"[iter] large axes alg={alg} dims={shape}", flush=True)
    vol, projs, ang, mk, sup = _setup(O, dim=d1, dim2=d2, P=13, seed=3)
    job = _job(ctx, mk, projs, ang, sup, iters, mode=1, alg=alg, monitor=1, L=L)
    ref = O.reconstruct(mk, mk != 0, sup, d1, d2, iters, mode=1, algorithm=alg, precision="fp64",
                        monitor_R=True, loop_length=L, projections=projs, angles=ang)
    rec = job.reconstruction()
    err = np.linalg.norm(rec - ref.reconstruction) / np.linalg.norm(ref.reconstruction)
    assert err < 1e-4, err
    np.testing.assert_allclose(job.errK(), ref.errK, atol=2e-5, rtol=0)
    errR, R1, R2 = job.errR()
    np.testing.assert_allclose(errR, ref.errR, atol=1e-4, rtol=0)
    np.testing.assert_allclose(R1, ref.Rarr_record, atol=1e-4, rtol=0)
    np.testing.assert_allclose(R2, ref.Rarr2_record, atol=1e-4, rtol=0)


def test_fixed_point(ctx, O):
    """M = fftn of a positive supported object, every point enforced: the object is
    reproduced and errK -> 0 (GENFIRE fixed point)."""
    print("[iter] fixed point", flush=True)
    dim, n = 32, 64
    rng = np.random.default_rng(2)
    obj = np.zeros((dim, dim, dim), np.float32)
    obj[8:24, 8:24, 8:24] = rng.uniform(0.1, 1.0, (16, 16, 16))
    sup = obj > 0
    M = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(O.paddzero(obj, (n, n, n))))).astype(np.complex64)
    P = 3
    projs = np.ones((dim, dim, P), np.float32)
    ang = np.zeros((P, 3))
    job = _job(ctx, M, projs, ang, sup, 20)
    rec = job.reconstruction()
    assert np.abs(rec - obj).max() < 1e-4 * np.abs(obj).max()
    assert job.errK()[-1] < 1e-5


@pytest.mark.parametrize("dims", [(64, 64), (32, 32)])
def test_graph_replays_match_direct_launches(ctx, O, dims, monkeypatch):
    """Steady-state iterations replay one captured HIP graph; with enforcement
    stages and R-factor evaluations launched directly between the replays, every
    errK entry and the result equal those of direct launches only (128^3 half
    spectrum and 64^3 full spectrum)."""
    d1, d2 = dims
    print(f"[iter] graph replays dims={dims}", flush=True)
    vol, projs, ang, mk, sup = _setup(O, dim=d1, dim2=d2, P=11, seed=4)
    out = {}
    for g in ("1", "0"):
        monkeypatch.setenv("GENFIRE_GRAPHS", g)
        job = _job(ctx, mk, projs, ang, sup, 40, mode=3, alg="dr", monitor=1, L=7)
        out[g] = (job.reconstruction(), job.errK(), job.errR()[0], job.graph_replays())
    assert out["1"][3] > 10 and out["0"][3] == 0, (out["1"][3], out["0"][3])
    for a, b in zip(out["1"][:3], out["0"][:3]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("alg", ["genfire", "dr"])
@pytest.mark.parametrize("shape", [(64, 64), (96, 80), (80, 96)])
def test_box_pruning_matches_full_passes(ctx, O, alg, shape, monkeypatch):
    """With the support on, the half-spectrum passes skip real-space rows outside
    the Dim box.  GENFIRE computes the same numbers as with every row
    transformed; DR, carried as x = P_R(x) + real(ifftn(D)), agrees to rounding."""
    d1, d2 = shape
    print(f"[iter] box pruning alg={alg} dims={shape}", flush=True)
    vol, projs, ang, mk, sup = _setup(O, dim=d1, dim2=d2, P=13, seed=6)
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("GENFIRE_PRUNE", flag)
        job = _job(ctx, mk, projs, ang, sup, 30, mode=1, alg=alg, monitor=1, L=10)
        out[flag] = (job.reconstruction(), job.errK(), job.errR()[0])
    if alg == "genfire":
        for a, b in zip(out["1"], out["0"]):
            np.testing.assert_array_equal(a, b)
    else:
        rec1, rec0 = out["1"][0], out["0"][0]
        assert np.linalg.norm(rec1 - rec0) / np.linalg.norm(rec0) < 1e-5
        np.testing.assert_allclose(out["1"][1], out["0"][1], atol=2e-6, rtol=0)
        np.testing.assert_allclose(out["1"][2], out["0"][2], atol=1e-5, rtol=0)


@pytest.mark.parametrize("dim,ratio,phi_psi", [(32, 2.0, 0.0), (33, 64 / 33, 0.0), (32, 2.0, 60.0), (33, 64 / 33, 60.0)])
def test_monitor_three_euler_matches_oracle(ctx, O, dim, ratio, phi_psi):
    """The R-factor monitor on tilts with phi and psi off zero (one ray per lane
    on the z-pair copies of the model): an even volume edge and an odd one,
    whose last even z pair ends in the zero padding; phi = psi = 60 degrees
    puts most projections on the x-fastest copies."""
    print(f"[iter] monitor three Euler dim={dim} phi_psi={phi_psi}", flush=True)
    vol, projs, ang, mk, sup = _setup(O, dim=dim, ratio=ratio, P=13, seed=9, euler=True, phi_psi=phi_psi)
    iters = 20
    job = _job(ctx, mk, projs, ang, sup, iters, mode=1, alg="dr", monitor=1, L=5)
    ref = O.reconstruct(mk, mk != 0, sup, dim, dim, iters, mode=1, algorithm="dr", precision="fp64",
                        monitor_R=True, loop_length=5, projections=projs, angles=ang)
    rec = job.reconstruction()
    assert np.linalg.norm(rec - ref.reconstruction) / np.linalg.norm(ref.reconstruction) < 1e-4
    _, R1, R2 = job.errR()
    np.testing.assert_allclose(R1, ref.Rarr_record, atol=1e-4, rtol=0)
    np.testing.assert_allclose(R2, ref.Rarr2_record, atol=1e-4, rtol=0)


@pytest.mark.parametrize("alg", ["genfire", "dr"])
@pytest.mark.parametrize("case", ["odd_box_63_in_128", "support_off_64"])
def test_box_edges_match_oracle(ctx, O, alg, case):
    """An odd box (63 of 128: real pairs straddle the box edge, y and z tiles are
    cut by it) and the support constraint off (nothing pruned, every row
    transformed) against the oracle's full-grid restatement."""
    print(f"[iter] box edges alg={alg} case={case}", flush=True)
    if case == "odd_box_63_in_128":
        dim, ratio, support = 63, 128 / 63, 1
    else:
        dim, ratio, support = 64, 2.0, 0
    vol, projs, ang, mk, sup = _setup(O, dim=dim, ratio=ratio, P=11, seed=8)
    assert mk.shape == (128, 128, 128)
    iters = 20
    job = _job(ctx, mk, projs, ang, sup, iters, mode=1, alg=alg, monitor=1, L=10, support=support)
    ref = O.reconstruct(mk, mk != 0, sup, dim, dim, iters, mode=1, algorithm=alg, precision="fp64",
                        support_on=bool(support), monitor_R=True, loop_length=10, projections=projs,
                        angles=ang)
    rec = job.reconstruction()
    err = np.linalg.norm(rec - ref.reconstruction) / np.linalg.norm(ref.reconstruction)
    assert err < 1e-4, err
    np.testing.assert_allclose(job.errK(), ref.errK, atol=2e-5, rtol=0)
    np.testing.assert_allclose(job.errR()[0], ref.errR, atol=1e-4, rtol=0)


@pytest.mark.parametrize("alg", ["genfire", "dr"])
def test_mode2_schedule_matches_oracle(ctx, O, alg):
    """constraintEnforcementMode 2 (GENFIRE_Reconstructor.m:160-161: the
    resolution cutoff falls from 0.95 and then holds all voxels) on the pruned
    half-spectrum path."""
    print(f"[iter] mode 2 alg={alg}", flush=True)
    vol, projs, ang, mk, sup = _setup(O, dim=64, P=11, seed=10)
    iters = 28
    job = _job(ctx, mk, projs, ang, sup, iters, mode=2, alg=alg, monitor=1, L=7)
    ref = O.reconstruct(mk, mk != 0, sup, 64, 64, iters, mode=2, algorithm=alg, precision="fp64",
                        monitor_R=True, loop_length=7, projections=projs, angles=ang)
    rec = job.reconstruction()
    err = np.linalg.norm(rec - ref.reconstruction) / np.linalg.norm(ref.reconstruction)
    assert err < 1e-4, err
    np.testing.assert_allclose(job.errK(), ref.errK, atol=2e-5, rtol=0)
    np.testing.assert_allclose(job.errR()[0], ref.errR, atol=1e-4, rtol=0)


@pytest.mark.parametrize("alg", ["genfire", "dr"])
def test_initial_object_matches_oracle(ctx, O, alg):
    """initialObject (reconstruct.m:76-80) on the box-pruned half-spectrum path:
    the first transform starts from the padded object, not from zero."""
    print(f"[iter] initial object alg={alg}", flush=True)
    vol, projs, ang, mk, sup = _setup(O, dim=64, P=11, seed=9)
    init = (0.5 * vol + np.random.default_rng(3).uniform(0, 0.1, vol.shape)).astype(np.float32)
    iters = 15
    job = _job(ctx, mk, projs, ang, sup, iters, mode=1, alg=alg, monitor=1, L=5, init=init)
    ref = O.reconstruct(mk, mk != 0, sup, 64, 64, iters, mode=1, algorithm=alg, precision="fp64",
                        initial_object=init, monitor_R=True, loop_length=5, projections=projs, angles=ang)
    rec = job.reconstruction()
    err = np.linalg.norm(rec - ref.reconstruction) / np.linalg.norm(ref.reconstruction)
    assert err < 1e-4, err
    np.testing.assert_allclose(job.errK(), ref.errK, atol=2e-5, rtol=0)
    np.testing.assert_allclose(job.errR()[0], ref.errR, atol=1e-4, rtol=0)


@pytest.mark.parametrize("alg", ["genfire", "dr"])
@pytest.mark.parametrize("shape", [(64, 64, 15), (96, 80, 10)])
def test_rfree_matches_oracle(ctx, O, alg, shape):
    """Rfree (reconstruct.m:46-69,137-147): the same withheld set as the oracle's
    restatement (same seeded permutation), the masked reconstruction and errK,
    and every Rfree_complex(shell, iteration) within 1e-4."""
    d1, d2, iters = shape
    print(f"[iter] Rfree alg={alg} dims={shape}", flush=True)
    vol, projs, ang, mk, sup = _setup(O, dim=d1, dim2=d2, P=13, seed=5)
    p = _lib.default_params()
    p.griddingMethod = 2   # DFT gridding
    p.numIterations = iters
    p.oversamplingRatio = mk.shape[0] / d1
    p.interpolationCutoffDistance = 0.3
    p.constraintEnforcementMode = 1
    p.algorithm = _lib.ALG_DR if alg == "dr" else _lib.ALG_GENFIRE
    p.calculate_Rfree = 1
    p.numBinsRfree = 8
    p.percentValuesForRfree = 0.05
    p.RandomSeed = 11
    job = _lib.Job(ctx, p, d1, d2, projs.shape[2])
    job.set_inputs(projs, ang, sup)
    job.set_measuredK(mk)
    job.reconstruct()
    ref = O.reconstruct(mk, mk != 0, sup, d1, d2, iters, mode=1, algorithm=alg, precision="fp64",
                        rfree=(8, 0.05, 11))
    rec = job.reconstruction()
    err = np.linalg.norm(rec - ref.reconstruction) / np.linalg.norm(ref.reconstruction)
    assert err < 1e-4, err
    np.testing.assert_allclose(job.errK(), ref.errK, atol=2e-5, rtol=0)
    rf = job.Rfree()
    assert rf.shape == ref.Rfree.shape
    np.testing.assert_array_equal(rf < 0, ref.Rfree < 0)
    np.testing.assert_allclose(rf, ref.Rfree, atol=1e-4, rtol=0)


@pytest.mark.parametrize("alg", ["genfire", "dr"])
@pytest.mark.parametrize("dim", [64, 32])
def test_snapshots_match_oracle(ctx, O, alg, dim):
    """save_temp snapshots (reconstruct.m:149-159) every 3 iterations through
    genfire_job_set_snapshots, on the half-spectrum path (128^3 grid) and the
    full-spectrum path (64^3), against the oracle's best-since-stage rule; the
    returned reconstruction equals the one without snapshots bit for bit."""
    print(f"[iter] snapshots alg={alg} dim={dim}", flush=True)
    vol, projs, ang, mk, sup = _setup(O, dim=dim, P=11, seed=12)
    iters, L = 20, 3
    d1, d2, P = projs.shape
    p = _lib.default_params()
    p.griddingMethod = 2   # DFT gridding
    p.numIterations = iters
    p.oversamplingRatio = 2.0
    p.interpolationCutoffDistance = 0.3
    p.constraintEnforcementMode = 3
    p.algorithm = _lib.ALG_DR if alg == "dr" else _lib.ALG_GENFIRE
    job = _lib.Job(ctx, p, d1, d2, P)
    job.set_inputs(projs, ang, sup)
    job.set_measuredK(mk)
    got = []
    job.set_snapshots(L, lambda it, v: got.append((it, v)))
    job.reconstruct()
    rec_snap = job.reconstruction()
    job.set_snapshots(0)
    job.reconstruct()
    np.testing.assert_array_equal(rec_snap, job.reconstruction())
    ref = O.reconstruct(mk, mk != 0, sup, d1, d2, iters, mode=3, algorithm=alg, precision="fp64", save_every=L)
    assert [t for t, _ in got] == [t for t, _ in ref.snapshots] == list(range(L, iters + 1, L))
    for (t, v), (_, w) in zip(got, ref.snapshots):
        err = np.linalg.norm(v - w) / np.linalg.norm(w)
        assert err < 1e-4, (t, err)


def test_save_temp_writes_mat_files(ctx, O, tmp_path):
    """The Python mirror's save_temp writes temp_reconstruction to
    [saveFilename num2str(it) '.mat'] every save_loopLength iterations
    (reconstruct.m:155-159), readable by scipy.io.loadmat."""
    import scipy.io
    from genfire_amd.reconstructor import GENFIRE_Reconstructor
    print("[iter] save_temp files", flush=True)
    vol, projs, ang, mk, sup = _setup(O, dim=32, P=9, seed=13)
    G = GENFIRE_Reconstructor()
    G.InputProjections, G.InputAngles, G.Support = projs, ang, sup
    G.griddingMethod, G.oversamplingRatio, G.interpolationCutoffDistance = 2, 2.0, 0.3
    G.constraintEnforcementMode, G.numIterations = 3, 10
    G.save_temp, G.save_loopLength = 1, 4
    G.saveFilename = str(tmp_path / "res" / "temp_rec")
    G = G.CheckPrepareData().runGridding().reconstruct()
    names = sorted(f.name for f in (tmp_path / "res").iterdir())
    assert names == ["temp_rec4.mat", "temp_rec8.mat"], names
    t8 = scipy.io.loadmat(str(tmp_path / "res" / "temp_rec8.mat"))["temp_reconstruction"]
    assert t8.shape == (32, 32, 32)
    ref = O.reconstruct(G.measuredK, G.measuredK != 0, sup, 32, 32, 10, mode=3, precision="fp64", save_every=4)
    assert np.linalg.norm(t8 - ref.snapshots[1][1]) / np.linalg.norm(ref.snapshots[1][1]) < 1e-4
