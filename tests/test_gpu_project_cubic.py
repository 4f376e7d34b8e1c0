"""The cubic-convolution projector (interp3 'cubic',
src/calculate3Dprojection_RealSpaceinterp.m:45,51) on the GPU against the
oracle's FP64 restatement (Keys' a = -1/2 kernel and boundary nodes), through
the C ABI: genfire_project_cubic, genfire_rfactor_cubic and the R-factor
monitor with projectorInterpolation = GENFIRE_INTERP_CUBIC."""
import numpy as np
import pytest

from genfire_amd import _lib

from test_gpu_iterate import _setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,angles", [
    ((32, 32, 32), (0.0, 0.0, 0.0)),
    ((32, 32, 32), (0.0, 37.0, 0.0)),       # a tilt about y: the cubic form, not the four-ray one
    ((33, 30, 31), (12.0, -47.0, 33.0)),
    ((64, 64, 64), (3.0, 65.0, -4.0)),
    ((33, 30, 31), (60.0, 60.0, 60.0)),     # 8 x 8 pixel patches
    ((31, 32, 33), (0.0, 30.0, 90.0)),      # lanes along i
    ((4, 5, 6), (20.0, 10.0, -30.0)),       # the smallest extents interp3 'cubic' takes
])
def test_project_cubic_matches_oracle(ctx, O, shape, angles):
    print(f"[cubic] {shape} {angles}", flush=True)
    vol = np.random.default_rng(1).uniform(0, 1, shape).astype(np.float32)
    got = ctx.project(vol, *angles, interp="cubic")
    ref = O.project(vol, *angles, interp="cubic")
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 2e-5, err


def test_project_cubic_of_a_quadratic(ctx, O):
    # interp3 'cubic' reproduces quadratics, boundary nodes included: the GPU
    # projection of an integer quadratic is the FP64 projection to FP32 rounding
    shape = (24, 20, 22)
    x, y, z = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    vol = (3 - 2 * x + y + 4 * z + x * x - y * y + 2 * z * z + x * y - 2 * y * z + x * z) / 100.0
    for angles in [(5.0, 33.0, -2.0), (0.5, -70.0, 4.0)]:
        print(f"[cubic] quadratic {angles}", flush=True)
        got = ctx.project(vol.astype(np.float32), *angles, interp="cubic")
        ref = O.project(vol.astype(np.float32), *angles, interp="cubic")
        assert np.abs(got - ref).max() <= 2e-5 * np.abs(ref).max()


@pytest.mark.parametrize("crop", [(30, 27), (40, 36), (7, 5)])
def test_rfactor_cubic_matches_oracle(ctx, O, crop):
    print(f"[cubic] rfactor {crop}", flush=True)
    rng = np.random.default_rng(5)
    vol = rng.uniform(0, 1, (40, 36, 44)).astype(np.float32)
    ang = np.stack([rng.uniform(-5, 5, 7), np.linspace(-63, 63, 7), rng.uniform(-5, 5, 7)], 1)
    m1, m2 = crop
    s1, s2 = 20 - m1 // 2, 18 - m2 // 2
    projs = np.stack([O.project(vol, *a, interp="cubic") for a in ang], 2)[s1:s1 + m1, s2:s2 + m2]
    projs = (projs * rng.uniform(0.9, 1.1, projs.shape)).astype(np.float32)
    R, Rarr, sims = ctx.rfactor(vol, projs, ang, interp="cubic")
    Rr, Rarr_r, sims_r = O.rfactor_posthoc(vol, projs, ang, interp="cubic")
    assert abs(R - Rr) < 1e-5
    np.testing.assert_allclose(Rarr, Rarr_r, atol=1e-5, rtol=0)
    assert np.abs(sims - sims_r).max() <= 2e-5 * np.abs(sims_r).max()


def test_cubic_rejects_small_extents(ctx):
    vol = np.ones((3, 8, 8), np.float32)
    with pytest.raises(_lib.GenfireError):
        ctx.project(vol, 0.0, 10.0, 0.0, interp="cubic")


@pytest.mark.parametrize("alg", ["genfire", "dr"])
def test_monitor_cubic_matches_oracle(ctx, O, alg):
    print(f"[cubic] monitor {alg}", flush=True)
    vol, projs, ang, mk, sup = _setup(O, dim=32, ratio=2.0, P=13, seed=9, euler=True)
    iters = 20
    d1, d2, P = projs.shape
    p = _lib.default_params()
    p.griddingMethod = 2
    p.numIterations = iters
    p.oversamplingRatio = 2.0
    p.interpolationCutoffDistance = 0.3
    p.constraintEnforcementMode = 1
    p.algorithm = _lib.ALG_DR if alg == "dr" else _lib.ALG_GENFIRE
    p.monitor_R = 1
    p.monitorR_loopLength = 5
    p.projectorInterpolation = _lib.INTERP_CUBIC
    job = _lib.Job(ctx, p, d1, d2, P)
    job.set_inputs(projs, ang, sup)
    job.set_measuredK(mk)
    job.reconstruct()
    ref = O.reconstruct(mk, mk != 0, sup, 32, 32, iters, mode=1, algorithm=alg, precision="fp64",
                        monitor_R=True, loop_length=5, projections=projs, angles=ang, projector_interp="cubic")
    rec = job.reconstruction()
    assert np.linalg.norm(rec - ref.reconstruction) / np.linalg.norm(ref.reconstruction) < 1e-4
    _, R1, R2 = job.errR()
    np.testing.assert_allclose(R1, ref.Rarr_record, atol=1e-4, rtol=0)
    np.testing.assert_allclose(R2, ref.Rarr2_record, atol=1e-4, rtol=0)
    # the linear monitor on the same run differs: the interpolant is really switched
    ref_lin = O.reconstruct(mk, mk != 0, sup, 32, 32, iters, mode=1, algorithm=alg, precision="fp64",
                            monitor_R=True, loop_length=5, projections=projs, angles=ang)
    assert np.abs(R1 - ref_lin.Rarr_record).max() > 1e-4


@pytest.mark.parametrize("vectors,angles", [
    (((0, 0, 1), (0, 1, 0), (1, 0, 0)), (10.0, -40.0, 25.0)),
    (((1, 1, 0), (0, 1, 1), (1, 0, 1)), (17.0, 33.0, -21.0)),
])
def test_project_general_cubic_matches_oracle(ctx, O, vectors, angles):
    """calculate3Dprojection_RealSpaceinterp_general.m:37,43 ('cubic') with the
    general rotation axes, and its post-hoc R-factor."""
    print(f"[cubic] general {vectors} {angles}", flush=True)
    rng = np.random.default_rng(6)
    vol = rng.uniform(0, 1, (40, 36, 38)).astype(np.float32)
    got = ctx.project_general(vol, *angles, *vectors, interp="cubic")
    ref = O.project_general(vol, *angles, *vectors, interp="cubic")
    assert np.abs(got - ref).max() / np.abs(ref).max() < 2e-5
    ang = np.array([angles, (angles[0], -angles[1], angles[2])])
    projs = np.stack([O.project_general(vol, *a, *vectors, interp="cubic") for a in ang], 2)
    projs = (projs * rng.uniform(0.9, 1.1, projs.shape)).astype(np.float32)
    R, Rarr, _ = ctx.rfactor_general(vol, projs, ang, *vectors, interp="cubic")
    for i, a in enumerate(ang):
        back = O.project_general(vol, *a, *vectors, interp="cubic")
        pj = projs[:, :, i].astype(np.float64)
        ref_r = np.sum(np.abs(np.abs(back) - np.abs(pj))) / np.sum(np.abs(pj))
        assert abs(Rarr[i] - ref_r) < 1e-5
