"""GPU projector and R-factor (src/calculate3Dprojection_RealSpaceinterp.m,
src/Tian_calc_Rfactor_realspace.m) against the oracle, through the C ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,angles", [
    ((32, 32, 32), (0.0, 0.0, 0.0)),
    ((32, 32, 32), (0.0, 90.0, 0.0)),
    ((33, 30, 31), (12.0, -47.0, 33.0)),
    ((64, 64, 64), (3.0, 65.0, -4.0)),
    ((33, 30, 31), (60.0, 60.0, 60.0)),    # no pairing within 0.9: 8 x 8 pixel patches
    ((32, 36, 30), (70.0, 60.0, 50.0)),    # patches, x-fastest copies
    ((31, 32, 33), (0.0, 30.0, 90.0)),     # |R10| = 1: lanes along i, y-fastest copies
    ((30, 34, 32), (90.0, -30.0, 0.0)),    # |R01| = 1: lanes along j, x-fastest copies
    ((30, 34, 32), (0.0, 0.0, 0.0)),       # identity: y exact, four-ray form off (d2 % 4 != 0)
])
def test_project_matches_oracle(ctx, O, shape, angles):
    print(f"[proj] {shape} {angles}", flush=True)
    rng = np.random.default_rng(1)
    vol = rng.uniform(0, 1, shape).astype(np.float32)
    got = ctx.project(vol, *angles)
    ref = O.project(vol, *angles)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 2e-5, err


@pytest.mark.parametrize("shape,theta", [
    ((33, 32, 31), 37.0),      # d2 % 4 == 0: tilts about y take the four-ray form
    ((31, 64, 29), -70.0),
    ((32, 30, 32), 37.0),      # d2 % 4 != 0: the pair form
    ((24, 16, 20), 90.0),
])
def test_project_tilt_about_y_matches_oracle(ctx, O, shape, theta):
    print(f"[proj] tilt about y {shape} {theta}", flush=True)
    vol = np.random.default_rng(2).uniform(0, 1, shape).astype(np.float32)
    got = ctx.project(vol, 0.0, theta, 0.0)
    ref = O.project(vol, 0.0, theta, 0.0)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 2e-5, err
    # the last y row (cell d2-2 with weight 1) and the first
    assert np.abs(got[:, -1] - ref[:, -1]).max() <= 2e-5 * np.abs(ref).max()
    assert np.abs(got[:, 0] - ref[:, 0]).max() <= 2e-5 * np.abs(ref).max()


@pytest.mark.parametrize("crop", [(30, 27), (40, 36), (7, 5)])
def test_rfactor_tilt_series_about_y_matches_oracle(ctx, O, crop):
    # windows with an odd offset along y (s2 = 18 - 27 // 2 = 5), the full
    # volume, and a small one: the four-ray groups straddle the window's edges
    print(f"[proj] rfactor tilt series {crop}", flush=True)
    rng = np.random.default_rng(5)
    vol = rng.uniform(0, 1, (40, 36, 44)).astype(np.float32)
    ang = np.stack([np.zeros(7), np.linspace(-63, 63, 7), np.zeros(7)], 1)
    m1, m2 = crop
    s1, s2 = 20 - m1 // 2, 18 - m2 // 2
    projs = np.stack([O.project(vol, *a) for a in ang], 2)[s1:s1 + m1, s2:s2 + m2]
    projs = (projs * rng.uniform(0.9, 1.1, projs.shape)).astype(np.float32)
    R, Rarr, sims = ctx.rfactor(vol, projs, ang)
    Rr, Rarr_r, sims_r = O.rfactor_posthoc(vol, projs, ang)
    assert abs(R - Rr) < 1e-5
    np.testing.assert_allclose(Rarr, Rarr_r, atol=1e-5)
    np.testing.assert_allclose(sims, sims_r, rtol=0, atol=2e-5 * np.abs(sims_r).max())


def test_rfactor_matches_oracle(ctx, O):
    print("[proj] rfactor", flush=True)
    rng = np.random.default_rng(4)
    vol = rng.uniform(0, 1, (40, 40, 40)).astype(np.float32)
    ang = rng.uniform([-10, -70, -10], [10, 70, 10], (6, 3))
    projs = np.stack([O.project(vol, *a) for a in ang], 2)
    projs = (projs[4:36, 4:36] * rng.uniform(0.9, 1.1, (32, 32, 6))).astype(np.float32)
    R, Rarr, sims = ctx.rfactor(vol, projs, ang)
    Rr, Rarr_r, sims_r = O.rfactor_posthoc(vol, projs, ang)
    assert abs(R - Rr) < 1e-5
    np.testing.assert_allclose(Rarr, Rarr_r, atol=1e-5)
    np.testing.assert_allclose(sims, sims_r, rtol=0, atol=2e-5 * np.abs(sims_r).max())


@pytest.mark.parametrize("single_axis", [False, True])
@pytest.mark.parametrize("P", [17, 53])
def test_rfactor_many_projections_matches_oracle(ctx, O, P, single_axis):
    # more than 8 projections: each XCD takes a block of ceil(P/8) and
    # interleaves groups of neighbours (17: groups of 3, the last blocks short
    # or empty; 53: groups of 6, every block's second group partly filled);
    # a single-axis series takes the four-ray form
    print(f"[proj] rfactor {P} projections, single axis {single_axis}", flush=True)
    rng = np.random.default_rng(7)
    vol = rng.uniform(0, 1, (24, 24, 24)).astype(np.float32)
    pp = np.zeros(P) if single_axis else rng.uniform(-5, 5, P)
    ps = np.zeros(P) if single_axis else rng.uniform(-5, 5, P)
    ang = np.stack([pp, np.linspace(-70, 70, P), ps], 1)
    projs = np.stack([O.project(vol, *a) for a in ang], 2)[2:22, 3:21]
    projs = (projs * rng.uniform(0.9, 1.1, projs.shape)).astype(np.float32)
    R, Rarr, sims = ctx.rfactor(vol, projs, ang)
    Rr, Rarr_r, sims_r = O.rfactor_posthoc(vol, projs, ang)
    assert abs(R - Rr) < 1e-5
    np.testing.assert_allclose(Rarr, Rarr_r, atol=1e-5)
    np.testing.assert_allclose(sims, sims_r, rtol=0, atol=2e-5 * np.abs(sims_r).max())


@pytest.mark.parametrize("vectors,angles", [
    (((0, 0, 1), (0, 1, 0), (1, 0, 0)), (10.0, -40.0, 25.0)),
    (((0, 0, 1), (0, 1, 0), (0, 0, 1)), (3.0, 65.0, -4.0)),
    (((1, 1, 0), (0, 1, 1), (1, 0, 1)), (17.0, 33.0, -21.0)),
])
def test_project_general_matches_oracle(ctx, O, vectors, angles):
    """src/calculate3Dprojection_RealSpaceinterp_general.m (rotation axes
    vector1..3, MatrixQuaternionRot) against the oracle's FP64 trilinear form."""
    print(f"[proj] general {vectors} {angles}", flush=True)
    rng = np.random.default_rng(6)
    vol = rng.uniform(0, 1, (40, 36, 38)).astype(np.float32)
    got = ctx.project_general(vol, *angles, *vectors)
    ref = O.project_general(vol, *angles, *vectors)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 2e-5, err
    # the post-hoc R-factor with the same projector
    ang = np.array([angles, (angles[0], -angles[1], angles[2])])
    projs = np.stack([O.project_general(vol, *a, *vectors) for a in ang], 2)
    projs = (projs * rng.uniform(0.9, 1.1, projs.shape)).astype(np.float32)
    R, Rarr, _ = ctx.rfactor_general(vol, projs, ang, *vectors)
    for i, a in enumerate(ang):
        back = O.project_general(vol, *a, *vectors)
        pj = projs[:, :, i].astype(np.float64)
        ref_r = np.sum(np.abs(np.abs(back) - np.abs(pj))) / np.sum(np.abs(pj))
        assert abs(Rarr[i] - ref_r) < 1e-5


def test_rfactor_rejects_empty_projection_set(ctx):
    # Tian_calc_Rfactor_realspace over no projections has no R-factor: an
    # argument error, before any launch
    from genfire_amd._lib import GenfireError
    vol = np.zeros((8, 8, 8), np.float32)
    with pytest.raises(GenfireError):
        ctx.rfactor(vol, np.zeros((8, 8, 0), np.float32), np.zeros((0, 3)))
    # the context still serves a valid call afterwards
    R, Rarr, _ = ctx.rfactor(vol + 1.0, np.ones((8, 8, 1), np.float32), np.zeros((1, 3)))
    assert Rarr.shape == (1,) and np.isfinite(R)
