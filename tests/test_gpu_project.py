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
])
def test_project_matches_oracle(ctx, O, shape, angles):
    print(f"[proj] {shape} {angles}", flush=True)
    rng = np.random.default_rng(1)
    vol = rng.uniform(0, 1, shape).astype(np.float32)
    got = ctx.project(vol, *angles)
    ref = O.project(vol, *angles)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 2e-5, err


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
