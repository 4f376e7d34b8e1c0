"""The reference's data_turb3 tilt series (its only fixture, Genfire_test.m:3-4)
at the driver's settings: oversampling 4, cutoff 0.125, vector3 = [1 0 0]."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def turb3():
    return np.load(os.path.join(HERE, "golden", "turb3.npz"))


def test_fixture_invariants(turb3):
    pj, ang = turb3["projections"], turb3["angles"]
    assert pj.shape == (64, 64, 41) and ang.shape == (41, 3)
    np.testing.assert_array_equal(ang[:, 1], np.linspace(-70, 70, 41))
    assert np.all(ang[:, 0] == 0) and np.all(ang[:, 2] == 0)
    # constant projected mass, the only invariant the data carries
    s = turb3["proj_sums"]
    assert np.all(np.abs(s / s.mean() - 1) < 1e-3) and abs(s.mean() - 87396) < 50


def test_oracle_grid_turb3_closed_forms(O, turb3):
    """The theta = 0 projection (index 20) hits exactly the k3 = 0 plane of the
    256^3 grid; every tilt contains the k2 axis with D = 0."""
    pj, ang = turb3["projections"], turb3["angles"]
    assert ang[20, 1] == 0
    R, _ = O.rotation(*ang[20])
    ind, D, u, v, _ = O.select(256, 256, True, R, 0.125)
    assert len(ind) == 129 * 257 and np.all(D == 0)
    R, _ = O.rotation(*ang[3])
    ind, D, u, v, _ = O.select(256, 256, True, R, 0.125)
    i1, i2, i3 = np.unravel_index(ind, (129, 257, 257), order="F")
    axis = (i1 == 128) & (i3 == 128)
    assert axis.sum() == 257 and np.all(D[axis] == 0)


@pytest.mark.gpu
def test_gpu_grid_turb3_bitexact(ctx, O, turb3):
    from genfire_amd import _lib
    print("[turb3] gridding 64x64x41 at oversampling 4", flush=True)
    pj, ang = turb3["projections"], turb3["angles"]
    p = _lib.default_params()
    p.griddingMethod = 2   # DFT gridding
    p.oversamplingRatio = 4
    p.interpolationCutoffDistance = 0.125
    p.constraintEnforcementMode = 3
    p.numIterations = 5
    job = _lib.Job(ctx, p, 64, 64, 41)
    job.set_inputs(pj, ang)
    job.grid()
    ref = O.grid_dft(pj, ang, 256, 256, 0.125, values=False)
    pts = job.points()
    order = np.lexsort((pts["ind"], pts["proj"]))
    np.testing.assert_array_equal(pts["ind"][order], ref.master_ind)
    np.testing.assert_array_equal(pts["dist"][order].view(np.uint32), ref.master_dist.view(np.uint32))
    np.testing.assert_array_equal(job.hits(), ref.hits)
    # values on a sample of points of three projections
    for q in (0, 20, 40):
        g = np.nonzero(ref.master_proj == q)[0][::7]
        F64 = O.dft64(pj[:, :, q], ref.master_un[g], ref.master_vn[g])
        got = pts["val"][order][g]
        assert np.linalg.norm(got - F64) / np.linalg.norm(F64) < 1e-5
