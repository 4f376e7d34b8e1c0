"""Slab sweep of the trilinear projector (src/calculate3Dprojection_RealSpaceinterp.m,
reconstruct.m:168-185): models whose z-duplicated form passes 128 MiB are
projected one z slab per launch with each ray's running sum kept
between launches.  The sweep must give the one-pass projector's results bit for
bit; GENFIRE_PROJECTOR_SLABS sets the slab count (1: one pass) on small models."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# three-Euler, near-y (pair form), no pairing within 0.9 (8 x 8 patches),
# lanes along i, rays pointing down in z (rz < 0), a ray plane at 90 degrees
# (rz rounding to a tiny value) and an exact tilt about y
ANGLES = np.array([
    (3.0, 35.0, -4.0),
    (-2.0, -68.0, 5.0),
    (61.0, 58.0, 63.0),
    (0.0, 30.0, 90.0),
    (10.0, 120.0, -15.0),
    (-5.0, 170.0, 2.0),
    (4.0, 90.0, -3.0),
    (0.0, -90.0, 0.0),
    (0.0, 25.0, 0.0),
])


def _rfactor(ctx, vol, projs, ang, slabs, monkeypatch):
    monkeypatch.setenv("GENFIRE_PROJECTOR_SLABS", str(slabs))
    return ctx.rfactor(vol, projs, ang)


@pytest.mark.parametrize("slabs", [2, 7, 44])
def test_slab_sweep_matches_one_pass(ctx, O, monkeypatch, slabs):
    print(f"[slab] post-hoc R-factor, {slabs} slabs", flush=True)
    rng = np.random.default_rng(11)
    vol = rng.uniform(0, 1, (40, 36, 44)).astype(np.float32)
    projs = np.stack([O.project(vol, *a) for a in ANGLES], 2)[4:36, 2:34]
    projs = (projs * rng.uniform(0.9, 1.1, projs.shape)).astype(np.float32)
    R1, Rarr1, sims1 = _rfactor(ctx, vol, projs, ANGLES, 1, monkeypatch)
    Rn, Rarrn, simsn = _rfactor(ctx, vol, projs, ANGLES, slabs, monkeypatch)
    assert np.array_equal(sims1.view(np.uint32), simsn.view(np.uint32))
    assert np.array_equal(Rarr1, Rarrn) and R1 == Rn
    # and against the oracle (the angles keep every sample off the model's edge,
    # where FP32 and FP64 rounding can decide inside/outside differently)
    Rr, Rarr_r, sims_r = O.rfactor_posthoc(vol, projs, ANGLES)
    err = np.abs(simsn - sims_r).max(axis=(0, 1)) / np.abs(sims_r).max()
    assert err.max() < 2e-5, err
    np.testing.assert_allclose(Rarrn, Rarr_r, atol=1e-5)


def test_slab_sweep_in_the_monitor(ctx, O, monkeypatch):
    """A DR job's R-factor records with the monitor's projections in 5 slabs
    equal those of one pass."""
    print("[slab] DR job monitor, 5 slabs", flush=True)
    from test_gpu_iterate import _job, _setup
    vol, projs, ang, mk, sup = _setup(O, euler=True)
    out = []
    for slabs in (1, 5):
        monkeypatch.setenv("GENFIRE_PROJECTOR_SLABS", str(slabs))
        job = _job(ctx, mk, projs, ang, sup, 12, alg="dr", monitor=1, L=4)
        out.append((job.errR(), job.reconstruction()))
    (e1, a1, b1), rec1 = out[0]
    (e5, a5, b5), rec5 = out[1]
    assert np.array_equal(np.asarray(e1), np.asarray(e5))
    assert np.array_equal(np.asarray(a1), np.asarray(a5)) and np.array_equal(np.asarray(b1), np.asarray(b5))
    assert np.array_equal(rec1, rec5)


def test_slab_sweep_at_its_own_size_matches_oracle(ctx, O, monkeypatch):
    """A 336^3 model's z-duplicated form (305 MB) passes the 128 MiB threshold:
    the projector sweeps it in slabs without any override, bit for bit as one
    pass, and within the oracle's bar except at rays with a sample within an
    FP32 rounding of the model's edge (about 3e-5 of the 336 x 336 x 2 rays),
    where FP32 and FP64 decide inside/outside differently: each such sample
    moves its ray sum by at most one voxel's value."""
    print("[slab] 336^3 model, natural slab count", flush=True)
    assert 8 * 337 * 337 * 336 > 128 * 2 ** 20   # two floats per element, a zero row past x and y
    rng = np.random.default_rng(12)
    vol = rng.uniform(0, 1, (336, 336, 336)).astype(np.float32)
    ang = np.array([(3.0, 35.0, -4.0), (-5.0, 150.0, 2.0)])
    projs = np.stack([O.project(vol, *a) for a in ang], 2).astype(np.float32)
    projs = (projs * rng.uniform(0.9, 1.1, projs.shape)).astype(np.float32)
    monkeypatch.setenv("GENFIRE_PROJECTOR_SLABS", "1")
    R1, Rarr1, sims1 = ctx.rfactor(vol, projs, ang)
    monkeypatch.delenv("GENFIRE_PROJECTOR_SLABS")
    R, Rarr, sims = ctx.rfactor(vol, projs, ang)
    assert np.array_equal(sims1.view(np.uint32), sims.view(np.uint32))
    assert np.array_equal(Rarr1, Rarr) and R1 == R
    Rr, Rarr_r, sims_r = O.rfactor_posthoc(vol, projs, ang)
    d = np.abs(sims - sims_r)
    off = d > 2e-5 * np.abs(sims_r).max()
    print(f"[slab] {off.sum()} of {off.size} rays off the 2e-5 bar, largest {d.max():.3g}", flush=True)
    assert off.sum() <= 3e-4 * off.size
    assert d.max() <= vol.max()
    np.testing.assert_allclose(Rarr, Rarr_r, atol=1e-5)
