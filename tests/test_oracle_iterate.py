"""Pins the oracle's iteration (reconstruct.m) and projector without MATLAB:
MATLAB round/linspace/colon semantics, a fixed point, idempotence of P_R,
the last-iterate property of the schedule, DR fixed point and projector
closed forms."""
import numpy as np
import pytest


def test_schedule_matlab_round_half_away(O):
    # round(linspace(1,200,5)) = [1 51 101 150 200]; numpy's half-to-even gives 100
    np.testing.assert_array_equal(O.cutoff_schedule(200, 5), [1, 51, 101, 150, 200])
    np.testing.assert_array_equal(O.cutoff_schedule(100, 5), [1, 26, 51, 75, 100])
    assert O.cutoff_schedule(50, 27)[-1] == 50 and O.cutoff_schedule(50, 27)[0] == 1


def test_delay_indicators(O):
    d1 = O.delay_indicators(1)
    assert len(d1) == 27 and len(O.delay_indicators(2)) == 14
    np.testing.assert_allclose(d1[:12], np.arange(0.95, -0.16, -0.1), atol=1e-12)
    assert d1[0] == 0.95 and d1[11] == -0.15 and d1[15] == -0.15 and d1[-1] == 0.95
    np.testing.assert_array_equal(O.delay_indicators(3), [-999] * 5)
    with pytest.raises(ValueError):
        O.delay_indicators(4)


def test_colon_symmetric(O):
    v = O.matlab_colon(0.95, -0.1, -0.15)
    assert len(v) == 12 and v[0] == 0.95 and v[-1] == -0.15
    np.testing.assert_allclose(v, 0.95 - 0.1 * np.arange(12), atol=1e-15)


def test_kspace_indices_closed_form(O):
    Q = O.make_kspace_indices((8, 6, 5))
    assert Q[4, 3, 2] == 0.0                    # centre (n/2+1 even, (n+1)/2 odd, 1-based)
    assert Q[0, 3, 2] == 1.0 and Q[4, 0, 2] == 1.0 and Q[4, 3, 0] == 1.0


def test_paddzero_stripzero_roundtrip(O):
    rng = np.random.default_rng(0)
    for small, big in [((4, 5, 6), (9, 10, 11)), ((5, 5, 5), (8, 8, 8)), ((6, 6, 6), (7, 9, 8))]:
        v = rng.standard_normal(small).astype(np.float32)
        np.testing.assert_array_equal(O.stripzero(O.paddzero(v, big), small), v)


def test_iffshift_ind_matches_ifftshift(O):
    shape = (6, 5, 4)
    a = np.arange(np.prod(shape)).reshape(shape, order="F")
    inds = np.arange(a.size)
    sh = O.iffshift3_ind(shape, inds)
    b = np.fft.ifftshift(a)
    np.testing.assert_array_equal(b.ravel(order="F")[sh], a.ravel(order="F")[inds])


def _obj(n=16, d=8, seed=0):
    rng = np.random.default_rng(seed)
    obj = np.zeros((d, d, d), np.float32)
    obj[2:6, 2:6, 2:6] = rng.uniform(0.1, 1.0, (4, 4, 4))
    M = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(O_pad(obj, n))))
    return obj, M


def O_pad(obj, n):
    from oracle import oracle as O
    return O.paddzero(obj, (n, n, n))


@pytest.mark.parametrize("alg", ["genfire", "dr"])
def test_fixed_point(O, alg):
    """Every voxel measured and enforced with M = F(x*), x* positive in the
    support: the iteration returns x* and errK reaches ~0."""
    obj, M = _obj()
    r = O.reconstruct(M, np.ones(M.shape, bool), obj > 0, 8, 8, 30, mode=3, algorithm=alg, precision="fp64")
    np.testing.assert_allclose(r.reconstruction, obj, atol=1e-5)
    assert r.errK[-1] < 1e-6


def test_reconstruction_is_last_iterate(O):
    """The schedule always changes at it = N (reconstruct.m:103,123), so the
    returned object is P_R of the last iterate even when errK rose."""
    obj, M = _obj(seed=2)
    meas = np.random.default_rng(3).random(M.shape) < 0.5
    Mm = np.where(meas, M, 0)
    r10 = O.reconstruct(Mm, meas, obj > 0, 8, 8, 10, mode=3, precision="fp64")
    r9 = O.reconstruct(Mm, meas, obj > 0, 8, 8, 9, mode=3, precision="fp64")
    x9 = np.fft.fftshift(r9.last_iterate)
    sup = O.paddzero((obj > 0).astype(np.float32), M.shape)
    expect = O.stripzero(np.where(x9 < 0, 0, x9) * sup, (8, 8, 8))
    np.testing.assert_allclose(r10.reconstruction, expect, atol=1e-6)


def test_projector_axis_aligned_sums(O):
    """phi = theta = psi = 0 samples voxel centres exactly: the projection is the
    plain sum over z; theta = 90 sums along x (ZYZ about y)."""
    rng = np.random.default_rng(1)
    v = rng.uniform(0, 1, (6, 7, 8)).astype(np.float32)
    np.testing.assert_allclose(O.project(v, 0, 0, 0), v.astype(np.float64).sum(axis=2), rtol=1e-12)
    c = rng.uniform(0, 1, (7, 7, 7)).astype(np.float32)
    c64 = c.astype(np.float64)
    p90 = O.project(c, 0, 90, 0)
    # R*[X;Y;Z] at theta = 90 is (Z, Y, -X): pixel i sums the line z = -X_i over x
    ref = np.array([[c64[:, j, (7 // 2) - (i - 7 // 2)].sum() for j in range(7)] for i in range(7)])
    np.testing.assert_allclose(p90, ref, rtol=1e-12)


def test_projector_gaussian_closed_form(O):
    """A 3D Gaussian projects to a 2D Gaussian of amplitude A*sigma*sqrt(2 pi)."""
    n, s = 48, 3.0
    k = np.arange(n) - n // 2
    g = np.exp(-(k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2) / (2 * s * s))
    proj = O.project(g.astype(np.float32), 20.0, 35.0, -10.0)
    expect = s * np.sqrt(2 * np.pi) * np.exp(-(k[:, None] ** 2 + k[None, :] ** 2) / (2 * s * s))
    assert np.abs(proj - expect).max() < 0.02 * expect.max()


def test_rfactor_zero_for_self_consistent(O):
    rng = np.random.default_rng(5)
    v = rng.uniform(0, 1, (12, 12, 12)).astype(np.float32)
    ang = rng.uniform(-30, 30, (3, 3))
    projs = np.stack([O.project(v, *a) for a in ang], 2)
    R1, R2 = O.rfactors(v, projs, ang)
    assert np.all(R1 < 1e-12) and np.all(R2 < 1e-12)


def test_rfree_selection(O):
    """reconstruct.m:46-69 restated: per shell [edge - 0.05, edge'), floor(count
    * percent) withheld (at least one), all among the measured voxels and inside
    the shell, the same set for the same seed and a different one for another."""
    rng = np.random.default_rng(4)
    n = 16
    M = (rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))).astype(np.complex64)
    M[rng.random(M.shape) < 0.4] = 0
    Q = O.make_kspace_indices(M.shape).ravel(order="F")
    nz = np.flatnonzero(M.ravel(order="F") != 0)
    edges = O.matlab_linspace(0.0, 1.0, 11)
    a = O.rfree_select(M, 10, 0.05, 7)
    b = O.rfree_select(M, 10, 0.05, 7)
    c = O.rfree_select(M, 10, 0.05, 8)
    assert len(a) == 10
    differ = False
    for s, ((sel, vals), (sel_b, _), (sel_c, _)) in enumerate(zip(a, b, c)):
        cand = nz[(Q[nz] >= edges[s] - 0.05) & (Q[nz] < edges[s + 1])]
        assert sel.size == (max(1, int(np.floor(cand.size * 0.05))) if cand.size else 0)
        assert np.all(np.isin(sel, cand))
        np.testing.assert_array_equal(vals, M.ravel(order="F")[sel])
        np.testing.assert_array_equal(sel, sel_b)
        differ |= not np.array_equal(sel, sel_c)
    assert differ


@pytest.mark.parametrize("alg", ["genfire", "dr"])
def test_rfree_vanishes_at_fixed_point(O, alg):
    """With M = F(x*) for a positive supported x*, the withheld points are
    predicted by the iterate: Rfree of every non-empty shell falls to ~0, while
    it starts at 1 (k = 0 at the zero initial object).  An error in the index
    shift or in the value pairing would leave it large."""
    obj, M = _obj()
    r = O.reconstruct(M, np.ones(M.shape, bool), obj > 0, 8, 8, 60, mode=3, algorithm=alg, precision="fp64",
                      rfree=(5, 0.1, 3))
    filled = r.Rfree[:, 0] >= 0
    assert filled.any()
    np.testing.assert_allclose(r.Rfree[filled, 0], 1.0, atol=1e-6)
    assert np.all(r.Rfree[filled, -1] < 1e-3), r.Rfree[:, -1]


def test_general_projector_reduces_to_zyz(O):
    """calculate3Dprojection_RealSpaceinterp.m:3-12 states that its ZYZ matrix is
    Q([0 0 1],phi) Q([0 1 0],theta) Q([0 0 1],psi); the general projector with
    those axes gives the ZYZ projection."""
    rng = np.random.default_rng(8)
    vol = rng.uniform(0, 1, (12, 10, 11)).astype(np.float32)
    for ang in [(0.0, 0.0, 0.0), (20.0, -35.0, 50.0), (-7.0, 64.0, 3.0)]:
        Rq = O.quat_rot((0, 0, 1), ang[0]) @ O.quat_rot((0, 1, 0), ang[1]) @ O.quat_rot((0, 0, 1), ang[2])
        np.testing.assert_allclose(Rq, O.zyz(*ang), atol=1e-14)
        np.testing.assert_allclose(O.project_general(vol, *ang, (0, 0, 1), (0, 1, 0), (0, 0, 1)),
                                   O.project(vol, *ang), rtol=0, atol=1e-9)


def test_general_projector_quarter_turn_about_x(O):
    """Q([1 0 0], 90): the sum along z of the rotated volume equals an axis sum
    of the unrotated one (closed form, integer grid points only)."""
    rng = np.random.default_rng(9)
    n = 9
    vol = np.zeros((n, n, n), np.float32)   # zero faces: samples on the box faces may round outside
    vol[1:-1, 1:-1, 1:-1] = rng.uniform(0, 1, (n - 2, n - 2, n - 2))
    R = O.quat_rot((1, 0, 0), 90.0)
    assert np.allclose(np.abs(R), [[1, 0, 0], [0, 0, 1], [0, 1, 0]], atol=1e-15)
    got = O.project_general(vol, 90.0, 0.0, 0.0, (1, 0, 0), (0, 1, 0), (0, 0, 1))
    # sample (X, Y, Z) reads vol at R (X, Y, Z): y' = +-Z, z' = +-Y, so the sum over Z runs over volume y
    s = np.sign(R[1, 2])
    expect = vol.sum(axis=1)            # (x, z')
    expect = expect if np.sign(R[2, 1]) > 0 else expect[:, ::-1]
    np.testing.assert_allclose(got, expect, rtol=0, atol=1e-5)
    assert s != 0


# Nonzero closed forms of the errK and R-factor normalisations
# (reconstruct.m:135,176-177): a wrong denominator, a missing absolute value or
# a missing square root changes each of these values.

@pytest.mark.parametrize("alg", ["genfire", "dr"])
def test_errk_is_one_at_zero_iterate(O, alg):
    """x = 0 at the first iteration: k = fftn(0) = 0, so errK(1) = sum|M| / sum|M| = 1
    over any measured set."""
    obj, M = _obj(seed=4)
    meas = np.random.default_rng(6).random(M.shape) < 0.3
    r = O.reconstruct(np.where(meas, M, 0), meas, obj > 0, 8, 8, 1, mode=3, algorithm=alg, precision="fp64")
    assert abs(r.errK[0] - 1.0) < 1e-6


@pytest.mark.parametrize("c", [0.25, 0.6, 1.5, 3.0])
def test_errk_scaled_object(O, c):
    """initialObject = c x* (x* >= 0 in the support), M = fftn(x*) on a measured
    subset: errK(1) = sum||c k| - |k|| / sum|k| = |1 - c| (a denominator of
    sum|c k|, or squared magnitudes, gives a different value)."""
    obj, M = _obj(seed=5)
    meas = np.random.default_rng(7).random(M.shape) < 0.4
    r = O.reconstruct(np.where(meas, M, 0), meas, obj > 0, 8, 8, 1, mode=3, precision="fp64",
                      initial_object=(c * obj).astype(np.float32))
    assert abs(r.errK[0] - abs(1.0 - c)) < 1e-6, (r.errK[0], c)


def test_errk_uses_magnitudes_only(O):
    """errK compares magnitudes (abs(abs(k) - abs(M))): measured values with their
    phases scrambled leave errK(1) at 0 for the object they came from, and a
    measured set whose magnitudes are doubled gives exactly 1/2."""
    obj, M = _obj(seed=6)
    meas = np.abs(M) > 1e-9
    ph = np.exp(1j * np.random.default_rng(8).uniform(0, 2 * np.pi, M.shape))
    r = O.reconstruct(np.where(meas, M * ph, 0), meas, obj > 0, 8, 8, 1, mode=3, precision="fp64",
                      initial_object=obj)
    assert r.errK[0] < 1e-6
    r2 = O.reconstruct(np.where(meas, 2 * M, 0), meas, obj > 0, 8, 8, 1, mode=3, precision="fp64",
                       initial_object=obj)
    assert abs(r2.errK[0] - 0.5) < 1e-6


def test_rfactors_zero_model(O):
    """A zero model projects to zero: R1 = sum|pj| / sum|pj| = 1 and
    R2 = ||pj|| / ||pj|| = 1 for every projection."""
    rng = np.random.default_rng(10)
    v = rng.uniform(0, 1, (12, 12, 12)).astype(np.float32)
    ang = rng.uniform(-30, 30, (4, 3))
    projs = np.stack([O.project(v, *a) for a in ang], 2)
    R1, R2 = O.rfactors(np.zeros_like(v), projs, ang)
    np.testing.assert_allclose(R1, 1.0, atol=1e-12)
    np.testing.assert_allclose(R2, 1.0, atol=1e-12)
    Rm, Rarr, _ = O.rfactor_posthoc(np.zeros_like(v), projs, ang)
    np.testing.assert_allclose(Rarr, 1.0, atol=1e-12)


@pytest.mark.parametrize("c", [0.5, 0.9, 1.25, 2.0])
def test_rfactors_scaled_model(O, c):
    """The projector is linear: a model c x* on data projected from x* gives
    back = c pj, so R1 = R2 = |1 - c| (an R2 without the square root would give
    (1 - c)^2, a denominator over back would give |1 - c| / c).  With pj >= 0,
    Tian's post-hoc ||back| - |pj|| form gives the same |1 - c|."""
    rng = np.random.default_rng(11)
    v = rng.uniform(0, 1, (12, 12, 12)).astype(np.float32)
    ang = rng.uniform(-30, 30, (4, 3))
    projs = np.stack([O.project(v, *a) for a in ang], 2)
    R1, R2 = O.rfactors((c * v.astype(np.float64)).astype(np.float32), projs, ang)
    np.testing.assert_allclose(R1, abs(1 - c), atol=1e-6)
    np.testing.assert_allclose(R2, abs(1 - c), atol=1e-6)
    Rm, Rarr, _ = O.rfactor_posthoc((c * v.astype(np.float64)).astype(np.float32), projs, ang)
    np.testing.assert_allclose(Rarr, abs(1 - c), atol=1e-6)
    assert abs(Rm - abs(1 - c)) < 1e-6


def test_rfactor_monitor_records_scaled_iterate(O):
    """The in-loop monitor (reconstruct.m:168-185) on a fixed point whose
    measured values are c F(x*): after the first iteration the iterate is c x*
    (every voxel enforced), so Rarr = Rarr2 = |1 - c| per projection and errR is
    their mean."""
    c = 0.7
    obj, M = _obj(seed=9)
    ang = np.random.default_rng(12).uniform(-30, 30, (3, 3))
    projs = np.stack([O.project(obj, *a) for a in ang], 2)
    r = O.reconstruct(c * M, np.ones(M.shape, bool), obj > 0, 8, 8, 4, mode=3, precision="fp64",
                      monitor_R=True, loop_length=2, projections=projs, angles=ang)
    np.testing.assert_allclose(r.Rarr_record, abs(1 - c), atol=1e-5)
    np.testing.assert_allclose(r.Rarr2_record, abs(1 - c), atol=1e-5)
    np.testing.assert_allclose(r.errR, abs(1 - c), atol=1e-5)


@pytest.mark.parametrize("alg,mode", [("genfire", 3), ("dr", 1)])
def test_snapshots_follow_best_since_stage(O, alg, mode):
    """save_temp (reconstruct.m:149-159): the snapshot of iteration t is the
    iterate whose errK is the first minimum since the last enforcement stage
    (bestErr reset at :123, replaced only on a strictly lower errK); the last
    snapshot is the returned reconstruction."""
    obj, M = _obj(seed=7)
    meas = np.random.default_rng(9).random(M.shape) < 0.45
    N = 17
    r = O.reconstruct(np.where(meas, M, 0), meas, obj > 0, 8, 8, N, mode=mode, algorithm=alg, precision="fp64",
                      save_every=1)
    snaps = dict(r.snapshots)
    assert sorted(snaps) == list(range(1, N + 1))
    stages = set(O.cutoff_schedule(N, len(O.delay_indicators(mode))).tolist())
    start = 1
    for t in range(1, N + 1):
        if t in stages:
            start = t
        seg = r.errK[start - 1:t]
        b = start + int(np.argmin(seg))          # first occurrence of the minimum
        np.testing.assert_array_equal(snaps[t], snaps[b])
    np.testing.assert_array_equal(snaps[N], r.reconstruction)
    r3 = O.reconstruct(np.where(meas, M, 0), meas, obj > 0, 8, 8, N, mode=mode, algorithm=alg, precision="fp64",
                       save_every=3)
    assert [t for t, _ in r3.snapshots] == [3, 6, 9, 12, 15]
    for t, v in r3.snapshots:
        np.testing.assert_array_equal(v, snaps[t])
