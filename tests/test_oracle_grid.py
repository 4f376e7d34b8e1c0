"""Pins the CPU oracle's gridding (fillInFourierGrid_DFT.m, CentroSymmetricity.m)
against closed forms, invariants, independent libraries and brute force.

None of these re-types the oracle's formula: each check would fail for a
plausible mistake (wrong rotation order or sign, transposed index, dropped
term, wrong in-plane test, FMA contraction, wrong mirror)."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation


# ---------------------------------------------------------------- rotations
@pytest.mark.parametrize("ang", [(0, 0, 0), (10, 20, 30), (-75, 33, 160), (5, -70, -5)])
def test_rotation_is_zyx_euler(O, ang):
    """R = Q(z,phi) Q(y,theta) Q(x,psi) (:110-114) is scipy's intrinsic 'ZYX'."""
    _, Rd = O.rotation(*ang)
    ref = Rotation.from_euler("ZYX", ang, degrees=True).as_matrix()
    np.testing.assert_allclose(Rd, ref, atol=1e-15)


def test_rotation_zyz_with_vector3_z(O):
    _, Rd = O.rotation(10, 20, 30, v3=(0, 0, 1))
    np.testing.assert_allclose(Rd, Rotation.from_euler("ZYZ", (10, 20, 30), degrees=True).as_matrix(), atol=1e-15)
    # the projector's hard-coded ZYZ matrix (calculate3Dprojection_RealSpaceinterp.m:13-15)
    np.testing.assert_allclose(O.zyz(10, 20, 30), Rd, atol=1e-15)


def test_rotation_single_axis_closed_form(O):
    for t in (-60.0, -3.0, 0.0, 45.0, 57.0):
        _, Rd = O.rotation(0, t, 0)
        c, s = np.cos(np.radians(t)), np.sin(np.radians(t))
        np.testing.assert_allclose(Rd, [[c, 0, s], [0, 1, 0], [-s, 0, c]], atol=1e-15)


def test_rotation_axis_scaling_invariant(O):
    """MatrixQuaternionRot normalises the axis (:356)."""
    a, _ = O.rotation(10, 20, 30, v1=(0, 0, 1))
    b, _ = O.rotation(10, 20, 30, v1=(0, 0, 7))
    np.testing.assert_array_equal(a, b)


# ---------------------------------------------------------------- index pass
def test_theta0_is_exactly_the_k3_plane(O):
    """theta = 0: D = |k3|, the kept set is the k3 = 0 plane of the half grid
    (65 x 129 = 8385 points at n = 128), every point an exact hit, u = k1, v = k2."""
    R, _ = O.rotation(0, 0, 0)
    ind, D, u, v, wrong = O.select(128, 128, True, R, 0.3)
    assert len(ind) == 65 * 129 and wrong == 0
    assert np.all(D == 0)
    i1, i2, i3 = np.unravel_index(ind, (65, 129, 129), order="F")
    assert np.all(i3 == 64)
    np.testing.assert_array_equal(u, (i1 - 64).astype(np.float32))
    np.testing.assert_array_equal(v, (i2 - 64).astype(np.float32))


def test_45deg_exact_plane_closed_form(O):
    """theta = 45: the FP32 normal has equal x and z components, so the plane
    k1 = -k3 is hit exactly (D == 0); the in-plane radius |k1|*sqrt(2) <= 64 keeps
    |k1| <= 45, i.e. 46 x 129 = 5934 exact points and nothing else below 1e-7."""
    R, _ = O.rotation(0, 45, 0)
    assert R[0, 2] == R[2, 2]
    ind, D, u, v, _ = O.select(128, 128, True, R, 0.3)
    exact = D < 1e-7
    assert exact.sum() == 46 * 129
    i1, i2, i3 = np.unravel_index(ind[exact], (65, 129, 129), order="F")
    np.testing.assert_array_equal((i1 - 64), -(i3 - 64))


def test_single_axis_tilt_contains_k2_axis(O):
    """Any tilt about y contains the k2 axis (k1 = k3 = 0) with D = 0."""
    for t in (-57.0, -21.0, 12.0, 60.0):
        R, _ = O.rotation(0, t, 0)
        ind, D, u, v, _ = O.select(64, 64, True, R, 0.3)
        i1, i2, i3 = np.unravel_index(ind, (33, 65, 65), order="F")
        axis = (i1 == 32) & (i3 == 32)
        assert axis.sum() == 65
        assert np.all(D[axis] == 0)


def test_distance_centrosymmetric_bitwise(O):
    """D(-k) == D(k) bit for bit (odd-symmetric FP32 ops), on the full grid
    (DFT_CentroSymmetricity = 0) so both k and -k are enumerated."""
    R, _ = O.rotation(17, -38, 9)
    ind, D, u, v, _ = O.select(63, 63, False, R, 0.7)
    i = np.stack(np.unravel_index(ind, (63, 63, 63), order="F"), 1) - 31
    key = {tuple(k): (d, uu, vv) for k, d, uu, vv in zip(i, D, u, v)}
    for k, (d, uu, vv) in list(key.items())[:4000]:
        d2, u2, v2 = key[tuple(-np.array(k))]
        assert d2 == d and u2 == -uu and v2 == -vv


def test_brute_force_float32_python(O):
    """Tiny grid, point by point with numpy float32 scalars: the same selection and
    D bits, and (u, v) within a few ulp of LAPACK's sgesv (numpy.linalg.solve)."""
    R, _ = O.rotation(23, -51, 14, v3=(0, 0, 1))
    n1 = n2 = 12
    cutoff = 0.7
    ind, D, u, v, _ = O.select(n1, n2, True, R, cutoff)
    f = np.float32
    nv = R[:, 2].astype(np.float32)
    nrm = f(np.sqrt(np.sum(nv.astype(np.float64) ** 2)))
    nh = nv / nrm
    nn = f(f(nv[0] * nv[0]) + f(nv[1] * nv[1])) + f(nv[2] * nv[2])
    got = []
    for k3 in range(-6, 7):
        for k2 in range(-6, 7):
            for k1 in range(-6, 1):
                K = np.array([k1, k2, k3], np.float32)
                t = f(f(f(K[0] * nh[0]) + f(K[1] * nh[1])) + f(K[2] * nh[2]))
                d = abs(t)
                if not float(d) < cutoff:
                    continue
                s = f(f(f(K[0] * nv[0]) + f(K[1] * nv[1])) + f(K[2] * nv[2]))
                vv = f(f(0) - s) / nn
                cp = np.array([f(K[i] + f(vv * nv[i])) for i in range(3)], np.float32)
                x = np.linalg.solve(R.astype(np.float32), cp)
                if abs(x[0]) <= n1 / 2 and abs(x[1]) <= n2 / 2:
                    lin = (k1 + 6) + 7 * ((k2 + 6) + 13 * (k3 + 6))
                    got.append((lin, d, x[0], x[1]))
    got.sort()
    assert [g[0] for g in got] == list(ind)
    np.testing.assert_array_equal(np.array([g[1] for g in got], np.float32), D)
    # LAPACK's sgesv (OpenBLAS here, MKL in MATLAB) orders its FP32 operations
    # differently: agreement to a few ulp of the |u| <= n/2 decision scale.
    tol = 4 * np.spacing(np.float32(n1 / 2))
    assert np.all(np.abs(np.array([g[2] for g in got]) - u) <= tol)
    assert np.all(np.abs(np.array([g[3] for g in got]) - v) <= tol)


def test_lu_solve_matches_exact_inverse(O):
    """R \\ b: the restated sgesv agrees with R^T b (R orthogonal) to a few ulp."""
    rng = np.random.default_rng(0)
    for _ in range(50):
        R, Rd = O.rotation(*rng.uniform(-90, 90, 3))
        b = rng.uniform(-100, 100, 3).astype(np.float32)
        x = O.lu_solve(R, b)
        exact = np.linalg.solve(R.astype(np.float64), b.astype(np.float64))
        assert np.all(np.abs(x - exact) <= 4 * np.spacing(np.float32(128)))


def test_kept_count_scales_with_slab(O):
    """The kept set of a tilted plane is the slab |D| < c intersected with the
    in-plane window: about 2c points per unit plane area."""
    R, _ = O.rotation(0, 30, 0)
    for c in (0.3, 0.7):
        ind, D, u, v, _ = O.select(128, 128, True, R, c)
        area_half = 129 * 129 / 2.0     # half of the |u|,|v| <= 64 window
        assert abs(len(ind) / (2 * c * area_half) - 1) < 0.1


def test_tie_census_closed_form(O):
    """theta = 0 with cutoff 1: D = |k3| exactly, so the planes k3 = +-1 sit on the
    cutoff (2 x 65 x 129 points); the kept set is the k3 = 0 plane (65 x 129) with
    u = k1, v = k2, whose edge |u| = 64 or |v| = 64 holds 129 + 2 x 65 - 2 points."""
    R, _ = O.rotation(0, 0, 0)
    nd, nuv, nkept, nexact = O.tie_census(128, 128, True, R, 1.0, ulps=1)
    assert (nd, nuv, nkept) == (2 * 65 * 129, 129 + 2 * 65 - 2, 65 * 129)
    assert nexact == nuv   # u and v are the integers k1, k2 there
    # the kept count is the selection's
    R, _ = O.rotation(3, -33, 2, v3=(0, 0, 1))
    nd, nuv, nkept, nexact = O.tie_census(64, 64, True, R, 0.3, ulps=2)
    assert nkept == len(O.select(64, 64, True, R, 0.3)[0])
    assert 0 <= nd and 0 <= nexact <= nuv <= nkept


# ---------------------------------------------------------------- DFT values
def test_dft_integer_frequencies_equal_fft(O):
    """At integer in-plane frequencies the DFT equals the FFT of the projection
    zero-padded to n1 x n2, with the centred pixel grid (k1o) as a phase."""
    rng = np.random.default_rng(1)
    d1, d2, n1, n2 = 8, 6, 16, 16     # powers of two: m/n1 and l/n2 are exact in FP32
    P = rng.standard_normal((d1, d2)).astype(np.float32)
    m = rng.integers(-8, 8, 20)
    l = rng.integers(-8, 8, 20)
    un = (m / n1).astype(np.float32)
    vn = (l / n2).astype(np.float32)
    F = O.dft64(P, un, vn)
    pad = np.zeros((n1, n2))
    # pixel a sits at x = a - floor(d1/2); place it at index x mod n1
    for a in range(d1):
        for b in range(d2):
            pad[(a - d1 // 2) % n1, (b - d2 // 2) % n2] = P[a, b]
    FF = np.fft.fft2(pad)
    ref = FF[m % n1, l % n2]
    np.testing.assert_allclose(F, ref, rtol=0, atol=1e-12 * np.abs(FF).max())


def test_dft_separable_equals_brute(O):
    rng = np.random.default_rng(2)
    P = rng.standard_normal((7, 9)).astype(np.float32)
    un = rng.uniform(-0.5, 0.5, 15).astype(np.float32)
    vn = rng.uniform(-0.5, 0.5, 15).astype(np.float32)
    np.testing.assert_allclose(O.dft64(P, un, vn), O.dft64(P, un, vn, brute=True), rtol=1e-13, atol=1e-12)


def test_dft_hermitian(O):
    rng = np.random.default_rng(3)
    P = rng.standard_normal((10, 10)).astype(np.float32)
    un = rng.uniform(-0.5, 0.5, 25).astype(np.float32)
    vn = rng.uniform(-0.5, 0.5, 25).astype(np.float32)
    np.testing.assert_allclose(O.dft64(P, -un, -vn), np.conj(O.dft64(P, un, vn)), atol=1e-12)


# ---------------------------------------------------------------- averaging
def test_weighted_average_closed_form(O):
    """:267-291: inverse-distance average, exact hits override, last exact wins."""
    ind = np.array([0, 0, 1, 2, 2, 2, 3], np.int64)
    D = np.array([0.25, 0.5, 0.1, 0.0, 0.2, 0.0, 0.0], np.float32)
    F = np.array([1 + 1j, 4, 7j, 10, 20, 30, 5j])
    FS = O.weighted_average(ind, D, F, 5)
    assert FS[0] == pytest.approx((4 * (1 + 1j) + 2 * 4) / 6)
    assert FS[1] == pytest.approx(7j)
    assert FS[2] == 30           # two exact hits: the later one wins
    assert FS[3] == 5j           # exact only
    assert FS[4] == 0


# ---------------------------------------------------------------- centrosymmetry
@pytest.mark.parametrize("n1,n2", [(8, 8), (7, 7), (8, 7), (7, 8), (10, 12), (9, 11)])
def test_centro_symmetricity_index_rule(O, n1, n2):
    """All four parity branches of CentroSymmetricity.m: rows k1 <= 0 copy FS(k),
    rows k1 > 0 are conj(FS(-k)), with the half grid's +n/2 extension used for
    the mirrored Nyquist entries."""
    (lo1, lo2, lo3), (l1, l2, l3) = O.half_ranges(n1, n2, True)
    rng = np.random.default_rng(n1 * 100 + n2)
    FS = rng.standard_normal(l1 * l2 * l3) + 1j * rng.standard_normal(l1 * l2 * l3)
    out = O.centro_symmetricity(FS, n1, n1 // 2, n2, dtype=np.complex128)
    H = FS.reshape(l1, l2, l3, order="F")
    for c1 in range(n1):
        for c2 in range(n2):
            for c3 in range(n1):
                k1, k2, k3 = c1 - n1 // 2, c2 - n2 // 2, c3 - n1 // 2
                if k1 <= 0:
                    exp = H[k1 - lo1, k2 - lo2, k3 - lo3]
                else:
                    exp = np.conj(H[-k1 - lo1, -k2 - lo2, -k3 - lo3])
                assert out[c1, c2, c3] == exp


def test_centro_grid_is_hermitian_off_self_paired_planes(O):
    """With real projections the gridded k-space is Hermitian:
    measuredK(k) = conj(measuredK(-k)) wherever both k and -k are in the grid
    and k1 != 0 (the k1 = 0 plane is filled independently from the half grid)."""
    rng = np.random.default_rng(4)
    projs = rng.uniform(0, 1, (16, 16, 5)).astype(np.float32)
    ang = rng.uniform([-10, -60, -10], [10, 60, 10], (5, 3))
    g = O.grid_dft(projs, ang, 32, 32, 0.5)
    mk = g.measuredK
    c = 16
    for k1 in range(1, 16):
        a = mk[c + k1, 1:, 1:]
        b = mk[c - k1, 1:, 1:][::-1, ::-1]
        np.testing.assert_array_equal(a, np.conj(b))


@pytest.mark.parametrize("d1,d2", [(64, 64), (65, 48), (128, 96)])
def test_cpu_nufft_matches_direct_dft(O, d1, d2):
    """The oracle's type-2 NUFFT (the CPU route of fillInFourierGrid_DFT.m:249-254,
    used by the CPU baseline) against the FP64 direct sum at random in-plane
    points, to nufft2d2's requested tolerance 1e-6; and at integer frequencies
    against the FFT of the projection (a textbook identity)."""
    rng = np.random.default_rng(d1 * 1000 + d2)
    P = rng.uniform(0.0, 1.0, (d1, d2)).astype(np.float32)
    un = rng.uniform(-0.5, 0.5, 2000).astype(np.float32)
    vn = rng.uniform(-0.5, 0.5, 2000).astype(np.float32)
    a = O.nufft2d2(P, un, vn)
    b = O.dft64(P, un, vn)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-6
    # u = p/d1, v = q/d2: F = sum P[a,b] exp(-2 pi i ((a - c1) p/d1 + (b - c2) q/d2)) = shifted fft2
    p = np.arange(-(d1 // 2), d1 - d1 // 2)
    q = np.arange(-(d2 // 2), d2 - d2 // 2)
    pp, qq = np.meshgrid(p, q, indexing="ij")
    ui = (pp / d1).astype(np.float32).ravel()
    vi = (qq / d2).astype(np.float32).ravel()
    exact = (np.fft.fft2(np.roll(P.astype(np.float64), (-(d1 // 2), -(d2 // 2)), axis=(0, 1)))
             [pp % d1, qq % d2]).ravel()
    # the FP32 coordinates p/d1 are not exactly the integer frequencies; compare with the direct sum there
    direct = O.dft64(P, ui, vi)
    fine = O.nufft2d2(P, ui, vi)
    assert np.linalg.norm(fine - direct) / np.linalg.norm(direct) < 1e-6
    if d1 & (d1 - 1) == 0 and d2 & (d2 - 1) == 0:   # p / d exact in FP32
        assert np.linalg.norm(direct - exact) / np.linalg.norm(exact) < 1e-12


@pytest.mark.parametrize("d1,d2", [(32, 32), (48, 40), (64, 64)])
def test_cpu_nufft_width16_is_fp64_grade(O, d1, d2):
    """With kernel width 16 (sigma 2) the oracle's NUFFT agrees with the FP64
    direct sum (fillInFourierGrid_DFT.m:236-248) to 1e-12: the FP64 reference
    that reruns the gridding at the 1024^3 grids, where the direct sum is out of
    reach on the CPU."""
    rng = np.random.default_rng(7 + d1)
    P = rng.uniform(0.0, 1.0, (d1, d2)).astype(np.float32)
    un = rng.uniform(-0.5, 0.5, 3000).astype(np.float32)
    vn = rng.uniform(-0.5, 0.5, 3000).astype(np.float32)
    b = O.dft64(P, un, vn)
    assert np.linalg.norm(O.nufft2d2(P, un, vn, w=16) - b) / np.linalg.norm(b) < 1e-12
    # a width-8 kernel is not: the check above has teeth
    assert np.linalg.norm(O.nufft2d2(P, un, vn, w=8) - b) / np.linalg.norm(b) > 1e-9
