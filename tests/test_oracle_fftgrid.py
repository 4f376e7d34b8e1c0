"""Pins of the FFT-gridding restatement (oracle.grid_fft, fillInFourierGrid.m:27-163;
oracle.hermitian_symmetrize, src/hermitianSymmetrize.m:22-65) by closed forms:
a mistake in the shift convention, the kx/ky orientation, the rotation's sign or
transposition, the mate rule or the zero-count rule fails one of them."""
import numpy as np
import pytest


def test_rotation_is_the_projector_matrix_transposed(O):
    # fillInFourierGrid.m:80-84 restated in Python equals the transpose of the
    # projector's ZYZ matrix (C restatement of calculate3Dprojection_RealSpaceinterp.m:13-15)
    rng = np.random.default_rng(1)
    angs = [(0, 0, 0), (90, -90, 180), (270, 45, -90)] + [tuple(a) for a in rng.uniform(-400, 400, (20, 3))]
    for a in angs:
        np.testing.assert_array_equal(O.fft_grid_rotation(*a), O.zyz(*a).T)


def test_my_fft_of_centre_delta_is_flat(O):
    for d in (8, 9):
        img = np.zeros((d, d))
        img[d // 2, d // 2] = 1.0      # centre pixel n/2+1 (even) or (n+1)/2 (odd), 1-based
        np.testing.assert_allclose(O.my_fft(img), np.ones((d, d)), atol=1e-15)


def test_hermitian_symmetrize_closed_forms(O):
    d = 6                               # even: index 0 of each edge has no mate
    c = d // 2
    a = np.zeros((d, d, d), np.complex128)
    a[c + 1, c - 2, c + 2] = 2 + 3j      # k = (1, -2, 2), mate (-1, 2, -2) empty
    out = O.hermitian_symmetrize(a)
    assert out[c + 1, c - 2, c + 2] == 2 + 3j and out[c - 1, c + 2, c - 2] == 2 - 3j
    assert np.count_nonzero(out) == 2
    b = a.copy()
    b[c - 1, c + 2, c - 2] = 4 - 1j      # both present: (a + conj(b)) / 2
    out = O.hermitian_symmetrize(b)
    assert out[c + 1, c - 2, c + 2] == ((2 + 3j) + (4 + 1j)) / 2
    assert out[c - 1, c + 2, c - 2] == ((4 - 1j) + (2 - 3j)) / 2
    e = np.zeros((d, d, d), np.complex128)
    e[0, c, c] = 5j                      # k1 = -d/2 meets the padding: unchanged
    np.testing.assert_array_equal(O.hermitian_symmetrize(e), e)
    # odd edge: the centre voxel pairs with itself -> its real part
    f = np.zeros((5, 5, 5), np.complex128)
    f[2, 2, 2] = 3 + 4j
    assert O.hermitian_symmetrize(f)[2, 2, 2] == 3


def _spectrum(O, proj, d):
    return O.my_fft(O.paddzero(proj.astype(np.float32), (d, d)))


@pytest.mark.parametrize("allow", [True, False])
def test_identity_orientation_puts_the_2d_spectrum_on_the_kz0_plane(O, allow):
    # R = I: every lattice point sits on its voxel (distance 0, weight 1/1e-5)
    # and no other voxel lies within the cutoff 0.7
    rng = np.random.default_rng(2)
    proj = rng.uniform(0, 1, (12, 10, 1))
    r = O.grid_fft(proj, np.zeros((1, 3)), 2.0, 0.7, allow)
    d = r.d
    assert d == 20
    F = _spectrum(O, proj[:, :, 0], d)
    exp = np.zeros((d, d, d), np.complex128)
    exp[:, :, d // 2] = F
    np.testing.assert_allclose(r.measuredK, exp.astype(np.complex64), rtol=0, atol=1e-5 * np.abs(F).max())
    hits = np.zeros((d, d, d), np.int32)
    hits[:, :, d // 2] = 1
    np.testing.assert_array_equal(r.hits, hits)


def test_quarter_turn_about_y_maps_kx_to_minus_kz(O):
    # theta = 90: rotkCoords = (0, ky, -kx) exactly (cosd(90) = 0), so
    # measuredK(0, ky, -kx) = F(kx, ky); the row kx = -d/2 falls outside the grid
    rng = np.random.default_rng(3)
    proj = rng.uniform(0, 1, (8, 8, 1))
    r = O.grid_fft(proj, np.array([[0.0, 90.0, 0.0]]), 2.0, 0.7, True)
    d, c = r.d, r.d // 2
    F = _spectrum(O, proj[:, :, 0], d)
    exp = np.zeros((d, d, d), np.complex128)
    for kx in range(-c, d - c):
        for ky in range(-c, d - c):
            if 0 <= c - kx < d:
                exp[c, c + ky, c - kx] = F[c + kx, c + ky]
    np.testing.assert_allclose(r.measuredK, exp.astype(np.complex64), rtol=0, atol=1e-5 * np.abs(F).max())
    assert r.hits.sum() == d * (d - 1)


def test_single_match_rule_counts_every_lattice_point_once(O):
    # allowMultipleGridMatches = 0 with a cutoff above sqrt(3)/2: each lattice
    # point lands on its rounded voxel only, so the hits sum to the points in range
    rng = np.random.default_rng(4)
    proj = rng.uniform(0, 1, (8, 8, 3))
    ang = rng.uniform(-60, 60, (3, 3))
    r = O.grid_fft(proj, ang, 2.0, 0.9, False)
    assert 0 < r.hits.sum() <= 3 * r.d ** 2
    assert len(r.master_ind) == r.hits.sum()
    assert np.all(r.master_dist <= 0.9)
    r2 = O.grid_fft(proj, ang, 2.0, 0.9, True)   # shiftMax = 1 adds neighbours
    assert r2.hits.sum() > r.hits.sum()


def test_result_is_hermitian(O):
    rng = np.random.default_rng(5)
    proj = rng.uniform(0, 1, (9, 9, 4))
    r = O.grid_fft(proj, rng.uniform(-70, 70, (4, 3)), 2.0, 0.7, True)   # odd Dim2: d = 17
    K = r.measuredK
    assert r.d == 17
    np.testing.assert_allclose(K, np.conj(K[::-1, ::-1, ::-1]), rtol=0, atol=1e-6 * np.abs(K).max())
