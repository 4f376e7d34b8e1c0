"""Pins the oracle's cubic-convolution projector (interp3 'cubic',
src/calculate3Dprojection_RealSpaceinterp.m:45,51) without MATLAB.

MATLAB's kernel for uniform grids is Keys' cubic convolution (a = -1/2) with
Keys' boundary nodes; both reproduce quadratics exactly, so a quadratic volume
is interpolated exactly everywhere in the node box, edges included: a wrong
weight, sign or boundary rule fails that.  An independent numpy restatement
(1D convolutions along each axis in turn) checks the tensor-product order and
the cell choice on random data; the interpolant equals the nodes on the grid;
the projection of a quadratic volume equals the ray sums of the polynomial.
Agreement with MATLAB itself is parity unpinned (MATLAB is absent)."""
import numpy as np
import pytest


def quad_volume(shape, c):
    x, y, z = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    return (c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * x + c[5] * y * y + c[6] * z * z
            + c[7] * x * y + c[8] * y * z + c[9] * x * z), lambda p: (
        c[0] + c[1] * p[:, 0] + c[2] * p[:, 1] + c[3] * p[:, 2] + c[4] * p[:, 0] ** 2 + c[5] * p[:, 1] ** 2
        + c[6] * p[:, 2] ** 2 + c[7] * p[:, 0] * p[:, 1] + c[8] * p[:, 1] * p[:, 2] + c[9] * p[:, 0] * p[:, 2])


def test_cubic_reproduces_quadratics_edges_included(O):
    rng = np.random.default_rng(0)
    shape = (9, 7, 8)
    c = rng.uniform(-1, 1, 10) * np.array([8, 1, 1, 1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.1])
    vol, q = quad_volume(shape, c)
    # values exact in FP32 would need integers; keep FP32 rounding out of the
    # comparison by building the polynomial from the FP32 nodes' own values
    vol32 = vol.astype(np.float32)
    pts = rng.uniform(0, 1, (4000, 3)) * (np.array(shape) - 1)
    pts = np.concatenate([pts, np.array([[0, 0, 0], [8, 6, 7], [8, 0, 3.5], [0.25, 6, 7], [7.9, 5.99, 0.01]])])
    got = O.interp_cubic(vol32, pts)
    # the interpolant is linear in the node values: interpolate the FP32
    # rounding error separately and add it to the exact polynomial
    err_nodes = vol32.astype(np.float64) - vol
    got_err = O.interp_cubic(err_nodes.astype(np.float32), pts)
    np.testing.assert_allclose(got, q(pts) + got_err, rtol=0, atol=2e-5)
    # integer coefficients keep every node exact in FP32: then exactness is direct
    ci = np.array([3, -2, 1, 4, 1, -1, 2, 1, -2, 1], np.float64)
    voli, qi = quad_volume(shape, ci)
    np.testing.assert_allclose(O.interp_cubic(voli.astype(np.float32), pts), qi(pts), rtol=0, atol=1e-9)


def test_cubic_interpolates_the_nodes_and_is_zero_outside(O):
    rng = np.random.default_rng(1)
    vol = rng.uniform(-1, 1, (6, 5, 7)).astype(np.float32)
    idx = np.array([[i, j, k] for i in range(6) for j in range(5) for k in range(7)], np.float64)
    np.testing.assert_allclose(O.interp_cubic(vol, idx), vol[idx[:, 0].astype(int), idx[:, 1].astype(int),
                                                             idx[:, 2].astype(int)], rtol=0, atol=1e-12)
    outside = np.array([[-1e-9, 2, 3], [5 + 1e-9, 2, 3], [2, 4.5, 3], [2, 2, -0.5], [2, 2, 6.0001]])
    assert np.all(O.interp_cubic(vol, outside) == 0.0)


def _keys(t):
    t = np.asarray(t, np.float64)
    return np.stack([((-t + 2) * t - 1) * t / 2, ((3 * t - 5) * t * t + 2) / 2, ((-3 * t + 4) * t + 1) * t / 2,
                     (t - 1) * t * t / 2], -1)


def _numpy_cubic(vol, pts):
    """Independent restatement: 1D cubic convolution along z, then y, then x,
    on the nodes extended by Keys' rule one axis at a time."""
    v = vol.astype(np.float64)
    for ax in range(3):
        lo = 3 * np.take(v, [0], ax) - 3 * np.take(v, [1], ax) + np.take(v, [2], ax)
        hi = 3 * np.take(v, [-1], ax) - 3 * np.take(v, [-2], ax) + np.take(v, [-3], ax)
        v = np.concatenate([lo, v, hi], ax)
    out = np.zeros(len(pts))
    n = np.array(vol.shape)
    for q, p in enumerate(pts):
        if np.any(p < 0) or np.any(p > n - 1):
            continue
        i0 = np.minimum(np.floor(p).astype(int), n - 2)
        w = _keys(p - i0)
        blk = v[i0[0]:i0[0] + 4, i0[1]:i0[1] + 4, i0[2]:i0[2] + 4]   # padded index of node i0 - 1 is i0
        out[q] = np.einsum("abc,a,b,c->", blk, w[0], w[1], w[2])
    return out


def test_cubic_matches_an_independent_restatement(O):
    rng = np.random.default_rng(2)
    vol = rng.uniform(-1, 1, (11, 9, 10)).astype(np.float32)
    pts = rng.uniform(-0.5, 1.05, (3000, 3)) * (np.array(vol.shape) - 1)
    np.testing.assert_allclose(O.interp_cubic(vol, pts), _numpy_cubic(vol, pts), rtol=0, atol=1e-12)


@pytest.mark.parametrize("angles", [(0.0, 0.0, 0.0), (12.0, -47.0, 33.0), (3.0, 65.0, -4.0)])
def test_cubic_projection_of_a_quadratic_is_its_ray_sum(O, angles):
    shape = (12, 10, 11)
    ci = np.array([3, -2, 1, 4, 1, -1, 2, 1, -2, 1], np.float64)
    vol, q = quad_volume(shape, ci)
    got = O.project(vol.astype(np.float32), *angles, interp="cubic")
    R = O.zyz(*angles)
    d = np.array(shape)
    c = d // 2
    ref = np.zeros(shape[:2])
    for i in range(shape[0]):
        for j in range(shape[1]):
            Z = np.arange(shape[2]) - c[2]
            P = np.stack([np.full_like(Z, i - c[0]), np.full_like(Z, j - c[1]), Z], 1).astype(np.float64)
            f = (P @ R.T) + c
            inside = np.all((f >= 0) & (f <= d - 1), 1)
            ref[i, j] = q(f[inside]).sum()
    np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-9)


def test_cubic_projector_axis_aligned_sums(O):
    v = np.random.default_rng(3).uniform(0, 1, (8, 9, 7)).astype(np.float32)
    np.testing.assert_allclose(O.project(v, 0, 0, 0, interp="cubic"), v.astype(np.float64).sum(axis=2),
                               rtol=1e-12)
