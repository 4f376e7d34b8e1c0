"""genfire_rfactor_noalloc (device buffers, scratch in the library's static
device array, nothing allocated) against genfire_rfactor on the same model,
projections and angles: Rtotal, every Rarr entry and every back-projection
equal bit for bit (the same kernels on the same forms), on two calls in a row,
the second smaller; sizes past the scratch limit are refused."""
import numpy as np
import pytest

from genfire_amd import _lib

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _model(d1, d2, d3, seed):
    rng = np.random.default_rng(seed)
    vol = np.zeros((d1, d2, d3), np.float32)
    k = [np.arange(n) - n // 2 for n in (d1, d2, d3)]
    for _ in range(10):
        c = rng.uniform(-0.3, 0.3, 3) * np.array([d1, d2, d3])
        g = [np.exp(-((kk - q) ** 2) / 6.0) for kk, q in zip(k, c)]
        vol += (rng.uniform(0.5, 1.0) * g[0][:, None, None] * g[1][None, :, None] * g[2][None, None, :]).astype(np.float32)
    return vol


def _dev(a):
    t = torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32).ravel(order="F"))).to("cuda")
    return t


def _check(ctx, vol, projs, ang):
    d1, d2, d3 = vol.shape
    p1, p2, P = projs.shape
    R0, Rarr0, sims0 = ctx.rfactor(vol, projs, ang)
    dv, dp = _dev(vol), _dev(projs)
    ds = torch.full((p1 * p2 * P,), -1.0, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    R1, Rarr1 = ctx.rfactor_noalloc(dv.data_ptr(), vol.shape, dp.data_ptr(), projs.shape, ang, ds.data_ptr())
    sims1 = ds.cpu().numpy().reshape(p1, p2, P, order="F")
    assert R1 == R0, (R1, R0)
    np.testing.assert_array_equal(Rarr1, Rarr0)
    np.testing.assert_array_equal(sims1, sims0)
    # without the back-projections
    R2, Rarr2 = ctx.rfactor_noalloc(dv.data_ptr(), vol.shape, dp.data_ptr(), projs.shape, ang)
    assert R2 == R0
    np.testing.assert_array_equal(Rarr2, Rarr0)
    return R0


def test_noalloc_matches_rfactor_two_calls(ctx):
    print("[noalloc] 48 x 48 x 40 model, 9 three-Euler projections, 44 x 40 window", flush=True)
    vol = _model(48, 48, 40, 1)
    rng = np.random.default_rng(2)
    ang = np.stack([rng.uniform(-5, 5, 9), np.linspace(-70, 70, 9), rng.uniform(-5, 5, 9)], 1)
    ang[4] = (60.0, 45.0, 60.0)   # an orientation far from the axes: 8 x 8 patches, x-fastest form
    full = np.stack([ctx.project(vol, *a) for a in ang], 2)
    projs = (full[2:46, 4:44, :] * np.float32(1.01)).astype(np.float32)
    R = _check(ctx, vol, projs, ang)
    assert 0.0 < R < 0.2
    print("[noalloc] second call: 32 x 32 x 24 model, 5 tilts about y (the four-ray form)", flush=True)
    vol2 = _model(32, 32, 24, 3)
    ang2 = np.stack([np.zeros(5), np.linspace(-60, 60, 5), np.zeros(5)], 1)
    projs2 = (np.stack([ctx.project(vol2, *a) for a in ang2], 2) * np.float32(0.97)).astype(np.float32)
    _check(ctx, vol2, projs2, ang2)


def test_noalloc_refuses_beyond_scratch(ctx):
    print("[noalloc] limits", flush=True)
    big = torch.zeros(224 * 224 * 224, dtype=torch.float32, device="cuda")
    pj = torch.zeros(224 * 224, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(_lib.GenfireError, match="scratch"):
        ctx.rfactor_noalloc(big.data_ptr(), (224, 224, 224), pj.data_ptr(), (224, 224, 1), np.array([[3.0, 30.0, 5.0]]))
    small = torch.zeros(8 * 8 * 8, dtype=torch.float32, device="cuda")
    pjs = torch.zeros(8 * 8 * 1025, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(_lib.GenfireError, match="1024 projections"):
        ctx.rfactor_noalloc(small.data_ptr(), (8, 8, 8), pjs.data_ptr(), (8, 8, 1025), np.zeros((1025, 3)))
