"""GPU 3D FFT (fftn / ifftn semantics of reconstruct.m:133,165) against numpy's
pocketfft in FP64, through genfire_fftn."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16, 16), (64, 48, 32), (128, 128, 128), (45, 75, 45), (96, 160, 64),
          (256, 256, 16), (512, 64, 8), (192, 8, 320), (1024, 16, 4), (384, 16, 16),
          (640, 8, 8), (768, 8, 8)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("inverse", [False, True])
def test_fftn_matches_numpy(ctx, shape, inverse):
    print(f"[fft] {shape} inverse={inverse}", flush=True)
    rng = np.random.default_rng(sum(shape) + inverse)
    a = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    got = ctx.fftn(a, inverse=inverse)
    ref = np.fft.ifftn(a.astype(np.complex128)) if inverse else np.fft.fftn(a.astype(np.complex128))
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 2e-6, err


def test_fft_round_trip_512(ctx):
    print("[fft] 512^3 round trip", flush=True)
    rng = np.random.default_rng(3)
    a = rng.standard_normal((512, 512, 512)).astype(np.float32).astype(np.complex64)
    b = ctx.fftn(ctx.fftn(a), inverse=True)
    err = np.linalg.norm((b - a).ravel()) / np.linalg.norm(a.ravel())
    assert err < 2e-6, err
    # Parseval: sum |k|^2 = N sum |a|^2 (sums in FP64)
    k = ctx.fftn(a)
    ek = np.sum(np.abs(k.ravel()).astype(np.float64) ** 2)
    ea = np.sum(np.abs(a.ravel()).astype(np.float64) ** 2)
    assert abs(ek / (a.size * ea) - 1) < 1e-5
