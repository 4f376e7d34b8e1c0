"""Seeded synthetic workloads: Gaussian-blob phantoms and their tilt series.

Workloads (sizes fixed here, every one seeded with numpy PCG64):

* ``tilt64``  -- 64^3 phantom of 300 blobs (sigma U[1,2], amplitude U[0.5,1]) inside a
  sphere of radius 24; 41 single-axis tilts theta = -60:3:60, phi = psi = 0; 1%
  Gaussian noise (sigma = 0.01 * mean projection); oversampling 2, cutoff 0.3,
  100 GENFIRE iterations, enforcement mode 3, spherical support of radius 28.
* ``tilt256`` -- 256^3 phantom of 20000 blobs (sigma 1.5) inside radius 100; 91
  projections at theta = linspace(-70,70,91), phi and psi uniform in +-5 degrees
  with a ZYZ convention (vector3 = [0 0 1]); Poisson noise at 1e4 counts per
  pixel; oversampling 2, cutoff 0.3, 200 DR iterations, R-factor every 10.
* ``tilt512`` -- as tilt256 with a 512^3 phantom of 40000 blobs (radius 200) and
  512 x 512 projections.
* ``tilt512x121`` -- 512^3 phantom of 60000 blobs (radius 200), 121 three-Euler
  projections 512 x 512 with theta in +-72 degrees, 300 DR iterations, R-factor
  every 20 (the projections are meant to be stored as FP16 with a scale per
  projection).

The tilt series are computed by the library's projector (the trilinear
calculate3Dprojection_RealSpaceinterp) from the voxelised phantom.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Workload:
    name: str
    dim: int
    n_blobs: int
    sigma: tuple            # (lo, hi) of the blob sigma in voxels
    radius: float           # blob centres lie inside this sphere
    num_projs: int
    three_euler: bool
    tilt_max: float
    noise: str              # 'gauss1pct' | 'poisson1e4'
    oversampling: float
    cutoff: float
    iterations: int
    algorithm: str          # 'genfire' | 'dr'
    support_radius: float
    monitor_R: int = 0
    monitor_every: int = 10
    mode: int = 3
    seed: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def vectors(self):
        v3 = (0.0, 0.0, 1.0) if self.three_euler else (1.0, 0.0, 0.0)
        return (0.0, 0.0, 1.0), (0.0, 1.0, 0.0), v3


WORKLOADS = {
    "tilt64": Workload("tilt64", 64, 300, (1.0, 2.0), 24.0, 41, False, 60.0, "gauss1pct", 2.0, 0.3,
                       100, "genfire", 28.0, seed=1),
    "tilt256": Workload("tilt256", 256, 20000, (1.5, 1.5), 100.0, 91, True, 70.0, "poisson1e4", 2.0,
                        0.3, 200, "dr", 110.0, monitor_R=1, monitor_every=10, seed=2),
    "tilt512": Workload("tilt512", 512, 40000, (1.5, 1.5), 200.0, 91, True, 70.0, "poisson1e4", 2.0,
                        0.3, 200, "dr", 220.0, monitor_R=1, monitor_every=10, seed=3),
    "tilt512x121": Workload("tilt512x121", 512, 60000, (1.5, 1.5), 200.0, 121, True, 72.0, "poisson1e4", 2.0,
                            0.3, 300, "dr", 220.0, monitor_R=1, monitor_every=20, seed=4),
}


def blob_phantom(dim: int, n_blobs: int, sigma, radius: float, rng: np.random.Generator) -> np.ndarray:
    """Sum of isotropic Gaussian blobs sampled on the voxel grid (centred coordinates
    x = i - floor(dim/2)); centres uniform inside the sphere by rejection."""
    c = dim // 2
    centres = np.zeros((0, 3))
    while len(centres) < n_blobs:
        cand = rng.uniform(-radius, radius, size=(2 * n_blobs, 3))
        cand = cand[np.sum(cand * cand, axis=1) <= radius * radius]
        centres = np.concatenate([centres, cand])[:n_blobs]
    amps = rng.uniform(0.5, 1.0, size=n_blobs)
    lo, hi = sigma
    sig = rng.uniform(lo, hi, size=n_blobs) if hi > lo else np.full(n_blobs, float(lo))
    vol = np.zeros((dim, dim, dim), np.float32)
    for (x, y, z), a, s in zip(centres, amps, sig):
        h = int(np.ceil(4.0 * s))
        idx = []
        prof = []
        for q in (x, y, z):
            i0 = max(int(np.floor(q)) - h, -c)
            i1 = min(int(np.floor(q)) + h + 1, dim - 1 - c)
            if i1 < i0:
                break
            k = np.arange(i0, i1 + 1)
            idx.append(slice(i0 + c, i1 + c + 1))
            prof.append(np.exp(-((k - q) ** 2) / (2.0 * s * s)))
        if len(idx) < 3:
            continue
        block = a * prof[0][:, None, None] * prof[1][None, :, None] * prof[2][None, None, :]
        vol[idx[0], idx[1], idx[2]] += block.astype(np.float32)
    return vol


def tilt_angles(w: Workload, rng: np.random.Generator) -> np.ndarray:
    """P x 3 [phi theta psi] in degrees."""
    P = w.num_projs
    if w.name == "tilt64":
        theta = np.arange(-60.0, 60.0 + 1e-9, 3.0)
    else:
        theta = np.linspace(-w.tilt_max, w.tilt_max, P)
    ang = np.zeros((P, 3))
    ang[:, 1] = theta
    if w.three_euler:
        ang[:, 0] = rng.uniform(-5.0, 5.0, P)
        ang[:, 2] = rng.uniform(-5.0, 5.0, P)
    return ang


def add_noise(projs: np.ndarray, kind: str, rng: np.random.Generator) -> np.ndarray:
    if kind == "gauss1pct":
        s = 0.01 * float(np.mean(projs))
        return (projs + rng.normal(0.0, s, projs.shape)).astype(np.float32)
    if kind == "poisson1e4":
        scale = 1e4 / max(float(np.mean(projs)), 1e-30)
        lam = np.clip(projs.astype(np.float64) * scale, 0.0, None)
        return (rng.poisson(lam) / scale).astype(np.float32)
    return projs.astype(np.float32)


def spherical_support(dim: int, radius: float) -> np.ndarray:
    c = dim // 2
    k = np.arange(dim) - c
    r2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    return r2 <= radius * radius


def make_inputs(w: Workload, project_many, seed: int | None = None):
    """(phantom, projections Dim x Dim x P, angles P x 3, support) for workload w.
    ``project_many(vol, angles) -> Dim x Dim x P`` computes the clean tilt series."""
    rng = np.random.Generator(np.random.PCG64(w.seed if seed is None else seed))
    vol = blob_phantom(w.dim, w.n_blobs, w.sigma, w.radius, rng)
    ang = tilt_angles(w, rng)
    clean = np.asarray(project_many(vol, ang), np.float32)
    projs = add_noise(clean, w.noise, rng)
    sup = spherical_support(w.dim, w.support_radius)
    return vol, projs, ang, sup


# The serving mix: a Poisson stream of reconstruction requests of uneven size.
MIX_SIZES = (64, 96, 128, 192, 256, 320)


@dataclass
class MixRequest:
    arrival_s: float
    dim: int
    num_projs: int
    three_euler: bool
    iterations: int
    seed: int


def serving_mix(n_requests: int, rate_per_s: float = 4.0, seed: int = 7):
    """Seeded request stream: exponential inter-arrival times at rate_per_s; volume
    N uniform over MIX_SIZES, projections uniform in [31, 91], single-axis or
    three-Euler tilts within +-70 degrees, 50-200 DR iterations."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for i in range(n_requests):
        t += float(rng.exponential(1.0 / rate_per_s))
        out.append(MixRequest(t, int(rng.choice(MIX_SIZES)), int(rng.integers(31, 92)), bool(rng.integers(0, 2)),
                              int(rng.integers(50, 201)), seed * 1000 + i))
    return out


def mix_workload(r: MixRequest) -> Workload:
    """Workload of one serving-mix request: 1 blob per 800 voxels (sigma 1.5)
    inside a sphere of radius 0.39 N, support radius 0.43 N, Poisson noise."""
    N = r.dim
    return Workload(f"mix{N}", N, max(1, N ** 3 // 800), (1.5, 1.5), 0.39 * N, r.num_projs, r.three_euler, 70.0,
                    "poisson1e4", 2.0, 0.3, r.iterations, "dr", 0.43 * N, seed=r.seed)


def gpu_project_many(ctx):
    """project_many for make_inputs using the library's batched projector."""
    def f(vol, ang):
        d1, d2, _ = vol.shape
        dummy = np.ones((d1, d2, len(ang)), np.float32)
        _, _, sims = ctx.rfactor(vol, dummy, ang)
        return sims
    return f
