"""CPU oracle for the GENFIRE v2 reconstruction hot path.

TEST INFRASTRUCTURE ONLY: this module is the checker that the HIP library in
``genfire_amd/`` is compared against.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it; the product never does.

It restates, function by function, the MATLAB reference (paths relative to the
reference repository):

* ``@GENFIRE_Reconstructor/fillInFourierGrid_DFT.m:85-310``  -> :func:`grid_dft`
  (index pass in the FP32 reference order via ``genfire_oracle.c``; values as an
  FP64 direct DFT; FP64 weighted average, :267-291)
* ``@GENFIRE_Reconstructor/CentroSymmetricity.m:2-72``       -> :func:`centro_symmetricity`
* ``@GENFIRE_Reconstructor/fillInFourierGrid.m:27-163``       -> :func:`grid_fft`
  (FFT gridding; ``src/hermitianSymmetrize.m`` -> :func:`hermitian_symmetrize`,
  ``src/my_fft.m`` -> :func:`my_fft`)
* ``@GENFIRE_Reconstructor/GENFIRE_Reconstructor.m:115-186`` -> :func:`delay_indicators`,
  :func:`oversampled_size`, :func:`volumn_index`
* ``@GENFIRE_Reconstructor/reconstruct.m:28-194``            -> :func:`reconstruct`
* ``src/calculate3Dprojection_RealSpaceinterp.m``             -> :func:`project`
* ``src/My_paddzero.m``, ``src/My_stripzero.m``, ``src/My_volumn_index.m``,
  ``src/My_iffshift3_ind.m``, ``src/make_Kspace_indices.m``, ``src/my_ifft.m``

Parity status per function is recorded in each docstring; "parity unpinned" marks
what no golden vector, closed form or invariant fixes.

Array convention: a MATLAB array of size (n1, n2, n3) is a numpy array of the same
shape; MATLAB linear indices are the 0-based Fortran-order ravel index.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np
import scipy.fft

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def lib():
    """Load oracle/liboracle.so (built by ``make -C oracle``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        L = ctypes.CDLL(path)
        L.go_rotation.argtypes = [_f64p, _f64p, _f64p, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, _f32p, _f64p]
        L.go_rotation.restype = None
        L.go_select.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_double,
                                ctypes.c_int64, _i64p, _f32p, _f32p, _f32p,
                                ctypes.POINTER(ctypes.c_int64)]
        L.go_select.restype = ctypes.c_int64
        L.go_tie_census.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_double,
                                    ctypes.c_int, _i64p]
        L.go_tie_census.restype = None
        L.go_lu_solve.argtypes = [_f32p, _f32p, _f32p]
        L.go_lu_solve.restype = None
        for name in ("go_dft", "go_dft_brute"):
            fn = getattr(L, name)
            fn.argtypes = [ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_int64, _f32p, _f32p, _f64p]
            fn.restype = None
        L.go_zyz.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _f64p]
        L.go_zyz.restype = None
        L.go_project.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p, _f64p]
        L.go_project.restype = None
        L.go_project_cubic.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p, _f64p]
        L.go_project_cubic.restype = None
        L.go_interp_cubic.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_int64, _f64p]
        L.go_interp_cubic.restype = None
        L.go_nufft_interp.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_int64, _f32p, _f32p, _f64p]
        L.go_nufft_interp.restype = None
        L.go_set_threads.argtypes = [ctypes.c_int]
        L.go_set_threads.restype = None
        _LIB = L
    return _LIB


# ---------------------------------------------------------------------------
# MATLAB semantics (Appendix-A style helpers)
# ---------------------------------------------------------------------------

def matlab_round(x):
    """MATLAB ``round``: half away from zero."""
    x = np.asarray(x, dtype=np.float64)
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def matlab_linspace(d1: float, d2: float, n: int) -> np.ndarray:
    """MATLAB ``linspace(d1,d2,n)``: ``d1 + (0:n1).*(d2-d1)./n1`` with the endpoint
    set exactly to d2 (MathWorks linspace.m, finite case)."""
    n = int(math.floor(n))
    if n < 1:
        return np.zeros(0)
    if n == 1:
        return np.array([float(d2)])
    n1 = n - 1
    k = np.arange(n1 + 1, dtype=np.float64)
    y = d1 + (k * (d2 - d1)) / n1
    if d1 == d2:
        y[:] = d1
    else:
        y[-1] = d2
    return y


def matlab_colon(a: float, d: float, b: float) -> np.ndarray:
    """MATLAB ``a:d:b`` for non-integer steps, restated from MathWorks' published
    colon algorithm (colonop): n = round((b-a)/d), reduced by one if a+n*d overshoots
    b by more than tol = 2*eps*max(|a|,|b|); the right end c = a+n*d snaps to b
    within tol; elements k <= n/2 are a+k*d, the others c-(n-k)*d, and for even n
    the middle one is (a+c)/2.  Parity unpinned (no MATLAB here); only
    constraintEnforcementMode 1 and 2 use it."""
    tol = 2.0 * np.finfo(float).eps * max(abs(a), abs(b))
    sig = math.copysign(1.0, d)
    n = int(matlab_round((b - a) / d))
    if sig * (a + n * d - b) > tol:
        n -= 1
    if n < 0:
        return np.zeros(0)
    c = a + n * d
    if sig * (c - b) > -tol:
        c = b
    v = np.zeros(n + 1)
    for k in range(0, n // 2 + 1):
        v[k] = a + k * d
        v[n - k] = c - k * d
    if n % 2 == 0:
        v[n // 2] = (a + c) / 2
    return v


def delay_indicators(mode: int) -> np.ndarray:
    """constraintEnforcementDelayIndicators, GENFIRE_Reconstructor.m:156-165."""
    if mode == 1:
        return np.concatenate([matlab_colon(0.95, -0.1, -0.15), [-10.0, -10.0, -10.0],
                               matlab_colon(-0.15, 0.1, 0.95)])
    if mode == 2:
        return np.concatenate([matlab_colon(0.95, -0.1, -0.15), [-10.0, -10.0]])
    if mode == 3:
        return np.array([-999.0] * 5)
    raise ValueError(f"GENFIRE: ERROR! constraintEnforcementMode value {mode} not understood")


def cutoff_schedule(num_iterations: int, num_delays: int) -> np.ndarray:
    """iterationNumsToChangeCutoff = round(linspace(1,N,numel(delays))), reconstruct.m:103."""
    return matlab_round(matlab_linspace(1.0, float(num_iterations), num_delays)).astype(np.int64)


def oversampled_size(dim: int, ratio: float) -> int:
    """n_os = round(Dim*ratio), GENFIRE_Reconstructor.m:168-169."""
    return int(matlab_round(dim * ratio))


def find(mask: np.ndarray) -> np.ndarray:
    """MATLAB ``find`` on an array: ascending 0-based Fortran-order linear indices."""
    return np.flatnonzero(np.asarray(mask).ravel(order="F"))


def iffshift3_ind(shape, inds: np.ndarray) -> np.ndarray:
    """src/My_iffshift3_ind.m:1-7 on 0-based indices: s = (c + ceil(n/2)) mod n per dim."""
    s1, s2, s3 = np.unravel_index(inds, shape, order="F")
    s1 = (s1 + (shape[0] + 1) // 2) % shape[0]
    s2 = (s2 + (shape[1] + 1) // 2) % shape[1]
    s3 = (s3 + (shape[2] + 1) // 2) % shape[2]
    return np.ravel_multi_index((s1, s2, s3), shape, order="F")


def _pad_start(big: int, small: int) -> int:
    """0-based start of the centre-preserving window (My_paddzero.m:33-49,
    My_stripzero.m:30-46, My_volumn_index.m:12-28 all reduce to this)."""
    if big % 2 == 0:
        return (big - small) // 2 if small % 2 == 0 else (big - small + 1) // 2
    return (big - small - 1) // 2 if small % 2 == 0 else (big - small) // 2


def paddzero(vol: np.ndarray, shape) -> np.ndarray:
    """src/My_paddzero.m:24-59 (output single, :30)."""
    out = np.zeros(shape, dtype=np.float32)
    sl = tuple(slice(_pad_start(B, s), _pad_start(B, s) + s) for B, s in zip(shape, vol.shape))
    out[sl] = vol
    return out


def stripzero(vol: np.ndarray, shape) -> np.ndarray:
    """src/My_stripzero.m:22-55."""
    sl = tuple(slice(_pad_start(B, s), _pad_start(B, s) + s) for B, s in zip(vol.shape, shape))
    return vol[sl]


def volumn_index(big, small) -> np.ndarray:
    """src/My_volumn_index.m:3-34, 1-based [start end] rows like the reference."""
    return np.array([[_pad_start(B, s) + 1, _pad_start(B, s) + s] for B, s in zip(big, small)])


def make_kspace_indices(shape) -> np.ndarray:
    """src/make_Kspace_indices.m:16-62: Q = sqrt(Kx.^2+Ky.^2+Kz.^2) in double with
    Kx along dim 2 (meshgrid(vec2,vec1,vec3)), evaluated left to right."""
    vecs = []
    for n in shape:
        if n % 2 == 0:
            c = n // 2
            vecs.append(np.arange(-c, c, dtype=np.float64) / c)
        elif n == 1:
            vecs.append(np.zeros(1))
        else:
            c = (n - 1) // 2
            vecs.append(np.arange(-c, c + 1, dtype=np.float64) / c)
    v1 = vecs[0][:, None, None]
    v2 = vecs[1][None, :, None]
    v3 = vecs[2][None, None, :]
    return np.sqrt((v2 * v2 + v1 * v1) + v3 * v3)


def my_ifft(k: np.ndarray) -> np.ndarray:
    """src/my_ifft.m:10-12: fftshift(ifftn(ifftshift(k)))."""
    return np.fft.fftshift(scipy.fft.ifftn(np.fft.ifftshift(k), workers=-1))


# ---------------------------------------------------------------------------
# Gridding (kernel a)
# ---------------------------------------------------------------------------

def rotation(phi: float, theta: float, psi: float, v1=(0, 0, 1), v2=(0, 1, 0), v3=(1, 0, 0)):
    """fillInFourierGrid_DFT.m:108-118.  Returns (R_single, R_double), each (3,3)
    with R[i, j] = row i, column j."""
    Rf = np.zeros(9, np.float32)
    Rd = np.zeros(9, np.float64)
    lib().go_rotation(np.asarray(v1, np.float64), np.asarray(v2, np.float64),
                      np.asarray(v3, np.float64), float(phi), float(theta), float(psi), Rf, Rd)
    return Rf.reshape(3, 3, order="F"), Rd.reshape(3, 3, order="F")


def half_ranges(n1: int, n2: int, centro: bool):
    """fillInFourierGrid_DFT.m:122-136: (lo, len) per dimension of the k grid."""
    if centro:
        return (-(n1 // 2), -(n2 // 2), -(n1 // 2)), (n1 // 2 + 1, 2 * (n2 // 2) + 1, 2 * (n1 // 2) + 1)
    return (-(n1 // 2), -(n2 // 2), -(n1 // 2)), (n1, n2, n1)


def select(n1: int, n2: int, centro: bool, R: np.ndarray, cutoff: float):
    """fillInFourierGrid_DFT.m:175-192 for one projection, whole-grid scan.
    Returns (Gind, D, u, v, n_wrong); Gind ascending 0-based half-grid indices."""
    _, ln = half_ranges(n1, n2, centro)
    cap = ln[0] * ln[1] * ln[2]
    # the slab |D| < cutoff holds at most a few times cutoff*area points; grow on demand
    cap_try = min(cap, max(1 << 16, int(8 * (2 * cutoff + 2) * max(ln) ** 2)))
    Rf = np.ascontiguousarray(np.asarray(R, np.float32).ravel(order="F"))
    while True:
        ind = np.empty(cap_try, np.int64)
        D = np.empty(cap_try, np.float32)
        u = np.empty(cap_try, np.float32)
        v = np.empty(cap_try, np.float32)
        wrong = ctypes.c_int64(0)
        m = lib().go_select(n1, n2, int(bool(centro)), Rf, float(cutoff), cap_try, ind, D, u, v,
                            ctypes.byref(wrong))
        if m >= 0:
            return ind[:m].copy(), D[:m].copy(), u[:m].copy(), v[:m].copy(), int(wrong.value)
        cap_try = min(cap, cap_try * 4)


def tie_census(n1: int, n2: int, centro: bool, R: np.ndarray, cutoff: float, ulps: int = 1):
    """Near-ties of the two FP32 index decisions of fillInFourierGrid_DFT.m
    (:178 D < cutoff, :190 |u| <= n1/2 and |v| <= n2/2) for one projection:
    (grid points whose D lies within `ulps` ulps of the cutoff, points with
    D < cutoff whose |u| or |v| lies within `ulps` ulps of n/2, points kept,
    points with D < cutoff whose |u| or |v| equals n/2 exactly)."""
    out = np.zeros(4, np.int64)
    Rf = np.ascontiguousarray(np.asarray(R, np.float32).ravel(order="F"))
    lib().go_tie_census(n1, n2, int(bool(centro)), Rf, float(cutoff), int(ulps), out)
    return int(out[0]), int(out[1]), int(out[2]), int(out[3])


def lu_solve(R: np.ndarray, b) -> np.ndarray:
    x = np.zeros(3, np.float32)
    lib().go_lu_solve(np.ascontiguousarray(np.asarray(R, np.float32).ravel(order="F")),
                      np.asarray(b, np.float32), x)
    return x


def dft64(proj: np.ndarray, un: np.ndarray, vn: np.ndarray, brute: bool = False) -> np.ndarray:
    """FP64 direct DFT of one projection at in-plane points (un, vn) (:236-248)."""
    d1, d2 = proj.shape
    P = np.ascontiguousarray(np.asarray(proj, np.float32).ravel(order="F"))
    un = np.ascontiguousarray(un, np.float32)
    vn = np.ascontiguousarray(vn, np.float32)
    out = np.zeros(2 * len(un), np.float64)
    fn = lib().go_dft_brute if brute else lib().go_dft
    fn(d1, d2, P, len(un), un, vn, out)
    return out[0::2] + 1j * out[1::2]


def _es_hat(k: np.ndarray, M: int, w: int, beta: float) -> np.ndarray:
    """Fourier transform of the exponential-of-semicircle kernel at frequency
    k/M: 2 int_0^{w/2} phi(t) cos(2 pi k t / M) dt by 200-node Gauss-Legendre
    quadrature after t = (w/2) sin(s), which removes the square-root edge."""
    s, wt = np.polynomial.legendre.leggauss(200)
    s = (s + 1.0) * (np.pi / 4.0)              # s in (0, pi/2)
    wt = wt * (np.pi / 4.0)
    t = 0.5 * w * np.sin(s)
    phi = np.exp(beta * (np.sqrt(1.0 - (2.0 * t / w) ** 2) - 1.0)) * 0.5 * w * np.cos(s)
    return 2.0 * np.cos(2.0 * np.pi * np.outer(np.asarray(k, np.float64), t) / M) @ (wt * phi)


def nufft2d2(proj: np.ndarray, un: np.ndarray, vn: np.ndarray, w: int = 8, sigma: float = 2.0,
             workers: int = -1) -> np.ndarray:
    """FP64 type-2 NUFFT of one projection at the in-plane points (un, vn): the
    same sum as dft64 (fillInFourierGrid_DFT.m:236-248), by the route the
    reference takes on the CPU (nufft2d2, :249-254, a library absent here):
    divide each mode by the kernel's Fourier transform, zero-pad to an
    M1 x M2 fine grid (M >= sigma Dim), FFT, and interpolate every point from
    its w x w nearest nodes (go_nufft_interp).  With w = 8 and sigma = 2 the
    relative error is below 1e-6 (tests/test_oracle_grid.py)."""
    d1, d2 = proj.shape
    M1 = scipy.fft.next_fast_len(int(np.ceil(sigma * d1)))
    M2 = scipy.fft.next_fast_len(int(np.ceil(sigma * d2)))
    beta = 0.97 * np.pi * w * (1.0 - 1.0 / (2.0 * sigma))
    k1 = np.arange(d1) - d1 // 2
    k2 = np.arange(d2) - d2 // 2
    h1 = _es_hat(k1, M1, w, beta)
    h2 = _es_hat(k2, M2, w, beta)
    fine = np.zeros((M1, M2), np.complex128)
    fine[np.ix_(k1 % M1, k2 % M2)] = np.asarray(proj, np.float64) / np.outer(h1, h2)
    G = scipy.fft.fft2(fine, workers=workers, overwrite_x=True)
    out = np.zeros(2 * len(un), np.float64)
    lib().go_nufft_interp(np.ascontiguousarray(G).view(np.float64).reshape(-1), M1, M2, w, beta, len(un),
                          np.ascontiguousarray(un, np.float32), np.ascontiguousarray(vn, np.float32), out)
    return out[0::2] + 1j * out[1::2]


def set_threads(n: int) -> None:
    """OpenMP threads of the oracle's C loops."""
    lib().go_set_threads(int(n))


@dataclass
class GridResult:
    n1: int
    n2: int
    half_shape: tuple
    master_ind: np.ndarray      # 0-based half-grid linear index per point
    master_proj: np.ndarray     # projection number per point (0-based)
    master_dist: np.ndarray     # D per point (single)
    master_un: np.ndarray       # u/n1 per point (single)
    master_vn: np.ndarray       # v/n2 per point (single)
    master_val: np.ndarray      # F per point (complex128 from the FP64 DFT)
    hits: np.ndarray            # per half-grid voxel hit count (int32)
    FS: np.ndarray              # half grid after averaging (complex128)
    measuredK: np.ndarray       # centred full grid, complex64 (runGridding single cast)
    n_wrong: int = 0
    extra: dict = field(default_factory=dict)


def weighted_average(master_ind, master_dist, master_val, total: int):
    """fillInFourierGrid_DFT.m:267-291 with FP64 accumulation:
    w = 1/D (0 where D < 1e-7); FS = sum(w F)/sum(w) where sum(w) > 0;
    then FS(exact) = F_exact assigned in list order (the last one wins)."""
    dist = np.asarray(master_dist, np.float32)
    index_exact = dist.astype(np.float64) < 1e-7          # :271
    w = np.zeros(len(dist), np.float32)
    w[~index_exact] = np.float32(1.0) / dist[~index_exact]  # :275, single reciprocal
    w = w.astype(np.float64)
    vals = np.asarray(master_val, np.complex128)
    # accumarray in list order: bincount adds its weights in input order
    FS = (np.bincount(master_ind, weights=vals.real * w, minlength=total)
          + 1j * np.bincount(master_ind, weights=vals.imag * w, minlength=total))   # :285
    sw = np.bincount(master_ind, weights=w, minlength=total)                        # :286
    interp = sw > 0                                         # :287
    FS[interp] = FS[interp] / sw[interp]                    # :289
    ex_ind = master_ind[index_exact]
    FS[ex_ind] = vals[index_exact]                          # :291 (numpy: last write wins)
    return FS


def centro_symmetricity(FS: np.ndarray, n1: int, n_k1: int, n2: int, dtype=np.complex64):
    """CentroSymmetricity.m:2-72, restated literally with reshapes and flips
    (fliplr = dim 2, the permute/flipud/permute = dim 3, flipud = dim 1)."""
    out = np.zeros((n1, n2, n1), dtype=dtype)

    def flip_all(a):
        a = a[:, ::-1, :]      # fliplr
        a = a[:, :, ::-1]      # permute(flipud(permute(.,[3 1 2])),[2 3 1])
        return a[::-1, :, :]   # flipud

    if n2 % 2 == 1:
        if n1 % 2 == 1:
            half = FS.reshape(n_k1 + 1, n2, n1, order="F")
            out[: n_k1 + 1] = half
            out[n_k1 + 1:] = np.conj(flip_all(half[:-1]))
        else:
            ext = FS.reshape(n_k1 + 1, n2, n1 + 1, order="F")
            out[: n_k1 + 1] = ext[:, :, :-1]
            out[n_k1 + 1:] = np.conj(flip_all(ext[1:-1])[:, :, :-1])
    else:
        if n1 % 2 == 1:
            ext = FS.reshape(n_k1 + 1, n2 + 1, n1, order="F")
            out[: n_k1 + 1] = ext[:, :-1, :]
            out[n_k1 + 1:] = np.conj(flip_all(ext[:-1])[:, :-1, :])
        else:
            ext = FS.reshape(n_k1 + 1, n2 + 1, n1 + 1, order="F")
            out[: n_k1 + 1] = ext[:, :-1, :-1]
            out[n_k1 + 1:] = np.conj(flip_all(ext[1:-1])[:, :-1, :-1])
    return out


def grid_dft(projections: np.ndarray, angles: np.ndarray, n1: int, n2: int, cutoff: float,
             centro: bool = True, vectors=((0, 0, 1), (0, 1, 0), (1, 0, 0)),
             values: bool = True, value_fn=None) -> GridResult:
    """fillInFourierGrid_DFT.m:85-310 + runGridding's single cast (GENFIRE_Reconstructor.m:182).

    projections: (Dim1, Dim2, P) array; angles: (P, 3) [phi theta psi] degrees.
    ``value_fn(p, proj, un, vn)`` may replace the FP64 DFT (e.g. to feed the GPU's
    point values through the same averaging); ``values=False`` skips values."""
    P = projections.shape[2]
    lo, ln = half_ranges(n1, n2, centro)
    total = ln[0] * ln[1] * ln[2]
    inds, dists, uns, vns, vals, projs = [], [], [], [], [], []
    n_wrong = 0
    for p in range(P):
        Rf, _ = rotation(angles[p, 0], angles[p, 1], angles[p, 2], *vectors)
        ind, D, u, v, wr = select(n1, n2, centro, Rf, cutoff)
        n_wrong += wr
        un = (u / np.float32(n1)).astype(np.float32)   # :237, single / n1
        vn = (v / np.float32(n2)).astype(np.float32)   # :238
        inds.append(ind); dists.append(D); uns.append(un); vns.append(vn)
        projs.append(np.full(len(ind), p, np.int32))
        if values:
            proj = np.asarray(projections[:, :, p], np.float32)   # gpuArray product is single
            if value_fn is not None:
                vals.append(np.asarray(value_fn(p, proj, un, vn), np.complex128))
            else:
                vals.append(dft64(proj, un, vn))
    master_ind = np.concatenate(inds) if inds else np.zeros(0, np.int64)
    master_dist = np.concatenate(dists) if dists else np.zeros(0, np.float32)
    hits = np.bincount(master_ind, minlength=total).astype(np.int32)
    res = GridResult(n1=n1, n2=n2, half_shape=tuple(ln), master_ind=master_ind,
                     master_proj=np.concatenate(projs) if projs else np.zeros(0, np.int32),
                     master_dist=master_dist,
                     master_un=np.concatenate(uns) if uns else np.zeros(0, np.float32),
                     master_vn=np.concatenate(vns) if vns else np.zeros(0, np.float32),
                     master_val=np.concatenate(vals) if vals else np.zeros(0, np.complex128),
                     hits=hits, FS=None, measuredK=None, n_wrong=n_wrong)
    if values:
        FS = weighted_average(master_ind, master_dist, res.master_val, total)
        res.FS = FS
        if centro:
            res.measuredK = centro_symmetricity(FS, n1, n1 // 2, n2, dtype=np.complex128).astype(np.complex64)
        else:
            res.measuredK = FS.reshape(n1, n2, n1, order="F").astype(np.complex64)
    return res


def rec_ifft(measuredK: np.ndarray, dim1: int, dim2: int) -> np.ndarray:
    """recIFFT = My_stripzero(real(my_ifft(measuredK)), [Dim1 Dim2 Dim1]), :304."""
    return stripzero(np.real(my_ifft(measuredK)), (dim1, dim2, dim1))


# ---------------------------------------------------------------------------
# FFT gridding (griddingMethod = 1)
# ---------------------------------------------------------------------------

def sincosd(x: float):
    """MATLAB sind/cosd: the angle reduced in degrees (nearest multiple of 90),
    so multiples of 90 give exact 0 and +-1."""
    r = math.fmod(float(x), 360.0)
    q = round(r / 90.0)
    rem = r - 90.0 * q
    sr, cr = math.sin(rem * math.pi / 180.0), math.cos(rem * math.pi / 180.0)
    return [(sr, cr), (cr, -sr), (-sr, -cr), (-cr, sr)][int(q) % 4]


def fft_grid_rotation(phi: float, theta: float, psi: float) -> np.ndarray:
    """fillInFourierGrid.m:80-84 (the RELION/XMIPP Euler matrix), FP64, R[i, j] = row i.
    Evaluated left to right as written (no fused multiply-add)."""
    sp, cp = sincosd(phi)
    st, ct = sincosd(theta)
    ss, cs = sincosd(psi)
    return np.array([
        [cs * ct * cp - ss * sp, cs * ct * sp + ss * cp, -cs * st],
        [-ss * ct * cp - cs * sp, -ss * ct * sp + cs * cp, ss * st],
        [st * cp, st * sp, ct]])


def fft_grid_edge(dim2: int, ratio: float) -> int:
    """fillInFourierGrid.m:38-42: kMeasured is d x d with d = Dim2*ratio (even Dim2)
    or 1 + (Dim2-1)*ratio (odd), particleWindowSize(2) = Dim2 (GENFIRE_Reconstructor.m:121)."""
    e = dim2 * ratio if dim2 % 2 == 0 else 1 + (dim2 - 1) * ratio
    if e != math.floor(e):
        raise ValueError(f"grid edge {e} is not a whole number")
    return int(e)


def my_fft(img: np.ndarray) -> np.ndarray:
    """src/my_fft.m:10-12: fftshift(fftn(ifftshift(img))) (FP64)."""
    return np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(img)))


def hermitian_symmetrize(a: np.ndarray) -> np.ndarray:
    """src/hermitianSymmetrize.m:22-65, literally: zero-pad each even edge by one
    (at the end), average with the conjugate of the all-axes flip, dividing by the
    number of nonzero entries of the pair, crop back.  FP64."""
    shp = a.shape
    ext = tuple(n + 1 if n % 2 == 0 else n for n in shp)
    inp = np.zeros(ext, np.complex128)
    inp[: shp[0], : shp[1], : shp[2]] = a
    flip = inp[::-1, ::-1, ::-1]
    out = inp + np.conj(flip)
    cnt = (inp != 0).astype(np.float64)
    cnt = cnt + cnt[::-1, ::-1, ::-1]
    nz = cnt != 0
    out[nz] = out[nz] / cnt[nz]
    return out[: shp[0], : shp[1], : shp[2]]


@dataclass
class FftGridResult:
    d: int
    master_ind: np.ndarray      # 0-based linear index into the centred d^3 grid, list order
    master_dist: np.ndarray     # distances (single)
    master_val: np.ndarray      # kMeasured values (complex128)
    hits: np.ndarray            # entries of masterInd per voxel (d^3, int32, Fortran order)
    measuredK: np.ndarray       # centred d^3, complex64 (after hermitianSymmetrize and the single cast)


def grid_fft(projections: np.ndarray, angles: np.ndarray, ratio: float, cutoff: float,
             allow_multiple: bool = True) -> FftGridResult:
    """fillInFourierGrid.m:27-163 + runGridding's single cast (GENFIRE_Reconstructor.m:182).

    Point-centric and literal: the rotated lattice of every projection
    (rotkCoords = R'*[kx;ky;kz] in single, products rounded and summed left to
    right), the shift loops Y, X, Z (:125-143) appending (index, value,
    distance) in list order, weights 1/(D + 1e-5) in single, FP64 accumarray
    (:153-160), then :func:`hermitian_symmetrize`.  The spectra are FP64 FFTs of
    the single padded projections (:59-61).  Relational tests of single against
    double compare exactly."""
    dim1, dim2, P = projections.shape
    d = fft_grid_edge(dim2, ratio)
    nc = int(matlab_round((d + 1) / 2))                                  # :48
    r = (np.arange(1, d + 1) - nc).astype(np.float32)
    kx = np.tile(r, d)                                                   # :52-53 (rows fastest)
    ky = np.repeat(r, d)
    kz = np.zeros(d * d, np.float32)
    mX, mY, mZ, vals = [], [], [], []
    for p in range(P):
        spec = my_fft(paddzero(np.asarray(projections[:, :, p], np.float32), (d, d)))   # :59-61
        Rt = fft_grid_rotation(*angles[p]).T.astype(np.float32)         # R' in single (:87)
        m = [(Rt[i, 0] * kx + Rt[i, 1] * ky) + Rt[i, 2] * kz for i in range(3)]
        mX.append(m[0]); mY.append(m[1]); mZ.append(m[2])
        vals.append(spec.ravel(order="F"))
    mX, mY, mZ = np.concatenate(mX), np.concatenate(mY), np.concatenate(mZ)
    kM = np.concatenate(vals)
    good = ~(kM == -999)                                                 # :108-113
    mX, mY, mZ, kM = mX[good], mY[good], mZ[good], kM[good]
    S = int(matlab_round(cutoff)) if allow_multiple else 0               # :117-121
    rX = matlab_round(mX).astype(np.float32)
    rY = matlab_round(mY).astype(np.float32)
    rZ = matlab_round(mZ).astype(np.float32)
    inds, dists, mv = [], [], []
    for ys in range(-S, S + 1):
        for xs in range(-S, S + 1):
            for zs in range(-S, S + 1):
                tX = rX + np.float32(xs)
                tY = rY + np.float32(ys)
                tZ = rZ + np.float32(zs)
                dx, dy, dz = np.abs(mX - tX), np.abs(mY - tY), np.abs(mZ - tZ)
                dist = np.sqrt((dx * dx + dy * dy) + dz * dz)           # :132 (single)
                iX, iY, iZ = tX + nc, tY + nc, tZ + nc
                ok = ~((iX > d) | (iX < 1) | (iY > d) | (iY < 1) | (iZ > d) | (iZ < 1))
                ok &= dist.astype(np.float64) <= cutoff                 # :136
                ind = ((iX[ok] - 1).astype(np.int64) + d * ((iY[ok] - 1).astype(np.int64)
                                                            + d * (iZ[ok] - 1).astype(np.int64)))
                inds.append(ind); dists.append(dist[ok]); mv.append(kM[ok])
    master_ind = np.concatenate(inds)
    master_dist = np.concatenate(dists)
    master_val = np.concatenate(mv)
    w = (master_dist + np.float32(1e-5)).astype(np.float32)              # :153
    w = (np.float32(1.0) / w).astype(np.float32)                         # :154
    wd = w.astype(np.float64)
    tot = d ** 3
    acc = (np.bincount(master_ind, weights=master_val.real * wd, minlength=tot)
           + 1j * np.bincount(master_ind, weights=master_val.imag * wd, minlength=tot))   # :157
    sw = np.bincount(master_ind, weights=wd, minlength=tot)             # :158
    hits = np.bincount(master_ind, minlength=tot).astype(np.int32)
    K = np.zeros(tot, np.complex128)
    pos = sw > 0
    K[pos] = acc[pos] / sw[pos]                                          # :159
    K = hermitian_symmetrize(K.reshape(d, d, d, order="F"))              # :163
    return FftGridResult(d=d, master_ind=master_ind, master_dist=master_dist, master_val=master_val,
                         hits=hits.reshape(d, d, d, order="F"), measuredK=K.astype(np.complex64))


# ---------------------------------------------------------------------------
# Projector and R-factor (kernel c)
# ---------------------------------------------------------------------------

def zyz(phi: float, theta: float, psi: float) -> np.ndarray:
    R = np.zeros(9)
    lib().go_zyz(float(phi), float(theta), float(psi), R)
    return R.reshape(3, 3, order="F")


def project(vol: np.ndarray, phi: float, theta: float, psi: float, interp: str = "linear") -> np.ndarray:
    """calculate3Dprojection_RealSpaceinterp(Vol,phi,theta,psi) with extrapolation
    value 0 (FP64): trilinear interpolation (interp="linear", the library's
    default), or interp3's 'cubic' (:45,51), cubic convolution with
    Keys' a = -1/2 kernel and boundary nodes (interp="cubic", genfire_oracle.c
    go_project_cubic; MATLAB is absent, so its agreement with MATLAB itself is
    parity unpinned)."""
    d1, d2, d3 = vol.shape
    R = zyz(phi, theta, psi)
    out = np.zeros(d1 * d2, np.float64)
    fn = {"linear": lib().go_project, "cubic": lib().go_project_cubic}[interp]
    fn(np.ascontiguousarray(np.asarray(vol, np.float32).ravel(order="F")),
       d1, d2, d3, np.ascontiguousarray(R.ravel(order="F")), out)
    return out.reshape(d1, d2, order="F")


def interp_cubic(vol: np.ndarray, pts: np.ndarray) -> np.ndarray:
    """The cubic-convolution interpolant of vol at pts (k x 3, node units along
    x, y, z), 0 outside the closed node box (genfire_oracle.c go_interp_cubic)."""
    d1, d2, d3 = vol.shape
    pts = np.ascontiguousarray(np.asarray(pts, np.float64).reshape(-1, 3))
    out = np.zeros(len(pts), np.float64)
    lib().go_interp_cubic(np.ascontiguousarray(np.asarray(vol, np.float32).ravel(order="F")),
                          d1, d2, d3, pts.ravel(), len(pts), out)
    return out


def quat_rot(vec, theta_deg: float) -> np.ndarray:
    """MatrixQuaternionRot (src/MatrixQuaternionRot.m), FP64, R[i, j] = row i."""
    th = theta_deg * np.pi / 180.0
    a = np.asarray(vec, np.float64)
    a = a / np.linalg.norm(a)
    w, s = np.cos(th / 2.0), np.sin(th / 2.0)
    x, y, z = -s * a[0], -s * a[1], -s * a[2]
    return np.array([[1 - 2 * y * y - 2 * z * z, 2 * x * y + 2 * w * z, 2 * x * z - 2 * w * y],
                     [2 * x * y - 2 * w * z, 1 - 2 * x * x - 2 * z * z, 2 * y * z + 2 * w * x],
                     [2 * x * z + 2 * w * y, 2 * y * z - 2 * w * x, 1 - 2 * x * x - 2 * y * y]])


def project_general(vol: np.ndarray, phi: float, theta: float, psi: float, v1, v2, v3,
                    interp: str = "linear") -> np.ndarray:
    """src/calculate3Dprojection_RealSpaceinterp_general.m with R =
    Q(v1,phi) Q(v2,theta) Q(v3,psi) (:3-7) and the interpolant ``interp`` of
    ``project`` (FP64; the reference's is 'cubic', :37,43)."""
    d1, d2, d3 = vol.shape
    R = quat_rot(v1, phi) @ quat_rot(v2, theta) @ quat_rot(v3, psi)
    out = np.zeros(d1 * d2, np.float64)
    fn = {"linear": lib().go_project, "cubic": lib().go_project_cubic}[interp]
    fn(np.ascontiguousarray(np.asarray(vol, np.float32).ravel(order="F")),
       d1, d2, d3, np.ascontiguousarray(R.ravel(order="F")), out)
    return out.reshape(d1, d2, order="F")


def rfactors(model: np.ndarray, projections: np.ndarray, angles: np.ndarray, interp: str = "linear"):
    """R-factor monitor body, reconstruct.m:171-178: per projection
    R1 = sum|back-pj|/sum|pj| and R2 = ||back-pj||_F/||pj||_F (``project``'s
    interpolant ``interp``)."""
    P = projections.shape[2]
    R1 = np.zeros(P)
    R2 = np.zeros(P)
    for i in range(P):
        pj = np.asarray(projections[:, :, i], np.float64)
        back = project(model, *angles[i], interp=interp)
        R1[i] = np.sum(np.abs(back - pj)) / np.sum(np.abs(pj))
        R2[i] = np.linalg.norm(back - pj) / np.linalg.norm(pj)
    return R1, R2


def rfactor_posthoc(model: np.ndarray, projections: np.ndarray, angles: np.ndarray, interp: str = "linear"):
    """src/Tian_calc_Rfactor_realspace.m:1-24 with the projector defined (the
    reference calls a missing ..._RealSpaceinterpZYX, :15): mean of
    sum||back|-|pj|| / sum|pj| after cropping the projection to the pj size."""
    P = projections.shape[2]
    R = np.zeros(P)
    sims = np.zeros(projections.shape, np.float32)
    for i in range(P):
        pj = np.asarray(projections[:, :, i], np.float64)
        back = stripzero(project(model, *angles[i], interp=interp), pj.shape)
        R[i] = np.sum(np.abs(np.abs(back) - np.abs(pj))) / np.sum(np.abs(pj))
        sims[:, :, i] = back
    return float(np.mean(R)), R, sims


# ---------------------------------------------------------------------------
# Iteration (kernel b)
# ---------------------------------------------------------------------------

@dataclass
class ReconResult:
    reconstruction: np.ndarray
    errK: np.ndarray
    errR: np.ndarray | None = None
    Rarr_record: np.ndarray | None = None
    Rarr2_record: np.ndarray | None = None
    last_iterate: np.ndarray | None = None
    Rfree: np.ndarray | None = None
    snapshots: list | None = None


_U64 = np.uint64


def splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    x = np.asarray(x, _U64) + _U64(0x9E3779B97F4A7C15)
    x = (x ^ (x >> _U64(30))) * _U64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> _U64(27))) * _U64(0x94D049BB133111EB)
    return x ^ (x >> _U64(31))


def rfree_select(measuredK: np.ndarray, num_bins: int, percent: float, seed: int):
    """reconstruct.m:46-69 with the build's seeded permutation in place of
    MATLAB's randperm after rng(RandomSeed) (MATLAB's generator is not
    available): per shell s, the candidates find(measuredK ~= 0 & Q >= edge_s -
    0.05 & Q < edge_s+1) are ordered by splitmix64(splitmix64(seed) ^
    splitmix64(index)) (ties by index) and the first floor(count * percent),
    at least one, are withheld.  Returns [(centred linear indices, values)] per
    shell (empty shells give empty arrays)."""
    shape = measuredK.shape
    Q = make_kspace_indices(shape)
    edges = matlab_linspace(0.0, 1.0, num_bins + 1)                                  # :47
    nz = np.asarray(measuredK) != 0
    sd = splitmix64(np.array([seed], dtype=np.int64).astype(_U64))[0]
    Mf = np.asarray(measuredK).ravel(order="F")
    out = []
    with np.errstate(over="ignore"):
        for s in range(num_bins):
            cand = find(nz & (Q >= edges[s] + (-0.05)) & (Q < edges[s + 1]))          # :50
            if cand.size == 0:
                out.append((cand, Mf[cand]))
                continue
            keys = splitmix64(sd ^ splitmix64(cand.astype(_U64)))
            perm = cand[np.lexsort((cand, keys))]                                       # :55-57
            cut = int(np.floor(cand.size * percent))                                    # :58
            if cut == 0:
                cut = 1                                                                 # :59-61
            sel = np.sort(perm[:cut])
            out.append((sel, Mf[sel]))                                                  # :62, :67
    return out


def reconstruct(measuredK: np.ndarray, mask: np.ndarray, support: np.ndarray, dim1: int,
                dim2: int, num_iterations: int, mode: int = 3, positivity: bool = True,
                support_on: bool = True, algorithm: str = "genfire", precision: str = "fp32",
                initial_object: np.ndarray | None = None, monitor_R: bool = False,
                loop_length: int = 10, projections=None, angles=None, rfree=None,
                save_every: int = 0, projector_interp: str = "linear") -> ReconResult:
    """reconstruct.m:28-194 (GENFIRE, algorithm='genfire').

    algorithm='dr' is the library's Douglas-Rachford variant composed of the same two
    projections (P_R = positivity/support, :126-131; P_F = replacement + real(ifftn),
    :163-165): y = 2 P_R(x) - x, x <- x + P_F(y) - P_R(x); errK is evaluated on
    fftn(y), the input of P_F, exactly as GENFIRE evaluates it on fftn(P_R(x)); the
    reconstruction is P_R(x) of the best iteration.  The reference has no DR (its
    README claims one), so the DR path is parity unpinned against the reference and
    pinned only by its fixed-point and convex-feasibility properties.

    precision: 'fp32' (MATLAB single, as when the support constraint is on) or 'fp64'.
    rfree: None or (num_bins, percent, seed): withhold points per shell
    (rfree_select, :46-69) and record Rfree_complex (:137-147) in ``Rfree``.
    save_every: save_loopLength of save_temp (:155-159); ``snapshots`` then holds
    (iteration, My_stripzero(fftshift(best))) for every multiple of it.
    """
    cdt = np.complex64 if precision == "fp32" else np.complex128
    rdt = np.float32 if precision == "fp32" else np.float64
    shape = measuredK.shape
    M = np.asarray(measuredK, cdt)
    padded_support = paddzero(np.asarray(support, np.float32), shape).astype(rdt)   # :34
    Q = make_kspace_indices(shape)                                                    # :36
    measured = (M != 0) & np.asarray(mask, bool)
    indicators = np.zeros(shape)                                                      # :39
    indicators[measured] = 1 - Q[measured]                                            # :42
    rf = None
    if rfree is not None:
        rf = rfree_select(M, *rfree)
        mflat = measured.ravel(order="F").copy()
        for sel, _ in rf:
            mflat[sel] = False                                                        # :66
        measured = mflat.reshape(shape, order="F")
        rf = [(iffshift3_ind(shape, sel), vals) for sel, vals in rf]                  # :65
        Rfree = np.full((len(rf), num_iterations), -1.0, np.float32)                  # :92-94
    delays = delay_indicators(mode)
    if initial_object is None:
        x = np.zeros(shape, rdt)                                                      # :77
    else:
        x = paddzero(np.asarray(initial_object, np.float32), shape).astype(rdt)
    errK = np.zeros(num_iterations, np.float32)                                       # :95
    err_ind = find(measured)                                                          # :99
    err_sh = iffshift3_ind(shape, err_ind)                                            # :100
    Mf = M.ravel(order="F")
    M_err = Mf[err_ind]
    den = np.sum(np.abs(M_err).astype(np.float64))
    sched = cutoff_schedule(num_iterations, len(delays))                              # :103
    cur = 0
    S = np.fft.ifftshift(padded_support)                                              # :109
    x = np.fft.ifftshift(x)                                                           # :110
    nR = num_iterations // loop_length if monitor_R else 0
    errR = np.zeros(nR)
    Rrec = np.zeros((projections.shape[2] if monitor_R else 0, nR))
    R2rec = np.zeros_like(Rrec)
    vol_ind = volumn_index(shape, (dim1, dim2, dim1))
    best_err = 1e30
    recon = None
    snapshots = []
    c_sh = None
    c_val = None
    for it in range(1, num_iterations + 1):                                          # :117
        if cur < len(sched) and it == sched[cur]:                                     # :118
            cur = int(np.nonzero(sched == it)[0][-1])
            c_ind = find((indicators > delays[cur]) & measured)                       # :120
            c_sh = iffshift3_ind(shape, c_ind)                                        # :121
            c_val = Mf[c_ind]
            cur += 1
            best_err = 1e30                                                           # :123
        if algorithm == "genfire":
            if positivity:
                x = np.where(x < 0, rdt(0), x)                                        # :127
            if support_on:
                x = x * S                                                             # :130
            r = x
            y = x
        else:
            r = x
            if positivity:
                r = np.where(r < 0, rdt(0), r)
            if support_on:
                r = r * S
            y = (rdt(2) * r - x).astype(rdt)
        k = scipy.fft.fftn(y.astype(cdt), workers=-1)                                 # :133
        kf = k.reshape(-1, order="F")
        e = np.abs(np.abs(kf[err_sh]).astype(np.float64) - np.abs(M_err).astype(np.float64))
        errK[it - 1] = np.float32(np.sum(e) / den)                                    # :135
        if rf is not None:                                                            # :137-147
            for sh, (ind_sh, vals) in enumerate(rf):
                if ind_sh.size:
                    num = np.sum(np.abs(kf[ind_sh].astype(np.complex128) - vals.astype(np.complex128)))
                    Rfree[sh, it - 1] = np.float32(num / np.sum(np.abs(vals.astype(np.complex128))))
        if errK[it - 1] < best_err:                                                   # :149-153
            best_err = errK[it - 1]
            recon = r.copy()
        if save_every > 0 and it % save_every == 0:                                   # :155-159
            snapshots.append((it, np.asarray(stripzero(np.fft.fftshift(recon), (dim1, dim2, dim1)), np.float32)))
        kf = kf.copy()
        kf[c_sh] = c_val                                                              # :163
        k = kf.reshape(shape, order="F")
        pf = np.real(scipy.fft.ifftn(k, workers=-1)).astype(rdt)                      # :165
        x = pf if algorithm == "genfire" else (x + pf - r).astype(rdt)
        if monitor_R and it % loop_length == 0:                                       # :168
            model = np.fft.fftshift(pf)
            sl = tuple(slice(a - 1, b) for a, b in vol_ind)
            model = model[sl]
            R1, R2 = rfactors(model, projections, angles, interp=projector_interp)
            j = it // loop_length - 1
            errR[j] = np.mean(R1)
            Rrec[:, j] = R1
            R2rec[:, j] = R2
    reconstruction = stripzero(np.fft.fftshift(recon), (dim1, dim2, dim1))            # :192
    return ReconResult(reconstruction=np.asarray(reconstruction, np.float32), errK=errK,
                       errR=errR if monitor_R else None,
                       Rarr_record=Rrec if monitor_R else None,
                       Rarr2_record=R2rec if monitor_R else None,
                       last_iterate=x, Rfree=Rfree if rf is not None else None, snapshots=snapshots)
