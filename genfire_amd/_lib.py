"""ctypes binding of libgenfire.so (declared in include/genfire.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when genfire_amd/libgenfire.so has not been built (python -m genfire_amd.build).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgenfire.so")

GENFIRE_OK = 0
STATUS = {
    -1: "GENFIRE_ERR_ARG",
    -2: "GENFIRE_ERR_HIP",
    -3: "GENFIRE_ERR_NOMEM",
    -4: "GENFIRE_ERR_STATE",
    -5: "GENFIRE_ERR_UNSUPPORTED",
}
ALG_GENFIRE, ALG_DR = 0, 1
PREC_FP32, PREC_SPLIT_BF16, PREC_FP16 = 0, 1, 2
INTERP_LINEAR, INTERP_CUBIC = 0, 1
KSTAT_GRID_VALUES, KSTAT_FFT_PASS, KSTAT_FFT_FUSED, KSTAT_ITERATION, KSTAT_PROJECT = range(5)

# every symbol include/genfire.h declares
EXPORTS = [
    "genfire_version", "genfire_params_default", "genfire_ctx_create", "genfire_ctx_destroy",
    "genfire_last_error", "genfire_job_create", "genfire_job_destroy", "genfire_job_sizes",
    "genfire_job_set_inputs", "genfire_job_set_inputs_f16", "genfire_job_grid", "genfire_job_grid_count", "genfire_job_get_points",
    "genfire_job_get_measuredK", "genfire_job_get_hits", "genfire_job_get_recIFFT",
    "genfire_job_set_measuredK", "genfire_job_reconstruct", "genfire_job_get_reconstruction",
    "genfire_job_get_errK", "genfire_job_get_errR", "genfire_job_get_Rfree", "genfire_job_submit", "genfire_job_poll",
    "genfire_job_wait", "genfire_job_timing", "genfire_job_set_profiling",
    "genfire_job_kernel_stats", "genfire_job_graph_replays", "genfire_project", "genfire_rfactor", "genfire_fftn",
    "genfire_project_general", "genfire_rfactor_general", "genfire_project_cubic", "genfire_rfactor_cubic",
    "genfire_project_general_cubic", "genfire_rfactor_general_cubic", "genfire_rfactor_noalloc", "genfire_job_set_snapshots",
    "genfire_server_config_default", "genfire_server_create", "genfire_server_destroy",
    "genfire_server_submit", "genfire_server_status", "genfire_server_wait", "genfire_server_drain",
    "genfire_server_result", "genfire_server_release", "genfire_server_get_stats",
]
REQ_QUEUED, REQ_PREPARING, REQ_ITERATING, REQ_FINISHING, REQ_DONE, REQ_FAILED = range(6)


class GenfireParams(ctypes.Structure):
    _fields_ = [
        ("numIterations", ctypes.c_int32),
        ("oversamplingRatio", ctypes.c_double),
        ("interpolationCutoffDistance", ctypes.c_double),
        ("constraintEnforcementMode", ctypes.c_int32),
        ("constraintPositivity", ctypes.c_int32),
        ("constraintSupport", ctypes.c_int32),
        ("DFT_CentroSymmetricity", ctypes.c_int32),
        ("vector1", ctypes.c_double * 3),
        ("vector2", ctypes.c_double * 3),
        ("vector3", ctypes.c_double * 3),
        ("monitor_R", ctypes.c_int32),
        ("monitorR_loopLength", ctypes.c_int32),
        ("algorithm", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("compute_recIFFT", ctypes.c_int32),
        ("calculate_Rfree", ctypes.c_int32),
        ("numBinsRfree", ctypes.c_int32),
        ("percentValuesForRfree", ctypes.c_double),
        ("RandomSeed", ctypes.c_int64),
        ("griddingMethod", ctypes.c_int32),
        ("allowMultipleGridMatches", ctypes.c_int32),
        ("projectorInterpolation", ctypes.c_int32),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_int64),
        ("max_active", ctypes.c_int32),
        ("max_batch", ctypes.c_int32),
        ("lookahead", ctypes.c_int32),
        ("prep_threads", ctypes.c_int32),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_int64),
        ("completed", ctypes.c_int64),
        ("failed", ctypes.c_int64),
        ("queued", ctypes.c_int64),
        ("active", ctypes.c_int64),
        ("batched_launches", ctypes.c_int64),
        ("batched_jobs", ctypes.c_int64),
        ("peak_footprint_bytes", ctypes.c_int64),
        ("graph_replays", ctypes.c_int64),
        ("hbm_budget_bytes", ctypes.c_int64),
        ("max_active_seen", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
        ("p50_ms", ctypes.c_double),
        ("p99_ms", ctypes.c_double),
        ("mean_ms", ctypes.c_double),
        ("jobs_per_hour", ctypes.c_double),
    ]


# void (*)(void* user, int32_t iteration, const float* reconstruction)
SNAPSHOT_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_float))


class GenfireError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


_LIB = None


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype)) if a is not None else None


def load():
    """Load libgenfire.so once and declare its prototypes."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libgenfire.so not built: {LIB_PATH} (run python -m genfire_amd.build)")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    i32, i64, f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    P = ctypes.POINTER
    pf, pd, pi32, pi64, pu8 = P(ctypes.c_float), P(f64), P(i32), P(i64), P(ctypes.c_uint8)
    sig = {
        "genfire_version": ([], ctypes.c_int),
        "genfire_params_default": ([P(GenfireParams)], ctypes.c_int),
        "genfire_ctx_create": ([i32, P(vp)], ctypes.c_int),
        "genfire_ctx_destroy": ([vp], ctypes.c_int),
        "genfire_last_error": ([vp], ctypes.c_char_p),
        "genfire_job_create": ([vp, P(GenfireParams), i32, i32, i32, P(vp)], ctypes.c_int),
        "genfire_job_destroy": ([vp], ctypes.c_int),
        "genfire_job_sizes": ([vp, pi32, pi32, pi64], ctypes.c_int),
        "genfire_job_set_inputs": ([vp, pf, pd, pu8, pf], ctypes.c_int),
        "genfire_job_set_inputs_f16": ([vp, P(ctypes.c_uint16), pf, pd, pu8, pf], ctypes.c_int),
        "genfire_job_grid": ([vp], ctypes.c_int),
        "genfire_job_grid_count": ([vp, pi64], ctypes.c_int),
        "genfire_job_get_points": ([vp, pi64, pi32, pf, pf, pf, pf], ctypes.c_int),
        "genfire_job_get_measuredK": ([vp, pf], ctypes.c_int),
        "genfire_job_get_hits": ([vp, pi32], ctypes.c_int),
        "genfire_job_get_recIFFT": ([vp, pf], ctypes.c_int),
        "genfire_job_set_measuredK": ([vp, pf, pu8], ctypes.c_int),
        "genfire_job_reconstruct": ([vp], ctypes.c_int),
        "genfire_job_set_snapshots": ([vp, i32, SNAPSHOT_FN, vp], ctypes.c_int),
        "genfire_job_get_reconstruction": ([vp, pf], ctypes.c_int),
        "genfire_job_get_errK": ([vp, pf], ctypes.c_int),
        "genfire_job_get_errR": ([vp, pf, pd, pd], ctypes.c_int),
        "genfire_job_get_Rfree": ([vp, pf], ctypes.c_int),
        "genfire_job_submit": ([vp], ctypes.c_int),
        "genfire_job_poll": ([vp, pi32], ctypes.c_int),
        "genfire_job_wait": ([vp], ctypes.c_int),
        "genfire_job_timing": ([vp, pd], ctypes.c_int),
        "genfire_job_set_profiling": ([vp, i32], ctypes.c_int),
        "genfire_job_kernel_stats": ([vp, i32, pi64, pd], ctypes.c_int),
        "genfire_job_graph_replays": ([vp, pi64], ctypes.c_int),
        "genfire_project": ([vp, pf, i32, i32, i32, f64, f64, f64, pf], ctypes.c_int),
        "genfire_rfactor": ([vp, pf, i32, i32, i32, pf, i32, i32, i32, pd, pd, pd, pf], ctypes.c_int),
        "genfire_project_cubic": ([vp, pf, i32, i32, i32, f64, f64, f64, pf], ctypes.c_int),
        "genfire_rfactor_cubic": ([vp, pf, i32, i32, i32, pf, i32, i32, i32, pd, pd, pd, pf], ctypes.c_int),
        "genfire_fftn": ([vp, pf, i32, i32, i32, i32], ctypes.c_int),
        "genfire_project_general": ([vp, pf, i32, i32, i32, f64, f64, f64, pd, pd, pd, pf], ctypes.c_int),
        "genfire_rfactor_general": ([vp, pf, i32, i32, i32, pf, i32, i32, i32, pd, pd, pd, pd, pd, pd, pf],
                                    ctypes.c_int),
        "genfire_project_general_cubic": ([vp, pf, i32, i32, i32, f64, f64, f64, pd, pd, pd, pf], ctypes.c_int),
        "genfire_rfactor_general_cubic": ([vp, pf, i32, i32, i32, pf, i32, i32, i32, pd, pd, pd, pd, pd, pd, pf],
                                          ctypes.c_int),
        "genfire_rfactor_noalloc": ([vp, vp, i32, i32, i32, vp, i32, i32, i32, pd, pd, pd, vp], ctypes.c_int),
        "genfire_server_config_default": ([P(ServerConfig)], ctypes.c_int),
        "genfire_server_create": ([vp, P(ServerConfig), P(vp)], ctypes.c_int),
        "genfire_server_destroy": ([vp], ctypes.c_int),
        "genfire_server_submit": ([vp, P(GenfireParams), i32, i32, i32, pf, pd, pu8, pi64], ctypes.c_int),
        "genfire_server_status": ([vp, i64, pi32], ctypes.c_int),
        "genfire_server_wait": ([vp, i64], ctypes.c_int),
        "genfire_server_drain": ([vp], ctypes.c_int),
        "genfire_server_result": ([vp, i64, pf, pf, pf, pd], ctypes.c_int),
        "genfire_server_release": ([vp, i64], ctypes.c_int),
        "genfire_server_get_stats": ([vp, P(ServerStats)], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _LIB = L
    return L


def default_params() -> GenfireParams:
    p = GenfireParams()
    load().genfire_params_default(ctypes.byref(p))
    return p


class Context:
    """One GPU (genfire_ctx)."""

    def __init__(self, device: int = 0):
        self.L = load()
        h = ctypes.c_void_p()
        rc = self.L.genfire_ctx_create(int(device), ctypes.byref(h))
        if rc != GENFIRE_OK:
            raise GenfireError(rc, f"genfire_ctx_create(device={device}) failed")
        self.h = h

    def check(self, rc: int, what: str):
        if rc != GENFIRE_OK:
            msg = self.L.genfire_last_error(self.h)
            raise GenfireError(rc, f"{what}: {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            self.L.genfire_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- standalone entry points ----
    def fftn(self, a: np.ndarray, inverse: bool = False) -> np.ndarray:
        """MATLAB fftn / ifftn of a 3D complex array (Fortran-order semantics)."""
        a = np.asarray(a, np.complex64)
        n = list(a.shape) + [1] * (3 - a.ndim)
        buf = np.asfortranarray(a).ravel(order="F").view(np.float32).copy()
        self.check(self.L.genfire_fftn(self.h, _ptr(buf, ctypes.c_float), n[0], n[1], n[2],
                                       1 if inverse else 0), "genfire_fftn")
        return buf.view(np.complex64).reshape(a.shape, order="F")

    def project(self, vol: np.ndarray, phi: float, theta: float, psi: float, interp: str = "linear") -> np.ndarray:
        """calculate3Dprojection_RealSpaceinterp: trilinear (genfire_project) or
        interp3 'cubic' (genfire_project_cubic)."""
        d1, d2, d3 = vol.shape
        v = np.ascontiguousarray(np.asarray(vol, np.float32).ravel(order="F"))
        out = np.zeros(d1 * d2, np.float32)
        name = {"linear": "genfire_project", "cubic": "genfire_project_cubic"}[interp]
        self.check(getattr(self.L, name)(self.h, _ptr(v, ctypes.c_float), d1, d2, d3, float(phi),
                                         float(theta), float(psi), _ptr(out, ctypes.c_float)), name)
        return out.reshape(d1, d2, order="F")

    def project_general(self, vol: np.ndarray, phi: float, theta: float, psi: float, v1, v2, v3,
                        interp: str = "linear") -> np.ndarray:
        """calculate3Dprojection_RealSpaceinterp_general: trilinear, or interp3 'cubic'."""
        d1, d2, d3 = vol.shape
        v = np.ascontiguousarray(np.asarray(vol, np.float32).ravel(order="F"))
        vec = [np.ascontiguousarray(np.asarray(x, np.float64).reshape(3)) for x in (v1, v2, v3)]
        out = np.zeros(d1 * d2, np.float32)
        name = {"linear": "genfire_project_general", "cubic": "genfire_project_general_cubic"}[interp]
        self.check(getattr(self.L, name)(self.h, _ptr(v, ctypes.c_float), d1, d2, d3, float(phi),
                                         float(theta), float(psi), *[_ptr(x, ctypes.c_double) for x in vec],
                                         _ptr(out, ctypes.c_float)), name)
        return out.reshape(d1, d2, order="F")

    def rfactor_general(self, model: np.ndarray, projections: np.ndarray, angles: np.ndarray, v1, v2, v3,
                        interp: str = "linear"):
        """Tian_calc_Rfactor_realspace with the general-axis projector (trilinear or cubic)."""
        d1, d2, d3 = model.shape
        p1, p2, P = projections.shape
        v = np.ascontiguousarray(np.asarray(model, np.float32).ravel(order="F"))
        m = np.ascontiguousarray(np.asarray(projections, np.float32).ravel(order="F"))
        ang = np.ascontiguousarray(np.asarray(angles, np.float64).ravel(order="F"))
        vec = [np.ascontiguousarray(np.asarray(x, np.float64).reshape(3)) for x in (v1, v2, v3)]
        rt = np.zeros(1)
        ra = np.zeros(P)
        sim = np.zeros(p1 * p2 * P, np.float32)
        name = {"linear": "genfire_rfactor_general", "cubic": "genfire_rfactor_general_cubic"}[interp]
        self.check(getattr(self.L, name)(self.h, _ptr(v, ctypes.c_float), d1, d2, d3,
                                         _ptr(m, ctypes.c_float), p1, p2, P, _ptr(ang, ctypes.c_double),
                                         *[_ptr(x, ctypes.c_double) for x in vec],
                                         _ptr(rt, ctypes.c_double), _ptr(ra, ctypes.c_double),
                                         _ptr(sim, ctypes.c_float)), name)
        return float(rt[0]), ra, sim.reshape(p1, p2, P, order="F")

    def rfactor(self, model: np.ndarray, projections: np.ndarray, angles: np.ndarray, interp: str = "linear"):
        """Tian_calc_Rfactor_realspace(model, Projs, Angles) -> (Rtotal, Rarr, simuProjs),
        with the trilinear (genfire_rfactor) or cubic (genfire_rfactor_cubic) projector."""
        d1, d2, d3 = model.shape
        p1, p2, P = projections.shape
        v = np.ascontiguousarray(np.asarray(model, np.float32).ravel(order="F"))
        m = np.ascontiguousarray(np.asarray(projections, np.float32).ravel(order="F"))
        ang = np.ascontiguousarray(np.asarray(angles, np.float64).ravel(order="F"))
        rt = np.zeros(1)
        ra = np.zeros(P)
        sim = np.zeros(p1 * p2 * P, np.float32)
        name = {"linear": "genfire_rfactor", "cubic": "genfire_rfactor_cubic"}[interp]
        self.check(getattr(self.L, name)(self.h, _ptr(v, ctypes.c_float), d1, d2, d3,
                                         _ptr(m, ctypes.c_float), p1, p2, P,
                                         _ptr(ang, ctypes.c_double), _ptr(rt, ctypes.c_double),
                                         _ptr(ra, ctypes.c_double), _ptr(sim, ctypes.c_float)), name)
        return float(rt[0]), ra, sim.reshape(p1, p2, P, order="F")


    def rfactor_noalloc(self, d_model: int, model_shape, d_projections: int, proj_shape, angles: np.ndarray,
                        d_simu: int = 0):
        """genfire_rfactor_noalloc on device buffers (integer device addresses,
        column-major model and projections, d_simu 0 for none) -> (Rtotal, Rarr)."""
        d1, d2, d3 = (int(v) for v in model_shape)
        p1, p2, P = (int(v) for v in proj_shape)
        ang = np.ascontiguousarray(np.asarray(angles, np.float64).reshape(P, 3).ravel(order="F"))
        rt = np.zeros(1)
        ra = np.zeros(P)
        self.check(self.L.genfire_rfactor_noalloc(self.h, ctypes.c_void_p(int(d_model)), d1, d2, d3,
                                                  ctypes.c_void_p(int(d_projections)), p1, p2, P,
                                                  _ptr(ang, ctypes.c_double), _ptr(rt, ctypes.c_double),
                                                  _ptr(ra, ctypes.c_double),
                                                  ctypes.c_void_p(int(d_simu)) if d_simu else None),
                   "genfire_rfactor_noalloc")
        return float(rt[0]), ra


class Job:
    """One reconstruction's device-resident state (genfire_job)."""

    def __init__(self, ctx: Context, params: GenfireParams, dim1: int, dim2: int, num_projs: int):
        self.ctx = ctx
        self.L = ctx.L
        self.params = params
        h = ctypes.c_void_p()
        ctx.check(self.L.genfire_job_create(ctx.h, ctypes.byref(params), int(dim1), int(dim2),
                                            int(num_projs), ctypes.byref(h)), "genfire_job_create")
        self.h = h
        self.dim1, self.dim2, self.P = int(dim1), int(dim2), int(num_projs)
        n1 = ctypes.c_int32()
        n2 = ctypes.c_int32()
        hl = (ctypes.c_int64 * 3)()
        self.L.genfire_job_sizes(h, ctypes.byref(n1), ctypes.byref(n2), hl)
        self.n1, self.n2 = n1.value, n2.value
        self.half_shape = tuple(int(v) for v in hl)

    def close(self):
        if self.h:
            self.L.genfire_job_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc, what):
        self.ctx.check(rc, what)

    def set_inputs(self, projections, angles, support=None, initial_object=None):
        pj = np.ascontiguousarray(np.asarray(projections, np.float32).ravel(order="F"))
        ang = np.ascontiguousarray(np.asarray(angles, np.float64).reshape(self.P, 3).ravel(order="F"))
        sup = None if support is None else np.ascontiguousarray(
            (np.asarray(support) != 0).astype(np.uint8).ravel(order="F"))
        ini = None if initial_object is None else np.ascontiguousarray(
            np.asarray(initial_object, np.float32).ravel(order="F"))
        self._keep = (pj, ang, sup, ini)
        self._c(self.L.genfire_job_set_inputs(self.h, _ptr(pj, ctypes.c_float), _ptr(ang, ctypes.c_double),
                                              _ptr(sup, ctypes.c_uint8), _ptr(ini, ctypes.c_float)),
                "genfire_job_set_inputs")

    def set_inputs_f16(self, projections_f16, scales, angles, support=None, initial_object=None):
        """Projections as float16 (Dim1 x Dim2 x P) with one float scale per projection."""
        pj = np.ascontiguousarray(np.asarray(projections_f16, np.float16).ravel(order="F")).view(np.uint16)
        sc = np.ascontiguousarray(np.asarray(scales, np.float32).reshape(self.P))
        ang = np.ascontiguousarray(np.asarray(angles, np.float64).reshape(self.P, 3).ravel(order="F"))
        sup = None if support is None else np.ascontiguousarray(
            (np.asarray(support) != 0).astype(np.uint8).ravel(order="F"))
        ini = None if initial_object is None else np.ascontiguousarray(
            np.asarray(initial_object, np.float32).ravel(order="F"))
        self._keep = (pj, sc, ang, sup, ini)
        self._c(self.L.genfire_job_set_inputs_f16(self.h, _ptr(pj, ctypes.c_uint16), _ptr(sc, ctypes.c_float),
                                                  _ptr(ang, ctypes.c_double), _ptr(sup, ctypes.c_uint8),
                                                  _ptr(ini, ctypes.c_float)), "genfire_job_set_inputs_f16")

    def grid(self):
        self._c(self.L.genfire_job_grid(self.h), "genfire_job_grid")

    def num_points(self) -> int:
        n = ctypes.c_int64()
        self._c(self.L.genfire_job_grid_count(self.h, ctypes.byref(n)), "genfire_job_grid_count")
        return n.value

    def points(self):
        n = self.num_points()
        ind = np.zeros(n, np.int64)
        proj = np.zeros(n, np.int32)
        dist = np.zeros(n, np.float32)
        un = np.zeros(n, np.float32)
        vn = np.zeros(n, np.float32)
        val = np.zeros(2 * n, np.float32)
        self._c(self.L.genfire_job_get_points(self.h, _ptr(ind, ctypes.c_int64), _ptr(proj, ctypes.c_int32),
                                              _ptr(dist, ctypes.c_float), _ptr(un, ctypes.c_float),
                                              _ptr(vn, ctypes.c_float), _ptr(val, ctypes.c_float)),
                "genfire_job_get_points")
        return dict(ind=ind, proj=proj, dist=dist, un=un, vn=vn, val=val.view(np.complex64))

    def measuredK(self) -> np.ndarray:
        out = np.zeros(2 * self.n1 * self.n2 * self.n1, np.float32)
        self._c(self.L.genfire_job_get_measuredK(self.h, _ptr(out, ctypes.c_float)), "genfire_job_get_measuredK")
        return out.view(np.complex64).reshape(self.n1, self.n2, self.n1, order="F")

    def hits(self) -> np.ndarray:
        """Per-voxel hit counts: flat over the half grid (DFT gridding), or the
        centred n1 x n1 x n1 grid before hermitianSymmetrize (FFT gridding)."""
        if self.params.griddingMethod == 1:
            out = np.zeros(self.n1 ** 3, np.int32)
        else:
            out = np.zeros(int(np.prod(self.half_shape)), np.int32)
        self._c(self.L.genfire_job_get_hits(self.h, _ptr(out, ctypes.c_int32)), "genfire_job_get_hits")
        if self.params.griddingMethod == 1:
            return out.reshape(self.n1, self.n1, self.n1, order="F")
        return out

    def recIFFT(self) -> np.ndarray:
        out = np.zeros(self.dim1 * self.dim2 * self.dim1, np.float32)
        self._c(self.L.genfire_job_get_recIFFT(self.h, _ptr(out, ctypes.c_float)), "genfire_job_get_recIFFT")
        return out.reshape(self.dim1, self.dim2, self.dim1, order="F")

    def set_measuredK(self, measuredK: np.ndarray, mask=None):
        mk = np.ascontiguousarray(np.asarray(measuredK, np.complex64).ravel(order="F")).view(np.float32)
        m = None if mask is None else np.ascontiguousarray((np.asarray(mask) != 0).astype(np.uint8).ravel(order="F"))
        self._c(self.L.genfire_job_set_measuredK(self.h, _ptr(mk, ctypes.c_float), _ptr(m, ctypes.c_uint8)),
                "genfire_job_set_measuredK")

    def reconstruct(self):
        self._c(self.L.genfire_job_reconstruct(self.h), "genfire_job_reconstruct")

    def set_snapshots(self, save_loopLength: int, fn=None):
        """save_temp (reconstruct.m:155-159): fn(iteration, reconstruction) every
        save_loopLength iterations of reconstruct(), reconstruction a
        Dim1 x Dim2 x Dim1 copy; fn None (or a length <= 0) turns them off."""
        if fn is None or save_loopLength <= 0:
            self._snap_cb = None
            self._c(self.L.genfire_job_set_snapshots(self.h, 0, SNAPSHOT_FN(), None), "genfire_job_set_snapshots")
            return
        shape, n = (self.dim1, self.dim2, self.dim1), self.dim1 * self.dim2 * self.dim1

        def cb(_user, it, ptr):
            vol = np.ctypeslib.as_array(ptr, shape=(n,)).copy().reshape(shape, order="F")
            fn(int(it), vol)

        self._snap_cb = SNAPSHOT_FN(cb)   # kept alive while the job may call it
        self._c(self.L.genfire_job_set_snapshots(self.h, int(save_loopLength), self._snap_cb, None),
                "genfire_job_set_snapshots")

    def reconstruction(self) -> np.ndarray:
        out = np.zeros(self.dim1 * self.dim2 * self.dim1, np.float32)
        self._c(self.L.genfire_job_get_reconstruction(self.h, _ptr(out, ctypes.c_float)),
                "genfire_job_get_reconstruction")
        return out.reshape(self.dim1, self.dim2, self.dim1, order="F")

    def errK(self) -> np.ndarray:
        out = np.zeros(self.params.numIterations, np.float32)
        self._c(self.L.genfire_job_get_errK(self.h, _ptr(out, ctypes.c_float)), "genfire_job_get_errK")
        return out

    def errR(self):
        nR = self.params.numIterations // max(1, self.params.monitorR_loopLength) if self.params.monitor_R else 0
        e = np.zeros(nR, np.float32)
        r1 = np.zeros(nR * self.P)
        r2 = np.zeros(nR * self.P)
        self._c(self.L.genfire_job_get_errR(self.h, _ptr(e, ctypes.c_float), _ptr(r1, ctypes.c_double),
                                            _ptr(r2, ctypes.c_double)), "genfire_job_get_errR")
        return e, r1.reshape(self.P, nR, order="F"), r2.reshape(self.P, nR, order="F")

    def Rfree(self) -> np.ndarray:
        """Rfree_complex: numBinsRfree x numIterations (reconstruct.m:137-147)."""
        nb, N = self.params.numBinsRfree, self.params.numIterations
        out = np.zeros(nb * N, np.float32)
        self._c(self.L.genfire_job_get_Rfree(self.h, _ptr(out, ctypes.c_float)), "genfire_job_get_Rfree")
        return out.reshape(nb, N, order="F")

    def submit(self):
        self._c(self.L.genfire_job_submit(self.h), "genfire_job_submit")

    def poll(self) -> bool:
        d = ctypes.c_int32()
        self._c(self.L.genfire_job_poll(self.h, ctypes.byref(d)), "genfire_job_poll")
        return bool(d.value)

    def wait(self):
        self._c(self.L.genfire_job_wait(self.h), "genfire_job_wait")

    def timing(self) -> dict:
        ms = np.zeros(4)
        self._c(self.L.genfire_job_timing(self.h, _ptr(ms, ctypes.c_double)), "genfire_job_timing")
        return dict(grid_ms=ms[0], iterate_ms=ms[1], rfactor_ms=ms[2], total_ms=ms[3])

    def set_profiling(self, on: bool):
        self._c(self.L.genfire_job_set_profiling(self.h, 1 if on else 0), "genfire_job_set_profiling")

    def graph_replays(self) -> int:
        n = ctypes.c_int64()
        self._c(self.L.genfire_job_graph_replays(self.h, ctypes.byref(n)), "genfire_job_graph_replays")
        return n.value

    def kernel_stats(self, k: int):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        self._c(self.L.genfire_job_kernel_stats(self.h, int(k), ctypes.byref(n), ctypes.byref(ms)),
                "genfire_job_kernel_stats")
        return n.value, ms.value


def default_server_config() -> ServerConfig:
    c = ServerConfig()
    load().genfire_server_config_default(ctypes.byref(c))
    return c


class Server:
    """Request server (genfire_server): continuous batching of reconstruction jobs."""

    def __init__(self, ctx: Context, config: ServerConfig | None = None):
        self.ctx = ctx
        self.L = ctx.L
        h = ctypes.c_void_p()
        ctx.check(self.L.genfire_server_create(ctx.h, ctypes.byref(config) if config is not None else None,
                                               ctypes.byref(h)), "genfire_server_create")
        self.h = h
        self._shape = {}

    def close(self):
        if self.h:
            self.L.genfire_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def request_inputs(projections, angles, support=None):
        """The inputs of a request in the library's memory layout (column-major
        projections and support, P x 3 angles), for repeated submission."""
        d1, d2, P = projections.shape
        pj = np.ascontiguousarray(np.asarray(projections, np.float32).ravel(order="F"))
        ang = np.ascontiguousarray(np.asarray(angles, np.float64).reshape(P, 3).ravel(order="F"))
        sup = None if support is None else np.ascontiguousarray(
            (np.asarray(support) != 0).astype(np.uint8).ravel(order="F"))
        return (d1, d2, P), pj, ang, sup

    def submit(self, params: GenfireParams, projections, angles=None, support=None) -> int:
        """Queue a request: (projections, angles, support) arrays, or the tuple
        request_inputs returns as the second argument."""
        if isinstance(projections, tuple):
            (d1, d2, P), pj, ang, sup = projections
        else:
            (d1, d2, P), pj, ang, sup = self.request_inputs(projections, angles, support)
        rid = ctypes.c_int64()
        self.ctx.check(self.L.genfire_server_submit(self.h, ctypes.byref(params), d1, d2, P,
                                                    _ptr(pj, ctypes.c_float), _ptr(ang, ctypes.c_double),
                                                    _ptr(sup, ctypes.c_uint8), ctypes.byref(rid)),
                       "genfire_server_submit")
        nR = params.numIterations // max(1, params.monitorR_loopLength) if params.monitor_R else 0
        self._shape[rid.value] = (d1, d2, params.numIterations, nR)
        return rid.value

    def status(self, rid: int) -> int:
        st = ctypes.c_int32()
        self.ctx.check(self.L.genfire_server_status(self.h, int(rid), ctypes.byref(st)), "genfire_server_status")
        return st.value

    def wait(self, rid: int):
        self.ctx.check(self.L.genfire_server_wait(self.h, int(rid)), "genfire_server_wait")

    def drain(self):
        self.ctx.check(self.L.genfire_server_drain(self.h), "genfire_server_drain")

    def result(self, rid: int):
        """(reconstruction, errK, errR, latency_ms {queued, preparing, running, total})"""
        d1, d2, N, nR = self._shape[rid]
        rec = np.zeros(d1 * d2 * d1, np.float32)
        ek = np.zeros(N, np.float32)
        er = np.zeros(nR, np.float32)
        lat = np.zeros(4)
        self.ctx.check(self.L.genfire_server_result(self.h, int(rid), _ptr(rec, ctypes.c_float),
                                                    _ptr(ek, ctypes.c_float), _ptr(er, ctypes.c_float),
                                                    _ptr(lat, ctypes.c_double)), "genfire_server_result")
        return (rec.reshape(d1, d2, d1, order="F"), ek, er,
                dict(queued=lat[0], preparing=lat[1], running=lat[2], total=lat[3]))

    def release(self, rid: int):
        self.ctx.check(self.L.genfire_server_release(self.h, int(rid)), "genfire_server_release")
        self._shape.pop(rid, None)

    def stats(self) -> dict:
        st = ServerStats()
        self.ctx.check(self.L.genfire_server_get_stats(self.h, ctypes.byref(st)), "genfire_server_get_stats")
        return {k: getattr(st, k) for k, _ in ServerStats._fields_ if k != "pad_"}
