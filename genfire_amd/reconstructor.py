"""GENFIRE_Reconstructor: the host-side mirror of the reference's job object.

Same property names, defaults and call order as
``@GENFIRE_Reconstructor/GENFIRE_Reconstructor.m`` (properties :5-71, methods
:76-225) and the driver ``Genfire_test.m:41-44``::

    G = GENFIRE_Reconstructor()
    G.InputProjections = projs          # Dim1 x Dim2 x P
    G.InputAngles = angles              # P x 3, degrees
    G.griddingMethod = 2
    G = G.CheckPrepareData()
    G = G.runGridding()                 # fillInFourierGrid_DFT (2) / fillInFourierGrid (1) on the GPU
    G = G.reconstruct()                 # reconstruct.m on the GPU
    R, Rarr, simu = Tian_calc_Rfactor_realspace(G.reconstruction, projs, angles)

Every array-heavy step runs in libgenfire.so (include/genfire.h); the grid, the
support and the iterate stay on the device between runGridding and reconstruct.
Errors raise GenfireError with the reference's 'GENFIRE: ...' messages.
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib
from ._lib import GenfireError

_DEFAULT_CTX = None


def default_context(device: int = 0) -> _lib.Context:
    global _DEFAULT_CTX
    if _DEFAULT_CTX is None:
        _DEFAULT_CTX = _lib.Context(device)
    return _DEFAULT_CTX


def _matlab_round(x: float) -> int:
    return int(np.sign(x) * np.floor(abs(x) + 0.5))


def My_volumn_index(big_size, ori_size) -> np.ndarray:
    """src/My_volumn_index.m: 1-based [start end] of the centred box, per dimension."""
    out = np.zeros((len(big_size), 2), np.int64)
    for i, (B, s) in enumerate(zip(big_size, ori_size)):
        start = B // 2 - s // 2
        out[i] = (start + 1, start + s)
    return out


class GENFIRE_Reconstructor:
    """Value-object mirror of GENFIRE_Reconstructor.m; methods return self."""

    def __init__(self, ctx: _lib.Context | None = None):
        self._ctx = ctx
        # internal variables (:7-13)
        self.InputProjections = None
        self.InputAngles = None
        self._measuredK = None
        self._measuredK_dirty = False
        self.measuredK_mask = None
        self.Support = None
        self.initialObject = None
        self._recIFFT = None
        # kept after run (:16-28)
        self.reconstruction = None
        self.errK = None
        self.errR = None
        self.Dim1 = None
        self.Dim2 = None
        self.n1_oversampled = None
        self.n2_oversampled = None
        self.NumProjs = None
        self.Rfree_complex = None
        self.constraintEnforcementDelayIndicators = None
        self.Rarr_record = None
        self.Rarr2_record = None
        self.simuProjs = None
        # R factor monitoring (:31-32)
        self.monitor_R = 0
        self.monitorR_loopLength = 10
        # filenames (:35-39)
        self.filename_Projections = ""
        self.filename_Angles = ""
        self.filename_Support = ""
        self.filename_InitialModel = ""
        self.filename_Results = "./results/GENFIRE_rec.mat"
        # reconstruction parameters (:42-47)
        self.numIterations = 50
        self.oversamplingRatio = 3
        self.interpolationCutoffDistance = 0.7
        self.griddingMethod = 1
        self.allowMultipleGridMatches = 1
        # the R-factor monitor's interpolant: "linear" (trilinear) or "cubic"
        # (interp3 'cubic', src/calculate3Dprojection_RealSpaceinterp.m:45,51)
        self.projectorInterpolation = "linear"
        self.calculate_Rfree = 0
        # Rfree properties reconstruct.m:47-58 reads (undeclared in the reference class)
        self.numBinsRfree = 10
        self.percentValuesForRfree = 0.05
        self.RandomSeed = None
        # axes vectors (:50-52)
        self.vector1 = [0, 0, 1]
        self.vector2 = [0, 1, 0]
        self.vector3 = [1, 0, 0]
        self.particleWindowSize = None
        # constraints (:58-60)
        self.constraintEnforcementMode = 1
        self.constraintPositivity = 1
        self.constraintSupport = 1
        # DFT parameters (:63-65)
        self.DFT_CentroSymmetricity = 1
        self.DFT_doGPU = 1
        self.Vol_ind = np.zeros((3, 2))
        # save temp (:68-70)
        self.save_temp = 0
        self.save_loopLength = 50
        self.saveFilename = "./results/temp_rec"
        # additions beyond the reference
        self.algorithm = "genfire"      # 'genfire' | 'dr'
        self.precision = "fp32"         # gridding GEMM precision
        self._job = None
        self.timing = {}

    # ------------------------------------------------------------------
    @property
    def ctx(self) -> _lib.Context:
        if self._ctx is None:
            self._ctx = default_context()
        return self._ctx

    @property
    def measuredK(self):
        if self._measuredK is None and self._job is not None and not self._measuredK_dirty:
            self._measuredK = self._job.measuredK()
        return self._measuredK

    @measuredK.setter
    def measuredK(self, value):
        self._measuredK = None if value is None else np.asarray(value, np.complex64)
        self._measuredK_dirty = value is not None

    @property
    def recIFFT(self):
        if self._recIFFT is None and self._job is not None:
            self._recIFFT = self._job.recIFFT()
        return self._recIFFT

    # ------------------------------------------------------------------
    def readFiles(self):
        """GENFIRE_Reconstructor.m:82-112 (MAT v5 via scipy.io, or .npy)."""
        def load(path):
            if path.endswith(".npy"):
                return np.load(path)
            import scipy.io
            d = {k: v for k, v in scipy.io.loadmat(path).items() if not k.startswith("__")}
            if len(d) != 1:
                raise GenfireError(-1, f"GENFIRE: {path} must hold exactly one variable")
            return next(iter(d.values()))

        if os.path.isfile(self.filename_Projections):
            self.InputProjections = load(self.filename_Projections)
        else:
            raise GenfireError(-1, "GENFIRE: Projections file does not exist!")
        if os.path.isfile(self.filename_Angles):
            self.InputAngles = load(self.filename_Angles)
        else:
            raise GenfireError(-1, "GENFIRE: Angles file does not exist!")
        if self.filename_Support:
            if os.path.isfile(self.filename_Support):
                self.Support = load(self.filename_Support) != 0
            else:
                print("GENFIRE: Support file does not exist!Making a support based on Projections.")
                s = self.InputProjections.shape
                self.Support = np.ones((s[0], s[1], s[0]), bool)
        if self.filename_InitialModel:
            if os.path.isfile(self.filename_InitialModel):
                self.initialObject = load(self.filename_InitialModel)
            else:
                raise GenfireError(-1, "GENFIRE: Inital Model file does not exist!")
        return self

    def CheckPrepareData(self):
        """GENFIRE_Reconstructor.m:115-171."""
        P = np.asarray(self.InputProjections)
        if P.ndim == 2:
            P = P[:, :, None]
        self.InputProjections = P
        self.NumProjs = P.shape[2]
        self.Dim1, self.Dim2 = P.shape[0], P.shape[1]
        self.particleWindowSize = [self.Dim1, self.Dim2]
        A = np.asarray(self.InputAngles, np.float64)
        if A.ndim == 1:
            A = A[:, None]
        if self.NumProjs != A.shape[0]:
            raise GenfireError(-1, "GENFIRE: Number of projections and Angles does not match!")
        if A.shape[1] > 3:
            raise GenfireError(-1, "GENFIRE: Input Angle 2nd dimenstions larger than three!")
        if A.shape[1] == 1:
            # :135 stacks [zeros; A; zeros] vertically (a reference bug that breaks
            # later indexing); the intended P x 3 [0 theta 0] layout is built here.
            A = np.hstack([np.zeros_like(A), A, np.zeros_like(A)])
        self.InputAngles = A
        if self.initialObject is not None:
            if tuple(np.shape(self.initialObject)) != (self.Dim1, self.Dim2, self.Dim1):
                raise GenfireError(-1, "GENFIRE: Initial Model size does not match with projections!")
        if self.Support is None:
            self.Support = np.ones((self.Dim1, self.Dim2, self.Dim1), bool)
        if self.griddingMethod > 2:
            raise GenfireError(-1, "GENFIRE: Unrecognized gridding method.")
        mode = self.constraintEnforcementMode
        if mode not in (1, 2, 3):
            raise GenfireError(-1, f"GENFIRE: ERROR! constraintEnforcementMode value {mode} not understood")
        self.n1_oversampled = _matlab_round(self.Dim1 * self.oversamplingRatio)
        self.n2_oversampled = _matlab_round(self.Dim2 * self.oversamplingRatio)
        if self.griddingMethod == 1:
            # fillInFourierGrid.m:38-42 sizes kMeasured from Dim2; zeros() needs a whole number
            e = (self.Dim2 * self.oversamplingRatio if self.Dim2 % 2 == 0
                 else 1 + (self.Dim2 - 1) * self.oversamplingRatio)
            if e != int(e):
                raise GenfireError(-1, "GENFIRE: FFT gridding needs a whole-number grid edge "
                                       f"(Dim2*oversamplingRatio for even Dim2, 1+(Dim2-1)*oversamplingRatio "
                                       f"for odd), got {e}")
            if e < self.Dim1:
                raise GenfireError(-1, "GENFIRE: paddedsize should be equal to or larger than original "
                                       "volume in all dimensions!")
        return self

    def _params(self) -> _lib.GenfireParams:
        p = _lib.default_params()
        p.numIterations = int(self.numIterations)
        p.oversamplingRatio = float(self.oversamplingRatio)
        p.interpolationCutoffDistance = float(self.interpolationCutoffDistance)
        p.constraintEnforcementMode = int(self.constraintEnforcementMode)
        p.constraintPositivity = int(bool(self.constraintPositivity))
        p.constraintSupport = int(bool(self.constraintSupport))
        p.DFT_CentroSymmetricity = int(bool(self.DFT_CentroSymmetricity))
        for i in range(3):
            p.vector1[i] = float(self.vector1[i])
            p.vector2[i] = float(self.vector2[i])
            p.vector3[i] = float(self.vector3[i])
        p.monitor_R = int(bool(self.monitor_R))
        p.monitorR_loopLength = int(self.monitorR_loopLength)
        p.algorithm = {"genfire": _lib.ALG_GENFIRE, "dr": _lib.ALG_DR}[self.algorithm]
        p.precision = {"fp32": _lib.PREC_FP32, "split_bf16": _lib.PREC_SPLIT_BF16,
                       "fp16": _lib.PREC_FP16}[self.precision]
        p.compute_recIFFT = 1
        p.calculate_Rfree = int(bool(self.calculate_Rfree))
        p.numBinsRfree = int(self.numBinsRfree)
        p.percentValuesForRfree = float(self.percentValuesForRfree)
        p.RandomSeed = int(self.RandomSeed) if self.RandomSeed is not None else 0
        p.griddingMethod = int(self.griddingMethod)
        p.allowMultipleGridMatches = int(bool(self.allowMultipleGridMatches))
        p.projectorInterpolation = {"linear": _lib.INTERP_LINEAR, "cubic": _lib.INTERP_CUBIC}[self.projectorInterpolation]
        return p

    def _ensure_job(self):
        if self._job is None:
            self._job = _lib.Job(self.ctx, self._params(), self.Dim1, self.Dim2, self.NumProjs)
            self._job.set_inputs(self.InputProjections, self.InputAngles, self.Support, self.initialObject)
        return self._job

    def runGridding(self):
        """GENFIRE_Reconstructor.m:172-186."""
        if self.griddingMethod == 1:
            return self.fillInFourierGrid()
        return self.fillInFourierGrid_DFT()

    def fillInFourierGrid(self):
        """fillInFourierGrid.m:27-168 (FFT gridding + hermitianSymmetrize,
        runGridding :182-185) on the device; the grid is a cube of edge
        Dim2*oversamplingRatio (odd Dim2: 1 + (Dim2-1)*oversamplingRatio)."""
        return self._grid(1)

    def fillInFourierGrid_DFT(self):
        """fillInFourierGrid_DFT.m:85-310 (+ CentroSymmetricity, runGridding :182-185)."""
        return self._grid(2)

    def _grid(self, method: int):
        self.griddingMethod = method
        if self._job is not None:
            self._job.close()
            self._job = None
        job = self._ensure_job()
        job.grid()
        self._measuredK = None
        self._measuredK_dirty = False
        self._recIFFT = None
        self.measuredK_mask = None
        self.Vol_ind = My_volumn_index((job.n1, job.n2, job.n1), (self.Dim1, self.Dim2, self.Dim1))
        self.timing["grid_ms"] = job.timing()["grid_ms"]
        return self

    def CentroSymmetricity(self, *args):
        """CentroSymmetricity.m is fused into the gridding finalize kernel."""
        return self

    def reset_measuredK_mask(self):
        """GENFIRE_Reconstructor.m:188-192."""
        self.measuredK_mask = None
        return self

    def reconstruct(self):
        """reconstruct.m:28-194 on the device-resident grid."""
        job = self._ensure_job()
        if self._measuredK_dirty or self.measuredK_mask is not None:
            mk = self._measuredK if self._measuredK is not None else job.measuredK()
            job.set_measuredK(mk, self.measuredK_mask)
            self._measuredK_dirty = False
        if self.save_temp == 1:
            # reconstruct.m:155-159: temp_reconstruction = My_stripzero(fftshift(best)) saved
            # to [saveFilename num2str(iterationNum) '.mat'] every save_loopLength iterations
            import scipy.io
            d = os.path.dirname(self.saveFilename)
            if d:
                os.makedirs(d, exist_ok=True)

            def save(it, vol):
                scipy.io.savemat(f"{self.saveFilename}{it}.mat", {"temp_reconstruction": vol})
            job.set_snapshots(int(self.save_loopLength), save)
        else:
            job.set_snapshots(0)
        job.reconstruct()
        self.reconstruction = job.reconstruction()
        self.errK = job.errK()
        if self.monitor_R:
            self.errR, self.Rarr_record, self.Rarr2_record = job.errR()
        if self.calculate_Rfree:
            self.Rfree_complex = job.Rfree()
        self.timing.update(job.timing())
        return self

    def ClearCalcVariables(self):
        """GENFIRE_Reconstructor.m:195-203."""
        self.InputProjections = None
        self.InputAngles = None
        self.initialObject = None
        self._recIFFT = None
        return self

    def SaveResults(self):
        """GENFIRE_Reconstructor.m:205-207: save(filename_Results, 'obj') -- here a
        MAT v5 file (scipy.io) holding a struct 'obj' with the result fields and
        the parameters, readable by MATLAB's load."""
        import scipy.io
        os.makedirs(os.path.dirname(self.filename_Results) or ".", exist_ok=True)
        fields = {
            "reconstruction": self.reconstruction, "errK": self.errK, "errR": self.errR,
            "Rarr_record": self.Rarr_record, "Rarr2_record": self.Rarr2_record,
            "Rfree_complex": self.Rfree_complex, "InputAngles": self.InputAngles,
            "numIterations": self.numIterations, "oversamplingRatio": self.oversamplingRatio,
            "interpolationCutoffDistance": self.interpolationCutoffDistance,
            "constraintEnforcementMode": self.constraintEnforcementMode,
            "griddingMethod": self.griddingMethod, "calculate_Rfree": self.calculate_Rfree,
        }
        obj = {k: (np.zeros((0, 0)) if v is None else np.asarray(v)) for k, v in fields.items()}
        scipy.io.savemat(self.filename_Results, {"obj": obj}, do_compression=True)

    def set_parameters(self, *args):
        """GENFIRE_Reconstructor.m:210-225."""
        if len(args) % 2 != 0:
            raise GenfireError(-1, "GENFIRE: Additional argument list not divisible by 2. "
                                   "Options should be 'key','value' pairs.")
        for k, v in zip(args[0::2], args[1::2]):
            if not hasattr(self, k) and k not in ("measuredK", "recIFFT"):
                raise GenfireError(-1, f"GENFIRE: Invalid option {k} provided.")
            setattr(self, k, v)
        return self

    def close(self):
        if self._job is not None:
            self._job.close()
            self._job = None


def calculate3Dprojection_RealSpaceinterp(Vol, phi, theta, psi, ctx: _lib.Context | None = None,
                                          interp: str = "linear"):
    """src/calculate3Dprojection_RealSpaceinterp.m: interp "linear" (trilinear)
    or "cubic" (the reference's interp3 'cubic', :45,51)."""
    return (ctx or default_context()).project(Vol, phi, theta, psi, interp=interp)


def Tian_calc_Rfactor_realspace(model, Projs, Angles, ctx: _lib.Context | None = None, interp: str = "linear"):
    """src/Tian_calc_Rfactor_realspace.m:1-24 -> (Rtotal, Rarr, simuProjs)."""
    return (ctx or default_context()).rfactor(model, np.asarray(Projs), np.asarray(Angles), interp=interp)
