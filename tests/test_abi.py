"""The C ABI library loads on a machine without a GPU and exports every entry
point include/genfire.h declares; calls fail cleanly (no crash) without a device."""
import ctypes
import os
import re

import pytest

from genfire_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "genfire.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(genfire_[A-Za-z0-9_]+)\s*\(", src)))


def test_header_and_binding_agree():
    assert _declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(L, s)]
    assert not missing, missing


def test_defaults_mirror_reference_properties():
    p = _lib.default_params()   # GENFIRE_Reconstructor.m:30-70
    assert p.numIterations == 50 and p.oversamplingRatio == 3 and p.interpolationCutoffDistance == 0.7
    assert p.constraintEnforcementMode == 1 and p.constraintPositivity == 1 and p.constraintSupport == 1
    assert p.DFT_CentroSymmetricity == 1 and p.monitor_R == 0 and p.monitorR_loopLength == 10
    assert p.griddingMethod == 1 and p.allowMultipleGridMatches == 1   # :45-46
    assert list(p.vector1) == [0, 0, 1] and list(p.vector2) == [0, 1, 0] and list(p.vector3) == [1, 0, 0]


def test_status_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    L = _lib.load()
    h = ctypes.c_void_p()
    assert L.genfire_ctx_create(0, ctypes.byref(h)) == -2   # GENFIRE_ERR_HIP
    assert L.genfire_job_grid(None) == -1                   # GENFIRE_ERR_ARG
    assert L.genfire_version() >= 100
