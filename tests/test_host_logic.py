"""Host-side mirror of the GENFIRE_Reconstructor object (no GPU): property
defaults, CheckPrepareData validation and sizes, set_parameters, My_volumn_index."""
import numpy as np
import pytest

from genfire_amd._lib import GenfireError
from genfire_amd.reconstructor import GENFIRE_Reconstructor, My_volumn_index


def _obj(P=5, d1=16, d2=16):
    G = GENFIRE_Reconstructor(ctx=object())
    G.InputProjections = np.zeros((d1, d2, P), np.float32)
    G.InputAngles = np.zeros((P, 3))
    return G


def test_defaults():
    G = GENFIRE_Reconstructor(ctx=object())
    assert (G.numIterations, G.oversamplingRatio, G.interpolationCutoffDistance) == (50, 3, 0.7)
    assert G.griddingMethod == 1 and G.constraintEnforcementMode == 1 and G.DFT_CentroSymmetricity == 1
    assert G.monitor_R == 0 and G.monitorR_loopLength == 10


def test_check_prepare_sizes():
    G = _obj(P=7, d1=15, d2=12)
    G.oversamplingRatio = 2.5
    G.CheckPrepareData()
    assert (G.NumProjs, G.Dim1, G.Dim2) == (7, 15, 12)
    assert G.n1_oversampled == 38 and G.n2_oversampled == 30   # round(37.5) half away from zero
    assert G.Support.shape == (15, 12, 15) and G.Support.all()


@pytest.mark.parametrize("mutate,msg", [
    (lambda G: setattr(G, "InputAngles", np.zeros((4, 3))), "Number of projections and Angles does not match"),
    (lambda G: setattr(G, "InputAngles", np.zeros((5, 4))), "larger than three"),
    (lambda G: setattr(G, "initialObject", np.zeros((3, 3, 3))), "Initial Model size"),
    (lambda G: setattr(G, "griddingMethod", 3), "Unrecognized gridding method"),
    (lambda G: setattr(G, "constraintEnforcementMode", 4), "constraintEnforcementMode value 4"),
])
def test_check_prepare_errors(mutate, msg):
    G = _obj()
    mutate(G)
    with pytest.raises(GenfireError, match=msg):
        G.CheckPrepareData()


def test_single_angle_column_becomes_theta():
    G = _obj(P=3)
    G.InputAngles = np.array([10.0, 20.0, 30.0])
    G.CheckPrepareData()
    np.testing.assert_array_equal(G.InputAngles, [[0, 10, 0], [0, 20, 0], [0, 30, 0]])


def test_set_parameters():
    G = _obj()
    G.set_parameters("numIterations", 7, "oversamplingRatio", 2)
    assert G.numIterations == 7 and G.oversamplingRatio == 2
    with pytest.raises(GenfireError, match="not divisible by 2"):
        G.set_parameters("numIterations")
    with pytest.raises(GenfireError, match="Invalid option"):
        G.set_parameters("noSuchProperty", 1)


def test_fft_gridding_is_refused():
    G = _obj()
    G.CheckPrepareData()
    with pytest.raises(GenfireError, match="griddingMethod"):
        G.runGridding()


def test_volumn_index_matches_oracle(O):
    for big, small in [((128, 128, 128), (64, 64, 64)), ((45, 46, 45), (15, 16, 15)), ((9, 8, 9), (4, 3, 4))]:
        np.testing.assert_array_equal(My_volumn_index(big, small), O.volumn_index(big, small))
