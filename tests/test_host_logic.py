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


def test_fft_gridding_edge_rules():
    """fillInFourierGrid.m:38-42: the FFT-gridding grid edge comes from Dim2 and
    must be a whole number no smaller than Dim1 (My_paddzero.m:27-28)."""
    G = _obj()
    G.griddingMethod = 1
    G.oversamplingRatio = 2.5
    G.InputProjections = np.zeros((16, 15, 3), np.float32)
    G.InputAngles = np.zeros((3, 3))
    G.CheckPrepareData()                         # odd Dim2: 1 + 14*2.5 = 36
    G.InputProjections = np.zeros((16, 16, 3), np.float32)
    G.oversamplingRatio = 2.1
    with pytest.raises(GenfireError, match="whole-number grid edge"):
        G.CheckPrepareData()                     # 16*2.1 = 33.6
    G.oversamplingRatio = 1
    G.InputProjections = np.zeros((20, 16, 3), np.float32)
    with pytest.raises(GenfireError, match="paddedsize"):
        G.CheckPrepareData()


def test_volumn_index_matches_oracle(O):
    for big, small in [((128, 128, 128), (64, 64, 64)), ((45, 46, 45), (15, 16, 15)), ((9, 8, 9), (4, 3, 4))]:
        np.testing.assert_array_equal(My_volumn_index(big, small), O.volumn_index(big, small))


def test_save_results_mat_roundtrip(tmp_path):
    """SaveResults (GENFIRE_Reconstructor.m:205-207) writes a MAT v5 struct 'obj'
    that loadmat (as MATLAB's load) reads back field by field."""
    import scipy.io
    G = GENFIRE_Reconstructor()
    G.filename_Results = str(tmp_path / "res" / "GENFIRE_rec.mat")
    rng = np.random.default_rng(0)
    G.reconstruction = rng.random((4, 5, 4)).astype(np.float32)
    G.errK = rng.random(7).astype(np.float32)
    G.Rfree_complex = rng.random((3, 7)).astype(np.float32)
    G.InputAngles = rng.random((5, 3))
    G.SaveResults()
    d = scipy.io.loadmat(G.filename_Results, squeeze_me=False)["obj"][0, 0]
    np.testing.assert_array_equal(d["reconstruction"], G.reconstruction)
    np.testing.assert_array_equal(d["errK"].ravel(), G.errK)
    np.testing.assert_array_equal(d["Rfree_complex"], G.Rfree_complex)
    np.testing.assert_array_equal(d["InputAngles"], G.InputAngles)
    assert d["errR"].size == 0 and int(d["numIterations"].ravel()[0]) == G.numIterations


def test_rfree_parameters_reach_the_library():
    G = GENFIRE_Reconstructor()
    G.calculate_Rfree, G.numBinsRfree, G.percentValuesForRfree, G.RandomSeed = 1, 6, 0.1, 42
    p = G._params()
    assert (p.calculate_Rfree, p.numBinsRfree, p.percentValuesForRfree, p.RandomSeed) == (1, 6, 0.1, 42)


def test_read_files_mat_v5(tmp_path):
    """readFiles (GENFIRE_Reconstructor.m:82-112) on MAT v5 files, as the
    reference's data_turb3 set is stored: the fixture's projections (64 x 64 x 41,
    written as double like the reference's .mat), angles (41 x 3), a support and
    an initial model come back with their shapes and values; a missing support
    file makes the all-ones support of the projections' size (:99-102), a
    missing projections file is the reference's error."""
    import os
    import scipy.io
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "turb3.npz"))
    pj, ang = d["projections"].astype(np.float64), d["angles"]
    sup = np.zeros((64, 64, 64), np.uint8)
    sup[16:48, 16:48, 16:48] = 1
    init = np.random.default_rng(3).random((64, 64, 64)).astype(np.float32)
    for name, key, val in [("pj.mat", "pj_new_Turbu3_1015", pj), ("ang.mat", "euler_angles", ang),
                           ("sup.mat", "support", sup), ("init.mat", "model", init)]:
        scipy.io.savemat(str(tmp_path / name), {key: val})
    G = GENFIRE_Reconstructor()
    G.filename_Projections = str(tmp_path / "pj.mat")
    G.filename_Angles = str(tmp_path / "ang.mat")
    G.filename_Support = str(tmp_path / "sup.mat")
    G.filename_InitialModel = str(tmp_path / "init.mat")
    G.readFiles()
    np.testing.assert_array_equal(G.InputProjections, pj)
    np.testing.assert_array_equal(G.InputAngles, ang)
    np.testing.assert_array_equal(G.Support, sup != 0)
    np.testing.assert_array_equal(G.initialObject, init)
    np.testing.assert_allclose(G.InputProjections.sum(axis=(0, 1)), d["proj_sums"], rtol=1e-6)
    G2 = GENFIRE_Reconstructor()
    G2.filename_Projections = str(tmp_path / "pj.mat")
    G2.filename_Angles = str(tmp_path / "ang.mat")
    G2.filename_Support = str(tmp_path / "missing.mat")
    G2.readFiles()
    assert G2.Support.shape == (64, 64, 64) and G2.Support.all()
    G3 = GENFIRE_Reconstructor()
    G3.filename_Projections = str(tmp_path / "none.mat")
    with pytest.raises(GenfireError, match="Projections file does not exist"):
        G3.readFiles()
