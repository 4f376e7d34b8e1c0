"""Writes tests/golden/turb3.npz from the reference's data_turb3/*.mat files.

The MAT files are read as data with scipy.io.loadmat (MAT v5, no code executed).
Stored: the 64x64x41 projections as float32 (the gridding demotes the FP64
projections to single, fillInFourierGrid_DFT.m:243), their FP64 per-projection
sums, and both angle sets (41 x 3, degrees).

    python tests/golden/make_fixtures.py /root/reference/data_turb3
"""
import os
import sys

import numpy as np
import scipy.io

src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/data_turb3"
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "turb3.npz")
pj = scipy.io.loadmat(os.path.join(src, "pj_new_Turbu3_1015.mat"))["pj_new_Turbu3_1015"]
ang = scipy.io.loadmat(os.path.join(src, "angles_usedin_refine_1015.mat"))["euler_angles"]
gt = scipy.io.loadmat(os.path.join(src, "GroundTruth_angles_Turbu3_1015.mat"))["angles_Turbu3_1015"]
np.savez_compressed(out, projections=pj.astype(np.float32), proj_sums=pj.sum(axis=(0, 1)),
                    angles=ang.astype(np.float64), angles_truth=gt.astype(np.float64))
print(out, pj.shape, ang.shape, gt.shape)
