"""The product's host mesh setup (libkutta_amd.so) against the oracle's
restatement and the reference's own mesh code: generated faces, perturbed
points, RCM permutation, FP64 geometry and the FV stencil bit for bit."""
import numpy as np
import pytest

import kutta_amd
import oracle

INT = ["face_offsets", "face_nodes", "owner", "neighbour", "cells_new_to_old", "faces_new_to_old",
       "cell_face_offsets", "cell_faces", "row_ptr", "col", "diag", "adjacent", "matrix_entry",
       "patches"]
FLOAT = ["diffusion", "lsq_x", "lsq_y", "lsq_z", "lsq_inverse", "nonorth_x", "nonorth_y",
         "nonorth_z", "cell_face_sign"]
DOUBLE = ["points", "face_centroid", "face_area", "cell_centroid", "cell_volume"]


@pytest.mark.parametrize("tet,shape,amplitude", [(True, (5, 4, 3), 0.0), (False, (6, 5, 4), 0.0),
                                                 (True, (11, 9, 8), 0.2), (False, (9, 8, 7), 0.2),
                                                 (True, (17, 16, 15), 0.2)])
def test_product_mesh_setup_bitwise_vs_oracle(tet, shape, amplitude):
    got = kutta_amd.MeshSetup(*shape, tet=tet, amplitude=amplitude)
    want = oracle.pressure_problem(*shape, tet=tet, amplitude=amplitude)
    assert got["rounds"][0] == want["rounds"][0]
    for name in INT:
        assert np.array_equal(got[name], want[name]), name
    for name in FLOAT:
        assert np.array_equal(got[name].view(np.uint32), want[name].view(np.uint32)), name
    for name in DOUBLE:
        assert np.array_equal(got[name].view(np.uint64), want[name].view(np.uint64)), name


@pytest.mark.slow
def test_product_mesh_setup_full_c2_size():
    """The 55^3 Freudenthal box (998,250 tets) used by the tetrahedral Poisson
    workload: sizes from the workload definition and bitwise agreement of the
    CSR pattern and coefficients with the oracle."""
    got = kutta_amd.MeshSetup(55, 55, 55)
    assert len(got["cell_volume"]) == 998_250
    assert len(got["neighbour"]) == 1_978_350
    assert len(got["owner"]) == 2_014_650
    assert len(got["col"]) == 4_954_950
    want = oracle.pressure_problem(55, 55, 55)
    for name in ("cells_new_to_old", "row_ptr", "col", "adjacent", "matrix_entry"):
        assert np.array_equal(got[name], want[name]), name
    assert np.array_equal(got["diffusion"].view(np.uint32), want["diffusion"].view(np.uint32))
