"""Pins the oracle's mesh restatement: against the reference's own generators,
geometry and RCM compiled from its sources (oracle/_ref, this container only),
and against closed forms that hold for any correct implementation."""
import ctypes as C

import numpy as np
import pytest

import oracle

REF = oracle.lazy(oracle.ref_mesh)
needs_ref = pytest.mark.skipif(not oracle.ref_built("ref_mesh"), reason="oracle/_ref not built (reference tree absent)")


def ref_array(h, name, dtype):
    n = REF.ref_array(h, name.encode(), None)
    out = np.empty(n, dtype)
    REF.ref_array(h, name.encode(), C.c_void_p(out.ctypes.data))
    return out


@needs_ref
@pytest.mark.parametrize("tet,shape,amplitude", [(True, (6, 5, 4), 0.0), (False, (7, 6, 5), 0.0),
                                                 (True, (9, 8, 7), 0.2), (True, (13, 11, 12), 0.2)])
def test_mesh_pipeline_bitwise_vs_reference(tet, shape, amplitude):
    b = oracle.pressure_problem(*shape, tet=tet, amplitude=amplitude)
    h = REF.ref_box(*shape, int(tet))
    assert h, REF.ref_mesh_last_error().decode()
    try:
        # generated mesh, before perturbation and renumbering
        assert np.array_equal(ref_array(h, "face_nodes", np.int32), b["generated_face_nodes"])
        assert np.array_equal(ref_array(h, "owner", np.int32), b["generated_owner"])
        assert np.array_equal(ref_array(h, "neighbour", np.int32), b["generated_neighbour"])
        gen = ref_array(h, "points", np.float64)
        assert np.array_equal(gen.view(np.uint64),
                              b["points_generated"].view(np.uint64)) or amplitude > 0
        # same perturbed points into the reference, then its RCM and geometry
        pts = np.ascontiguousarray(b["points"])
        REF.ref_set_points(h, C.c_void_p(pts.ctypes.data))
        assert REF.ref_renumber(h) == 0, REF.ref_mesh_last_error().decode()
        assert REF.ref_geometry(h) == 0, REF.ref_mesh_last_error().decode()
        for name in ("cells_new_to_old", "faces_new_to_old", "face_offsets", "face_nodes",
                     "owner", "neighbour", "patches", "cell_face_offsets", "cell_faces"):
            assert np.array_equal(ref_array(h, name, np.int32), b[name]), name
        for name in ("face_centroid", "face_area", "cell_centroid", "cell_volume"):
            assert np.array_equal(ref_array(h, name, np.float64).view(np.uint64),
                                  b[name].view(np.uint64)), name
    finally:
        REF.ref_free(h)


def test_box_closed_forms():
    """Unperturbed box: tet volumes h^3/6, volume sums to 1, face areas sum to
    the box surface, affine LSQ rows reproduce exact gradients."""
    nx, ny, nz = 6, 5, 4
    b = oracle.pressure_problem(nx, ny, nz, tet=True, amplitude=0.0)
    vol = b["cell_volume"]
    assert len(vol) == 6 * nx * ny * nz
    assert np.allclose(vol, 1.0 / (nx * ny * nz) / 6, rtol=1e-12)
    assert abs(vol.sum() - 1.0) < 1e-12
    area = b["face_area"].reshape(-1, 3)
    n_int = len(b["neighbour"])
    boundary = np.abs(area[n_int:]).sum()
    assert abs(boundary - 6.0) < 1e-12


def test_pressure_operator_is_symmetric_m_matrix():
    """fv_pressure_assemble contract (tests/fv/test_pressure_matrix.cpp:25-54):
    symmetric, nonpositive off-diagonals, diagonal >= sum |off| (upward
    rounding), zero RHS for zero flux; diffusion |S|^2/(S.d) recomputed in FP64."""
    b = oracle.pressure_problem(9, 8, 7, amplitude=0.2)
    rp, col, diag, v = b["row_ptr"], b["col"], b["diag"], b["values"]
    n = len(rp) - 1
    dense = {}
    for i in range(n):
        off = 0.0
        for j in range(rp[i], rp[i + 1]):
            dense[(i, col[j])] = v[j]
            if col[j] != i:
                assert v[j] <= 0
                off -= float(v[j])
        assert float(v[diag[i]]) >= off
    for (i, j), val in dense.items():
        assert dense[(j, i)] == val
    assert np.all(b["rhs"] == 0)
    area = b["face_area"].reshape(-1, 3)
    cc = b["cell_centroid"].reshape(-1, 3)
    fc = b["face_centroid"].reshape(-1, 3)
    owner, nb = b["owner"], b["neighbour"]
    d = np.where((np.arange(len(owner)) < len(nb))[:, None],
                 cc[np.concatenate([nb, owner[len(nb):]])], fc) - cc[owner]
    coeff = (area * area).sum(1) / (area * d).sum(1)
    assert np.array_equal(coeff.astype(np.float32), b["diffusion"])


def test_rcm_is_deterministic_and_a_permutation():
    """tests/mesh/test_mesh.cpp:191-244: RCM is a permutation, deterministic,
    and keeps internal faces sorted by (owner, neighbour) with owner < neighbour."""
    a = oracle.pressure_problem(7, 6, 5, amplitude=0.2)
    b = oracle.pressure_problem(7, 6, 5, amplitude=0.2)
    perm = a["cells_new_to_old"]
    assert np.array_equal(np.sort(perm), np.arange(len(perm)))
    assert np.array_equal(perm, b["cells_new_to_old"])
    owner, nb = a["owner"], a["neighbour"]
    assert np.all(owner[:len(nb)] < nb)
    key = owner[:len(nb)].astype(np.int64) * (1 << 32) + nb
    assert np.all(np.diff(key) > 0)


@pytest.mark.parametrize("shape,amplitude,seed", [((9, 8, 7), 0.2, 20261202), ((12, 5, 6), 0.35, 7)])
def test_perturbation_and_reference_cell_properties(shape, amplitude, seed):
    """The seeded perturbation of the benchmark box (oracle/mesh.cpp, the
    workload definition): every node moves by at most amplitude x the cell
    size along each axis, nodes on a face of the box stay on its plane
    (corners and edges stay put), every cell keeps a positive volume, the
    same seed gives the same points and another seed others; the pinned
    reference cell (-2: the cell holding the origin) has the origin among its
    nodes."""
    b = oracle.pressure_problem(*shape, tet=True, amplitude=amplitude, seed=seed)
    p = b["points"].reshape(-1, 3)
    # The generator's lattice, node v = i + (nx+1) (j + (ny+1) k), before the noise.
    nx, ny, nz = shape
    v = np.arange(p.shape[0])
    g = np.stack([(v % (nx + 1)) / nx, (v // (nx + 1) % (ny + 1)) / ny, (v // ((nx + 1) * (ny + 1))) / nz], 1)
    assert np.array_equal(oracle.pressure_problem(*shape, tet=True, amplitude=0.0)["points"].reshape(-1, 3), g)
    h = 1.0 / np.asarray(shape, np.float64)
    assert np.all(np.abs(p - g) <= amplitude * h * (1 + 1e-12))
    assert np.any(np.abs(p - g) > 0.5 * amplitude * h)  # the noise is not degenerate
    for axis in range(3):
        for plane in (0.0, 1.0):
            on = g[:, axis] == plane
            assert on.any() and np.array_equal(p[on, axis], g[on, axis])
    assert np.all(b["cell_volume"] > 0)
    assert b["cell_volume"].sum() == pytest.approx(1.0, rel=1e-12)
    again = oracle.pressure_problem(*shape, tet=True, amplitude=amplitude, seed=seed)
    assert np.array_equal(again["points"].view(np.uint64), b["points"].view(np.uint64))
    other = oracle.pressure_problem(*shape, tet=True, amplitude=amplitude, seed=seed + 1)
    assert not np.array_equal(other["points"], b["points"])
    ref = int(b["reference"][0])
    faces = b["cell_faces"][b["cell_face_offsets"][ref]:b["cell_face_offsets"][ref + 1]]
    offsets, face_nodes = b["face_offsets"], b["face_nodes"]
    nodes = np.concatenate([face_nodes[offsets[f]:offsets[f + 1]] for f in faces])
    assert np.any(np.all(p[nodes] == 0.0, axis=1))
