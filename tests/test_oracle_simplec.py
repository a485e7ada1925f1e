"""The oracle's SIMPLEC restatement (oracle/simplec.cpp) pinned without a GPU:
its least-squares fits reproduce linear fields exactly, its BiCGStab is
independent of the convergence-check interval (a property the reference's own
tests assert), the step keeps planar and mirror-symmetric flows so, and the
65x65 Re=1000 cavity converges to the state the reference's CPU backend logs."""
import ctypes as C

import numpy as np
import pytest

import oracle
from cavity import REFERENCE_65_RE1000, centerline, interpolate, reference_error, run_to_steady

OMITTED, FACE, FOOT = 0, 1, 2


def lsq_gradient(nx, ny, nz, tet, amplitude, obs, coeffs):
    L = oracle.lib()
    L.oracle_lsq_gradient.argtypes = [C.c_int] * 4 + [C.c_double, C.c_ulonglong] + [C.c_void_p] * 4
    n = nx * ny * nz * (6 if tet else 1)
    grad = np.empty((n, 3), np.float32)
    cent = np.empty((n, 3), np.float64)
    o = np.array(obs, np.int32)
    c = np.array(coeffs, np.float64)
    got = L.oracle_lsq_gradient(nx, ny, nz, int(tet), amplitude, 7, o.ctypes.data, c.ctypes.data,
                                grad.ctypes.data, cent.ctypes.data)
    assert got == n, L.oracle_simplec_last_error()
    return grad, cent


@pytest.mark.parametrize("tet,amplitude", [(False, 0.0), (True, 0.0), (True, 0.2)])
@pytest.mark.parametrize("obs", [
    (FACE,) * 6,                                      # walls everywhere: velocity fit
    (FACE, FACE, FACE, FACE, FOOT, FOOT),             # cavity velocity: symmetry front/back
    (OMITTED,) * 6,                                   # closed-domain pressure fit
    (FOOT, OMITTED, FACE, FOOT, OMITTED, FACE),       # mixed
])
def test_least_squares_gradient_is_exact_for_linear_fields(tet, amplitude, obs):
    """fv/operators.cpp:97-137, 166-211 with fv_gradient: every row of a cell's
    fit is a difference of the linear field, so the fit returns the field's
    gradient projected onto the directions its rows span (a cell with fewer
    than three independent observations gets unit eigenvalues elsewhere,
    which contribute nothing). Per cell, the three unit-gradient fields give
    that projection P: it must be symmetric and idempotent, the identity
    wherever the observed rows span space (every cell of a hex box), and a
    fourth, oblique gradient must come back as P g. A transposed inverse plane,
    a dropped weight or a wrong normal-foot projection fails by orders of
    magnitude."""
    cols = [lsq_gradient(5, 4, 3, tet, amplitude, obs, c)[0].astype(np.float64)
            for c in ((0.3, 1, 0, 0), (-0.2, 0, 1, 0), (0.1, 0, 0, 1))]
    P = np.stack(cols, axis=2)  # P[c, i, j] = d(gradient i) / d(g_j)
    assert np.abs(P - P.transpose(0, 2, 1)).max() < 1e-4
    assert np.abs(P @ P - P).max() < 1e-4
    rank = np.rint(np.trace(P, axis1=1, axis2=2))
    if not tet:  # every hex cell has a neighbour or an observed face on each axis
            assert np.all(rank == 3)
    full = rank == 3
    assert full.mean() > 0.5
    assert np.abs(P[full] - np.eye(3)).max() < 1e-4
    g = np.array([0.7, -1.3, 0.45])
    grad, _ = lsq_gradient(5, 4, 3, tet, amplitude, obs, (0.25, *g))
    assert np.abs(grad - P @ g).max() < 1e-4


def test_bicgstab_result_is_independent_of_check_interval():
    """tests/linalg/test_solvers.cpp:359-389: reused BiCGStab-Jacobi solves
    with a residual check every 8 updates equal those checking every update
    (iterations, final residual and x)."""
    a = oracle.poisson(8, 6, 4, 1.0, False, False, 5.0)
    val = np.array(a.val, dtype=np.float64)
    for d in a.diag:
        val[d] += 100
    boosted = oracle.Csr.from_arrays(a.row_ptr, a.col, a.diag, val)
    rng = np.random.default_rng(3)
    for m, rel in ((boosted, 1e-4), (oracle.poisson(14, 12, 1, 1.0, wind=3.0), 1e-6)):
        b = rng.standard_normal(m.n_rows).astype(np.float32)
        x8, s8 = oracle.Solver("BicgstabJacobi", m).solve(b, rel_tol=rel, abs_tol=1e-20, max_iter=300, interval=8)
        x1, s1 = oracle.Solver("BicgstabJacobi", m).solve(b, rel_tol=rel, abs_tol=1e-20, max_iter=300, interval=1)
        assert s8["converged"] and s8["iterations"] == s1["iterations"]
        assert s8["final_residual"] == s1["final_residual"]
        assert np.array_equal(x8.view(np.uint32), x1.view(np.uint32))


def cavity(nx, ny, re, lid=1.0, paired=True, relaxation=1.0):
    kinds = (C.c_int32 * 6)(0, 0, 0, 0, 3, 3)
    wall = (C.c_double * 18)(*([0.0] * 9 + [lid, 0.0, 0.0] + [0.0] * 6))
    settings = (C.c_double * 11)(1.0, 1.0 / re, relaxation, 1.0, 1.0, 100, 1.0 if paired else 0.0, 1.0, 0, 0, 0)
    return oracle.SteadySolver(nx, ny, 1, kinds, wall, settings)


def test_planar_cavity_stays_planar_and_conservative():
    """With symmetry front and back the z velocity stays exactly zero, the
    momentum defects fall and the corrected continuity is near float roundoff."""
    s = cavity(12, 10, 100)
    s.iterate(5)
    first = s.report()
    s.iterate(35)
    last = s.report(monitor=True)
    assert np.all(s.field("w") == 0) and np.all(s.field("w_lo") == 0)
    assert last[2] < 0.05 * first[2] and last[6] < 0.05 * first[6]  # u, v scaled defects
    assert abs(last[17]) < 1e-6 * np.abs(s.field("mass")).sum()


def test_mirrored_lid_gives_mirrored_flow():
    """Reversing the lid mirrors the flow about x = 1/2: u(x,y) = -u'(1-x,y),
    v(x,y) = v'(1-x,y), p(x,y) = p'(1-x,y) up to the pinned gauge. A sign error
    in a boundary or upwind term breaks the mirror image."""
    a, b = cavity(10, 10, 100, lid=1.0), cavity(10, 10, 100, lid=-1.0)
    a.iterate(30)
    b.iterate(30)
    ca, cb = a.centroids(), b.centroids()
    key = {(round(x * 1e9), round(y * 1e9)): i for i, (x, y, _) in enumerate(cb)}
    mirror = np.array([key[(round((1 - x) * 1e9), round(y * 1e9))] for x, y, _ in ca])
    ua, va = a.field("u").astype(np.float64), a.field("v").astype(np.float64)
    ub, vb = b.field("u").astype(np.float64)[mirror], b.field("v").astype(np.float64)[mirror]
    scale = np.abs(ua).max()
    assert np.abs(ua + ub).max() < 1e-4 * scale
    assert np.abs(va - vb).max() < 1e-4 * scale


@pytest.mark.slow
def test_cavity_65_re1000_reaches_the_reference_state():
    """validation::cavity(65x65, Re=1000, update tolerance 2e-6)
    (tests/fv/test_cavity.cpp:40-60) on the oracle: steady after the same 600
    iterations as the reference's CPU backend, and the converged centreline
    errors against Erturk's tables and the centre velocities agree with its
    log to a few float ulps (the converged centre u is the same float)."""
    s = cavity(65, 65, 1000)
    names = ("iteration", "u_raw", "u_scaled")

    def report():
        r = s.report()
        return {"u_scaled": r[2], "v_scaled": r[6], "corrected_mass_imbalance": r[17], "iteration": r[0]}

    r, u, v, steady = run_to_steady(s, s.iterate, report, lambda: (s.field("u"), s.field("v")), 2e-6)
    assert steady and r["iteration"] == REFERENCE_65_RE1000["iterations"]
    s.iterate(300)  # to the fixed point
    cent = s.centroids()
    U, V = centerline(cent, s.field("u"), True), centerline(cent, s.field("v"), False)
    ref = REFERENCE_65_RE1000
    assert interpolate(U, .5) == pytest.approx(ref["u_center"], abs=1e-12)  # the log prints 12 digits
    assert interpolate(V, .5) == pytest.approx(ref["v_center"], abs=1e-7)
    assert reference_error(U, "erturk2005_u_vertical_centerline_Re1000.csv", 0, 1) == \
        pytest.approx(ref["u_error"], abs=2e-7)
    assert reference_error(V, "erturk2005_v_horizontal_centerline_Re1000.csv", 0, 1) == \
        pytest.approx(ref["v_error"], abs=2e-7)


UPLOAD_ARRAYS = ("area_x", "area_y", "area_z", "owner_to_face_x", "owner_to_face_y", "owner_to_face_z",
                 "neighbour_to_face_x", "neighbour_to_face_y", "neighbour_to_face_z", "owner_to_neighbour_x",
                 "owner_to_neighbour_y", "owner_to_neighbour_z", "weight", "volume", "sign",
                 "wall_normal_distance", "wall_distance")


@pytest.mark.skipif(not oracle.ref_built("ref_devmesh"), reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("nx,ny,nz,tet", [(6, 5, 4, False), (4, 3, 5, True), (9, 7, 1, False)])
def test_upload_mesh_matches_reference_build(nx, ny, nz, tet):
    """fv/device_mesh.cpp:208-256 built from the reference tree: every FP32
    face and cell array of upload_mesh equals the restatement's bit for bit."""
    wall = [1, 0, 1, 1, 0, 1]
    for name in UPLOAD_ARRAYS:
        a = oracle.upload_mesh_array(nx, ny, nz, tet, wall, name)
        b = oracle.upload_mesh_array(nx, ny, nz, tet, wall, name, reference=True)
        assert a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32)), name


@pytest.mark.parametrize("tet", [False, True])
def test_wall_distance_closed_form(tet):
    """fv/device_mesh.cpp:232-245: in the unit box every wall patch is a
    whole face of the cube, so a cell's distance to the walls is its
    centroid's distance to the nearest wall plane; with no wall it is
    infinite. A wrong triangle fan, a dropped edge case or a square root
    left out fails this."""
    nx, ny, nz = 5, 4, 3
    wall = [1, 0, 0, 1, 1, 0]  # xmin, ymax, zmin
    d = oracle.upload_mesh_array(nx, ny, nz, tet, wall, "wall_distance").astype(np.float64)
    s = oracle.SteadySolver(nx, ny, nz, (C.c_int32 * 6)(0, 3, 3, 0, 0, 3), (C.c_double * 18)(),
                            (C.c_double * 11)(1, .01, .7, 1, 1, 20, 1, 1, 0, 0, 0)) if not tet else None
    if s is not None:  # the same distances inside the solver, in its cell order
        assert np.array_equal(s.field("wall_distance"), d.astype(np.float32))
    cells = nx * ny * nz * (6 if tet else 1)
    assert d.shape == (cells,) and np.all(d > 0)
    # The cell centroids come back in the solver's order from a symmetric setup.
    t = oracle.SteadySolver(nx, ny, nz, (C.c_int32 * 6)(3, 3, 3, 3, 3, 3), (C.c_double * 18)(),
                            (C.c_double * 11)(1, .01, .7, 1, 1, 20, 1, 1, 0, 0, 0)) if not tet else None
    if t is not None:
        x = t.centroids()
        expect = np.minimum(np.minimum(x[:, 0], 1 - x[:, 1]), x[:, 2])
        assert np.allclose(d, expect, rtol=1e-6, atol=0)
        assert np.all(np.isinf(t.field("wall_distance")))
    else:
        assert d.max() < 1.0  # the reference build pins the tet values bit for bit


def test_steady_cavity_walls_balance_the_lid():
    """At steady state the fluid is in equilibrium, so the x forces it exerts
    on the four walls (pressure plus viscous, fv/steady_solver.cpp:1711-1736)
    cancel: the lid's drag (the fluid resists the lid, so along -x) returns
    through the other walls, up to the discretization error of the wall
    tractions (0.1% on this 16x16 grid). The area vector of a closed set of
    walls in a one-cell-deep box is the symmetry planes' missing share: zero
    in x and y."""
    s = cavity(16, 16, 100)
    s.iterate(400)
    lid = s.forces(["ymax"])
    total = s.forces(["xmin", "xmax", "ymin", "ymax"])
    fx_lid = lid["pressure"][0] + lid["viscous"][0]
    fx = total["pressure"][0] + total["viscous"][0]
    assert fx_lid < -0.1
    assert abs(fx) < 3e-3 * abs(fx_lid)
    assert total["area_vector"][:2] == pytest.approx([0.0, 0.0], abs=1e-15)
    assert lid["area_vector"] == pytest.approx([0.0, 1.0, 0.0], rel=1e-12)  # the unit lid


def test_k_cycle_pressure_reaches_the_same_steady_cavity():
    """The pressure AMG rebuilt as the K-cycle (no reference counterpart)
    changes each outer iteration's pressure solve but not the fixed point:
    both solvers reach the same steady 24x24 Re=100 cavity."""
    states, early = [], []
    for k_cycle in (False, True):
        s = cavity(24, 24, 100, relaxation=0.7)
        if k_cycle:
            packed = (C.c_int32 * 11)(1, 1, 4, 24, 6, 3, 0, 1, 1, 0, 0)  # V(1,1), 4 coarsest sweeps, 6 levels
            s.set_pressure_options(packed, 0.25)
        s.iterate(10)
        early.append(s.field("p"))

        def report():
            r = s.report()
            return {"u_scaled": r[2], "v_scaled": r[6], "corrected_mass_imbalance": r[17]}

        r, u, v, steady = run_to_steady(s, s.iterate, report, lambda: (s.field("u"), s.field("v")), 1e-6, 3000)
        assert steady
        s.iterate(200)
        u, v = s.field("u"), s.field("v")
        states.append((u, v, s.field("p")))
    assert not np.array_equal(early[0], early[1])  # the rebuilt hierarchy is the one in use
    for a, b in zip(*states):
        assert np.allclose(a, b, rtol=0, atol=2e-5 * np.abs(b).max())


def test_pressure_options_keep_the_m_matrix_structure():
    s = cavity(8, 8, 100)
    with pytest.raises(oracle.OracleError, match="M-matrix"):
        s.set_pressure_options((C.c_int32 * 11)(1, 1, 4, 24, 6, 3, 0, 0, 1, 0, 0), 0.25)
    s.iterate(2)
