"""The MI355X steady SIMPLEC solver against the oracle's serial restatement
(bitwise, every field, through the C ABI) and against the reference's own
logged cavity results (fv/steady_solver.cpp; validation/cases/cavity.hpp)."""
import ctypes as C

import numpy as np
import pytest

import kutta_amd
import oracle
from cavity import (REFERENCE_129_RE100, REFERENCE_65_RE1000, centerline, interpolate, reference_error,
                    run_to_steady)

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p", "mass", "face_pressure", "boundary_pressure", "coefficient",
          "physical_diagonal", "consistent_inverse", "correction")
PAIRED_FIELDS = ("u_lo", "v_lo", "w_lo", "p_lo", "mass_lo")
# upload_mesh's FP32 arrays and the least-squares fits (host setup, uploaded).
SETUP_FIELDS = ("volume", "weight", "area_x", "owner_to_face_x", "neighbour_to_face_y", "owner_to_neighbour_z",
                "wall_normal_distance", "wall_distance", "velocity_lsq_inverse", "pressure_lsq_inverse",
                "velocity_lsq_boundary_x", "pressure_lsq_boundary_y")


def make(device, nx, ny, nz, conditions, settings):
    kinds, wall = kutta_amd.pack_conditions(conditions)
    gpu = kutta_amd.SteadySolver(device, nx, ny, nz, conditions, settings)
    cpu = oracle.SteadySolver(nx, ny, nz, kinds, wall, settings.packed())
    return gpu, cpu


def assert_same_state(gpu, cpu, paired):
    for name in FIELDS + (PAIRED_FIELDS if paired else ()):
        a, b = gpu.field(name), cpu.field(name)
        assert a.shape == b.shape, name
        bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
        assert bad.size == 0, f"{name}: {bad.size} differ, first {bad[:4]} gpu {a[bad[:4]]} oracle {b[bad[:4]]}"


def assert_same_report(gpu, cpu):
    g, o = gpu.report(), cpu.report()
    keys = kutta_amd.REPORT_FIELDS
    for k, name in enumerate(keys):
        if name.endswith("iterations") or name in ("iteration", "pressure_converged", "maximum_velocity",
                                                   "pseudo_time_step"):
            assert g[name] == o[k], name  # counts and max reductions are exact
        else:  # FP64 sums in another order
            assert g[name] == pytest.approx(o[k], rel=1e-9, abs=1e-300), name


CASES = [
    # nx, ny, nz, Re, paired, iterations
    (16, 16, 1, 100, True, 25),
    (16, 16, 1, 100, False, 25),
    (24, 20, 1, 400, True, 21),
    (7, 6, 5, 50, True, 12),   # a 3-D box: every velocity component moves
]


@pytest.mark.parametrize("nx,ny,nz,re,paired,iterations", CASES)
def test_simplec_matches_oracle_bitwise(device, nx, ny, nz, re, paired, iterations):
    print(f"[simplec] {nx}x{ny}x{nz} Re={re} paired={paired}", flush=True)
    conditions = kutta_amd.cavity_conditions()
    if nz > 1:  # closed 3-D box: the lid drags diagonally, symmetry replaced by walls
        conditions = {name: kutta_amd.Wall() for name in kutta_amd.PATCHES}
        conditions["ymax"] = kutta_amd.Wall((1.0, 0.0, 0.5))
    settings = kutta_amd.SolverSettings(viscosity=1.0 / re, momentum_relaxation=0.7, paired_state=paired,
                                        monitor_interval=7)
    gpu, cpu = make(device, nx, ny, nz, conditions, settings)
    assert gpu.info()["reference"] == cpu.reference()
    for name in SETUP_FIELDS:
        a, b = gpu.field(name), cpu.field(name)
        assert a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32)), name
    assert_same_state(gpu, cpu, paired)  # after the constructor's flux
    for block in (1, 4, iterations - 5):
        gpu.iterate(block)
        cpu.iterate(block)
        assert_same_state(gpu, cpu, paired)
        assert_same_report(gpu, cpu)
    if nz > 1:
        assert np.abs(gpu.field("w")).max() > 0


def test_c1_cavity_first_iterations_match_oracle(device):
    """The 128x128x1 Re=1000 cavity (momentum and pressure relaxation 1,
    monitor every 100) for 6 iterations, then to iteration 120, past the
    first 100-iteration monitor, bit for bit."""
    print("[simplec] C1 128x128x1 Re=1000", flush=True)
    settings = kutta_amd.SolverSettings(viscosity=1e-3, momentum_relaxation=1.0, pressure_relaxation=1.0,
                                        monitor_interval=100)
    gpu, cpu = make(device, 128, 128, 1, kutta_amd.cavity_conditions(), settings)
    gpu.iterate(6)
    cpu.iterate(6)
    assert_same_state(gpu, cpu, True)
    assert_same_report(gpu, cpu)
    assert np.all(gpu.field("w") == 0)  # symmetry front and back keep the flow planar
    gpu.iterate(114)
    cpu.iterate(114)
    print("[simplec] C1 after 120 iterations", flush=True)
    assert_same_state(gpu, cpu, True)
    assert_same_report(gpu, cpu)


@pytest.mark.parametrize("nx,ny,nz,paired", [(24, 20, 1, True), (24, 20, 1, False), (7, 6, 5, True)])
def test_batched_momentum_equals_sequential(device, nx, ny, nz, paired):
    """The batched momentum solve (no reference counterpart; the default on
    axis-aligned meshes) leaves every field and every iteration count exactly
    as the reference's component-by-component order does."""
    print(f"[simplec batched] {nx}x{ny}x{nz} paired={paired}", flush=True)
    conditions = kutta_amd.cavity_conditions()
    if nz > 1:  # two symmetry faces, so the boundary momentum terms act
        conditions = {name: kutta_amd.Wall() for name in kutta_amd.PATCHES}
        conditions["ymax"] = kutta_amd.Wall((1.0, 0.0, 0.5))
        conditions["zmin"] = kutta_amd.Symmetry()
        conditions["xmax"] = kutta_amd.Symmetry()
    settings = kutta_amd.SolverSettings(viscosity=1e-2, momentum_relaxation=0.7, paired_state=paired,
                                        monitor_interval=7)
    batched = kutta_amd.SteadySolver(device, nx, ny, nz, conditions, settings)
    sequential = kutta_amd.SteadySolver(device, nx, ny, nz, conditions, settings)
    assert batched.set_batched_momentum(True)
    assert not sequential.set_batched_momentum(False)
    for block in (1, 6, 13):
        batched.iterate(block)
        sequential.iterate(block)
        for name in FIELDS + (PAIRED_FIELDS if paired else ()):
            a, b = batched.field(name), sequential.field(name)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name
        ra, rb = batched.report(), sequential.report()
        for name in kutta_amd.REPORT_FIELDS:
            assert ra[name] == rb[name], name
    assert batched.info()["host_waits"] < sequential.info()["host_waits"]


def steady_cavity(device, n, re):
    settings = kutta_amd.SolverSettings(viscosity=1.0 / re, momentum_relaxation=1.0, pressure_relaxation=1.0,
                                        monitor_interval=100)
    solver = kutta_amd.SteadySolver(device, n, n, 1, kutta_amd.cavity_conditions(), settings)
    tolerance = 2e-6 if re == 1000 else 2e-5
    r, u, v, steady = run_to_steady(solver, solver.iterate, solver.report,
                                    lambda: (solver.field("u"), solver.field("v")), tolerance)
    cent = solver.centroids()
    return r, centerline(cent, u, True), centerline(cent, v, False), steady


def test_cavity_65_re1000_matches_reference_log(device):
    """validation::cavity(65x65, Re=1000, update tolerance 2e-6) as in the
    reference's refinement test; expected values from its CPU backend's log."""
    print("[simplec] 65x65 Re=1000 to steady", flush=True)
    r, U, V, steady = steady_cavity(device, 65, 1000)
    ref = REFERENCE_65_RE1000
    assert steady and r["iteration"] == ref["iterations"]
    eu = reference_error(U, "erturk2005_u_vertical_centerline_Re1000.csv", 0, 1)
    ev = reference_error(V, "erturk2005_v_horizontal_centerline_Re1000.csv", 0, 1)
    print(f"[simplec] Erturk u {eu:.12f} v {ev:.12f} centre {interpolate(U, .5):.12f} {interpolate(V, .5):.12f}")
    # The reference's converged state, to a few float ulps of the velocity.
    assert eu == pytest.approx(ref["u_error"], abs=3e-7)
    assert ev == pytest.approx(ref["v_error"], abs=3e-7)
    assert interpolate(U, .5) == pytest.approx(ref["u_center"], abs=3e-7)
    assert interpolate(V, .5) == pytest.approx(ref["v_center"], abs=3e-7)


def test_cavity_129_re100_matches_reference_log(device):
    """validation::cavity(129x129, Re=100) against Ghia's tables; the
    reference logs max deviations 0.005044 (u) and 0.008972 (v)."""
    print("[simplec] 129x129 Re=100 to steady", flush=True)
    r, U, V, steady = steady_cavity(device, 129, 100)
    ref = REFERENCE_129_RE100
    assert steady
    eu = reference_error(U, "ghia1982_u_vertical_centerline.csv", 2, 3)
    ev = reference_error(V, "ghia1982_v_horizontal_centerline.csv", 2, 3)
    print(f"[simplec] iterations {r['iteration']} Ghia u {eu:.6f} v {ev:.6f}")
    # Converged only to the harness's 2e-5 update tolerance: a few 1e-6 apart.
    assert eu == pytest.approx(ref["ghia_u"], abs=5e-6)
    assert ev == pytest.approx(ref["ghia_v"], abs=5e-6)


def test_forces_match_oracle(device):
    """forces() (fv/steady_solver.cpp:1711-1736) on the GPU against the
    oracle: per-face forces are bitwise equal, so the FP64 sums agree to
    summation order; ForceReport::coefficient follows its closed form."""
    print("[simplec] forces 20x16x1 Re=200", flush=True)
    settings = kutta_amd.SolverSettings(viscosity=1.0 / 200, momentum_relaxation=0.8)
    gpu, cpu = make(device, 20, 16, 1, kutta_amd.cavity_conditions(), settings)
    gpu.iterate(40)
    cpu.iterate(40)
    for patches, pref in ((["ymax"], 0.0), (["xmin", "xmax", "ymin", "ymax"], 0.25), (["ymin", "xmin"], -1.0)):
        g, o = gpu.forces(patches, pref), cpu.forces(patches, pref)
        for key in ("pressure", "viscous", "area_vector"):
            assert g[key] == pytest.approx(o[key], rel=1e-9, abs=1e-14), (patches, key)
        force = np.add(g["pressure"], g["viscous"])
        adjusted = force + np.array(g["area_vector"]) * (pref - 0.1)
        expected = adjusted @ np.array([1.0, 0.5, 0.0]) / (0.5 * 1.2 * 2.0 ** 2 * 0.3)
        got = kutta_amd.SteadySolver.force_coefficient(g, (1.0, 0.5, 0.0), 1.2, 2.0, 0.3, 0.1)
        assert got == pytest.approx(expected, rel=1e-12)
    with pytest.raises(kutta_amd.KuttaError):
        gpu.forces(["zmin"])  # a symmetry plane is not a wall


@pytest.mark.parametrize("nx,ny,nz,re,options", [
    (48, 40, 1, 400, dict(pre_sweeps=1, post_sweeps=1, coarse_sweeps=4, max_levels=6, k_min_rows=0)),
    (20, 18, 12, 100, dict(pre_sweeps=2, post_sweeps=2, aggregation_passes=2, k_min_rows=300)),
])
def test_k_cycle_pressure_matches_oracle_bitwise(device, nx, ny, nz, re, options):
    """The pressure AMG rebuilt with the K-cycle (no reference counterpart,
    both sides rebuilt from the same pinned operator after three outer
    iterations): every field and the report bit for bit over the next seven."""
    print(f"[simplec] K-cycle pressure {nx}x{ny}x{nz} Re={re} {options}", flush=True)
    conditions = kutta_amd.cavity_conditions()
    if nz > 1:
        conditions = {name: kutta_amd.Wall() for name in kutta_amd.PATCHES}
        conditions["ymax"] = kutta_amd.Wall((1.0, 0.0, 0.5))
    settings = kutta_amd.SolverSettings(viscosity=1.0 / re, momentum_relaxation=0.7, monitor_interval=5)
    amg = kutta_amd.AmgOptions(structure=kutta_amd.OperatorStructure.SymmetricMMatrix,
                               cycle=kutta_amd.AmgCycle.K, **options)
    gpu, cpu = make(device, nx, ny, nz, conditions, settings)
    gpu.iterate(3)
    cpu.iterate(3)
    gpu.set_pressure_options(amg)
    cpu.set_pressure_options(amg.packed(), amg.strength_threshold)
    for block in (1, 6):
        gpu.iterate(block)
        cpu.iterate(block)
        assert_same_state(gpu, cpu, True)
        assert_same_report(gpu, cpu)
    assert gpu.report()["pressure_iterations"] > 0


def test_pressure_options_must_keep_the_m_matrix_structure(device):
    gpu = kutta_amd.SteadySolver(device, 8, 8, 1, kutta_amd.cavity_conditions(), kutta_amd.SolverSettings())
    with pytest.raises(kutta_amd.KuttaError, match="M-matrix"):
        gpu.set_pressure_options(kutta_amd.AmgOptions(cycle=kutta_amd.AmgCycle.K))
    gpu.iterate(2)  # the solver it had is kept
