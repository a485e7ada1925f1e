"""The oracle's SIMPLEC with velocity inlets and pressure outlets
(oracle/simplec.cpp), pinned by the closed forms of the reference's own tests
(tests/fv/test_steady.cpp:22-47, 84-102, 104-126, 131-157): a uniform stream
through an inlet and an outlet is an exact steady state, an outlet's default
backflow is along the face normal, a closed domain's gauge is zero at the
reference cell, and the float-pair state reaches a continuity floor the
float32 state cannot."""
import numpy as np
import pytest

import kutta_amd
import oracle
from kutta_amd import PATCHES, PressureOutlet, SolverSettings, Symmetry, VelocityInlet, Wall

# Report positions (ka_simplec_report order).
ITERATION, CONTINUITY_RAW, CONTINUITY_SCALED = 0, 13, 14


def approx(a, b, eps):
    """doctest::Approx(b).epsilon(eps) == a, elementwise."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.all(np.abs(a - b) < eps * (1 + np.maximum(np.abs(a), np.abs(b))))


def make(n, conditions, settings=None, extent=(1.0, 1.0, 0.1)):
    s = settings or SolverSettings()
    kinds, velocity, pressure, total = kutta_amd.pack_open_conditions(conditions)
    return oracle.SteadySolver(*n, kinds, velocity, s.packed(), pressure, total, extent)


def initialize(solver, velocity, pressure):
    u, v, w, p = kutta_amd.initial_fields(solver.centroids(), velocity, pressure)
    solver.initialize(u, v, w, p)


def through_flow(through):
    c = {}
    for name in PATCHES:
        if through and name == "xmin":
            c[name] = VelocityInlet((through, 0.0, 0.0))
        elif through and name == "xmax":
            c[name] = PressureOutlet(0.0, (through, 0.0, 0.0))
        elif through or name in ("zmin", "zmax"):
            c[name] = Symmetry()
        else:
            c[name] = Wall()
    return c


@pytest.mark.parametrize("through", [0, 1, -1])
def test_rest_and_uniform_through_flow_are_preserved(through):
    """tests/fv/test_steady.cpp:22-47: rest in a closed box, and a uniform
    stream entering at xmin and leaving through an outlet at xmax (slip sides),
    stay as they are: u equal to the stream, v and p zero, continuity zero."""
    s = make((8, 4, 1), through_flow(through))
    assert (s.reference() >= 0) == (through == 0)  # an outlet pins no cell
    initialize(s, (through, 0.0, 0.0), 0.0)
    s.iterate(5)
    report = s.report(monitor=True)
    s.iterate(1)
    assert s.report()[ITERATION] == 6
    assert approx(s.field("u"), through, 1e-5)
    assert np.abs(s.field("v")).max() < 1e-6
    assert np.abs(s.field("p")).max() < 1e-6
    assert report[CONTINUITY_RAW] < 1e-6


def test_outlet_default_backflow_is_normal():
    """tests/fv/test_steady.cpp:84-102: a stream leaving through the xmin
    'inlet' (u = -1) enters through a default PressureOutlet; its backflow
    velocity follows the face normal, so the reversed stream is steady at
    density 2."""
    c = {name: Symmetry() for name in PATCHES}
    c["xmin"] = VelocityInlet((-1.0, 0.0, 0.0))
    c["xmax"] = PressureOutlet()
    s = make((4, 3, 1), c, SolverSettings(density=2.0))
    initialize(s, (-1.0, 0.0, 0.0), 0.0)
    s.iterate(10)
    assert approx(s.field("u"), -1.0, 1e-6)
    assert np.abs(s.field("v")).max() < 1e-6
    assert np.abs(s.field("p")).max() < 1e-6


def test_closed_domain_gauge_is_zero_at_the_reference_cell():
    """tests/fv/test_steady.cpp:104-126: a lid-driven box initialised with
    p = x + 2y keeps zero pressure in the cell holding the reference point
    (0.75, 0.25, 0.05), and nonzero pressure elsewhere."""
    c = {name: Wall() for name in PATCHES}
    c["ymax"] = Wall((1.0, 0.0, 0.0))
    c["zmin"] = c["zmax"] = Symmetry()
    s = make((6, 6, 1), c, SolverSettings(reference_point=(0.75, 0.25, 0.05)))
    initialize(s, (0.0, 0.0, 0.0), lambda x: x[0] + 2 * x[1])
    s.iterate(10)
    centres = s.centroids()
    pinned = np.flatnonzero((np.abs(centres[:, 0] - 0.75) < 1e-9) & (np.abs(centres[:, 1] - 0.25) < 1e-9))
    assert len(pinned) == 1 and s.reference() == pinned[0]
    p = s.field("p")
    assert p[pinned[0]] == 0.0
    assert np.abs(p).max() > 1e-3


def channel():
    c = {}
    for name in PATCHES:
        c[name] = Symmetry() if name in ("zmin", "zmax") else Wall()
    c["xmin"] = VelocityInlet((1.0, 0.0, 0.0))
    c["xmax"] = PressureOutlet(0.0, (1.0, 0.0, 0.0))
    return c


def test_float_pair_state_falls_below_the_float32_continuity_floor():
    """tests/fv/test_steady.cpp:131-157: a 3 x 1 channel (24 x 8 cells, nu
    0.02) after 200 iterations: the float32 state stalls above 1e-6 scaled
    continuity, the float pair falls below 1e-7, and their velocities agree to
    1e-5."""
    u, continuity = {}, {}
    for paired in (False, True):
        s = make((24, 8, 1), channel(), SolverSettings(viscosity=0.02, paired_state=paired), (3.0, 1.0, 0.1))
        initialize(s, (1.0, 0.0, 0.0), 0.0)
        s.iterate(200)
        continuity[paired] = s.report(monitor=True)[CONTINUITY_SCALED]
        u[paired] = s.field("u")
    print(f"[open channel] continuity float32 {continuity[False]:.3g}, pair {continuity[True]:.3g}")
    assert continuity[False] > 1e-6 and continuity[True] < 1e-7
    assert approx(u[True], u[False], 1e-5)


def duct(outlet):
    c = {name: Wall() for name in PATCHES}
    c["xmin"] = VelocityInlet((1.0, 0.0, 0.0))
    c["xmax"] = outlet
    return c


def test_outlet_pressure_sets_the_pressure_level():
    """Incompressible flow sees only pressure differences: raising the outlet
    pressure by 0.5 shifts every cell's pressure by 0.5 and leaves the
    developing duct flow unchanged, whose cross-section mean stays the inlet's
    (a sign or scaling error in the prescribed outlet pressure breaks this)."""
    runs = []
    for level in (0.0, 0.5):
        s = make((14, 6, 5), duct(PressureOutlet(level)), SolverSettings(viscosity=0.02), (2.0, 1.0, 1.0))
        initialize(s, (0.0, 0.0, 0.0), 0.0)
        s.iterate(20)
        runs.append((s.field("u"), s.field("p")))
    (u0, p0), (u1, p1) = runs
    # Equal up to the float32 rounding of the shifted level, carried through
    # 20 inexact linear solves (observed 4e-4).
    assert np.abs(u1 - u0).max() < 2e-3
    assert np.abs((p1 - p0) - 0.5).max() < 2e-3
    assert abs(u0.mean() - 1.0) < 1e-3 and u0.max() > 1.5


def test_open_conditions_are_validated():
    c = through_flow(1)
    c["xmax"] = PressureOutlet(float("nan"))
    with pytest.raises(oracle.OracleError):
        make((4, 2, 1), c)
    with pytest.raises(kutta_amd.KuttaError):
        kutta_amd.pack_open_conditions({name: object() for name in PATCHES})


def couette():
    c = {name: Symmetry() for name in PATCHES}
    c["ymax"] = Wall((1.0, 4.0, 0.0))
    c["ymin"] = Wall()
    c["xmin"] = c["xmax"] = VelocityInlet(lambda x: (x[1], 0.0, 0.0))
    return c


def make_profiled(n, conditions, settings, extent=(1.0, 1.0, 0.1)):
    kinds, velocity, pressure, total = kutta_amd.pack_open_conditions(conditions)
    profiled, callback = kutta_amd.profile_callback(conditions)
    return oracle.SteadySolver(*n, kinds, velocity, settings.packed(), pressure, total, extent, profiled, callback)


def test_planar_couette_flow_with_moving_and_slip_walls():
    """tests/fv/test_steady.cpp:49-82: a lid moving at (1, 4, 0) over a wall
    at rest, inlets at both ends carrying the linear profile u = y: the steady
    flow is u = y to 2e-4 (the lid's normal component is discarded), v
    vanishes, the lid's viscous drag is -nu dU/dy A = -0.1 and its drag
    coefficient -2, with no pressure force along x; moving the reference
    pressure to 2 changes the lid's y pressure force by -2 x 0.1."""
    s = make_profiled((4, 12, 1), couette(), SolverSettings(viscosity=1.0))
    s.iterate(400)
    u, v = s.field("u"), s.field("v")
    y = s.centroids()[:, 1]
    print(f"[couette] max error {np.abs(u - y).max():.3g}")
    assert np.abs(u - y).max() < 2e-4
    assert np.abs(v).max() < 2e-4
    force = s.forces(["ymax"])
    assert force["viscous"][0] == pytest.approx(-0.1, rel=2e-4)
    assert abs(force["pressure"][0]) < 1e-6
    assert force["viscous"][0] / (0.5 * 1.0 * 1.0 * 0.1) == pytest.approx(-2.0, rel=2e-4)
    shifted = s.forces(["ymax"], 2.0)
    assert shifted["pressure"][1] - force["pressure"][1] == pytest.approx(-0.2, abs=1e-6)


def test_symmetry_constrains_normal_velocity_in_a_closed_box():
    """tests/fv/test_steady.cpp:9-20: one cell bounded by symmetry planes,
    started with u = (1, 0.4, -0.2): every velocity component is normal to a
    plane, so after 100 iterations the fluid is at rest to 1e-5."""
    s = make((1, 1, 1), {name: Symmetry() for name in PATCHES}, SolverSettings(viscosity=0.1), (1.0, 1.0, 1.0))
    initialize(s, (1.0, 0.4, -0.2), 0.0)
    s.iterate(100)
    assert s.report()[18] < 1e-5  # maximum velocity


@pytest.mark.parametrize("direction", [None, (1.0, 0.0, 0.0), (2.0, 0.0, 0.0)])
def test_mass_flow_inlet_gives_the_uniform_stream(direction):
    """fv/boundary_physics.cpp:14-26: 0.2 kg/s at density 2 through the 1 x 0.1
    xmin patch enters at 0.2 / (2 x 0.1) = 1 along the inward normal (or the
    given direction, of any length); with an outlet downstream the uniform
    stream u = 1 is steady (as tests/fv/test_steady.cpp:22-47)."""
    from kutta_amd import MassFlowInlet
    c = through_flow(1)
    c["xmin"] = MassFlowInlet(0.2, direction)
    s = make((8, 4, 1), c, SolverSettings(density=2.0))
    initialize(s, (1.0, 0.0, 0.0), 0.0)
    s.iterate(6)
    assert approx(s.field("u"), 1.0, 1e-5)
    assert np.abs(s.field("v")).max() < 1e-6 and np.abs(s.field("p")).max() < 1e-6


def test_mass_flow_inlet_validation():
    from kutta_amd import MassFlowInlet
    for bad in (MassFlowInlet(-1.0), MassFlowInlet(0.2, (-1.0, 0.0, 0.0))):
        c = through_flow(1)
        c["xmin"] = bad
        with pytest.raises(oracle.OracleError):
            make((4, 2, 1), c)


def test_profile_errors_reach_the_caller():
    """A boundary profile that raises, or returns the wrong number of
    components, is reported to the caller with its own exception; the face it
    was evaluated for gets NaN, which the solver's finiteness check rejects."""
    for bad, kind in ((lambda x: 1 / 0, ZeroDivisionError), (lambda x: (1.0, 0.0), ValueError)):
        c = couette()
        c["xmin"] = VelocityInlet(bad)
        with pytest.raises(kind):
            make_profiled((4, 12, 1), c, SolverSettings(viscosity=1.0))
