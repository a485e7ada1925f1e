"""The MI355X SIMPLEC with velocity inlets and pressure outlets against the
oracle (bitwise, every field, through the C ABI), and the closed forms of the
reference's own tests (tests/fv/test_steady.cpp:22-47, 84-102, 104-126,
131-157) on the device path."""
import numpy as np
import pytest

import kutta_amd
import oracle
from kutta_amd import PATCHES, PressureOutlet, SolverSettings, Symmetry, VelocityInlet, Wall
from test_gpu_simplec import assert_same_report, assert_same_state

pytestmark = pytest.mark.gpu


def approx(a, b, eps):
    """doctest::Approx(b).epsilon(eps) == a, elementwise."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.all(np.abs(a - b) < eps * (1 + np.maximum(np.abs(a), np.abs(b))))


def make(device, n, conditions, settings, extent):
    kinds, velocity, pressure, total = kutta_amd.pack_open_conditions(conditions)
    gpu = kutta_amd.SteadySolver(device, *n, conditions, settings, extent)
    cpu = oracle.SteadySolver(*n, kinds, velocity, settings.packed(), pressure, total, extent)
    return gpu, cpu


def initialize(gpu, cpu, velocity, pressure):
    gpu.initialize(velocity, pressure)
    centres = gpu.centroids()
    assert np.array_equal(centres, cpu.centroids())
    cpu.initialize(*kutta_amd.initial_fields(centres, velocity, pressure))


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


def channel(outlet):
    c = {name: Symmetry() if name in ("zmin", "zmax") else Wall() for name in PATCHES}
    c["xmin"] = VelocityInlet((1.0, 0.0, 0.0))
    c["xmax"] = outlet
    return c


def duct(outlet):
    """A square duct: inlet at xmin, the given outlet at xmax, walls around."""
    c = {name: Wall() for name in PATCHES}
    c["xmin"] = VelocityInlet((1.0, 0.0, 0.0))
    c["xmax"] = outlet
    return c


@pytest.mark.parametrize("through", [1, -1])
def test_uniform_through_flow_is_preserved_and_matches_the_oracle(device, through):
    print(f"[open] through-flow {through}", flush=True)
    gpu, cpu = make(device, (8, 4, 1), through_flow(through), SolverSettings(), (1.0, 1.0, 0.1))
    assert gpu.info()["reference"] == -1 and cpu.reference() == -1
    assert np.all(np.isinf(gpu.field("wall_distance")))  # no wall patch
    initialize(gpu, cpu, (through, 0.0, 0.0), 0.0)
    for s in (gpu, cpu):
        s.iterate(5)
    g, o = gpu.report(monitor=True), cpu.report(monitor=True)
    assert g["continuity_raw"] < 1e-6
    for s in (gpu, cpu):
        s.iterate(1)
    assert gpu.report()["iteration"] == 6
    assert_same_state(gpu, cpu, True)
    assert_same_report(gpu, cpu)
    assert approx(gpu.field("u"), through, 1e-5)
    assert np.abs(gpu.field("v")).max() < 1e-6 and np.abs(gpu.field("p")).max() < 1e-6


def test_outlet_normal_backflow_matches_the_oracle(device):
    print("[open] normal backflow", flush=True)
    c = {name: Symmetry() for name in PATCHES}
    c["xmin"] = VelocityInlet((-1.0, 0.0, 0.0))
    c["xmax"] = PressureOutlet()
    gpu, cpu = make(device, (4, 3, 1), c, SolverSettings(density=2.0), (1.0, 1.0, 0.1))
    initialize(gpu, cpu, (-1.0, 0.0, 0.0), 0.0)
    for s in (gpu, cpu):
        s.iterate(10)
    assert_same_state(gpu, cpu, True)
    assert approx(gpu.field("u"), -1.0, 1e-6)
    assert np.abs(gpu.field("v")).max() < 1e-6 and np.abs(gpu.field("p")).max() < 1e-6


def test_gauge_pressure_initialization_matches_the_oracle(device):
    print("[open] gauge", flush=True)
    c = {name: Wall() for name in PATCHES}
    c["ymax"] = Wall((1.0, 0.0, 0.0))
    c["zmin"] = c["zmax"] = Symmetry()
    gpu, cpu = make(device, (6, 6, 1), c, SolverSettings(reference_point=(0.75, 0.25, 0.05)), (1.0, 1.0, 0.1))
    initialize(gpu, cpu, (0.0, 0.0, 0.0), lambda x: x[0] + 2 * x[1])
    assert gpu.info()["reference"] == cpu.reference() >= 0
    for s in (gpu, cpu):
        s.iterate(10)
    assert_same_state(gpu, cpu, True)
    p = gpu.field("p")
    assert p[cpu.reference()] == 0.0 and np.abs(p).max() > 1e-3


@pytest.mark.parametrize("paired", [False, True])
def test_channel_matches_the_oracle_and_reaches_the_pair_floor(device, paired):
    """200 iterations of the 24 x 8 channel on [0,3] x [0,1] x [0,0.1], bitwise
    every 50; the pair's continuity falls below 1e-7, float32's stays above
    1e-6 (tests/fv/test_steady.cpp:131-157)."""
    print(f"[open] channel paired={paired}", flush=True)
    settings = SolverSettings(viscosity=0.02, paired_state=paired)
    gpu, cpu = make(device, (24, 8, 1), channel(PressureOutlet(0.0, (1.0, 0.0, 0.0))), settings, (3.0, 1.0, 0.1))
    initialize(gpu, cpu, (1.0, 0.0, 0.0), 0.0)
    for _ in range(4):
        for s in (gpu, cpu):
            s.iterate(50)
        assert_same_state(gpu, cpu, paired)
        assert_same_report(gpu, cpu)
    continuity = gpu.report(monitor=True)["continuity_scaled"]
    assert continuity < 1e-7 if paired else continuity > 1e-6


@pytest.mark.parametrize("outlet", [
    PressureOutlet(),                                   # normal backflow, static
    PressureOutlet(0.5, (0.3, 0.0, 0.0), True),         # backflow velocity, total pressure
])
def test_developing_duct_matches_the_oracle(device, outlet):
    """A 3-D duct starting from rest: the inlet drives a developing flow whose
    outlet boundary terms (backflow, prescribed pressure, the pressure speed
    scale) all act; 20 iterations bitwise, with the batched momentum solve."""
    print(f"[open] duct {outlet}", flush=True)
    gpu, cpu = make(device, (14, 6, 5), duct(outlet), SolverSettings(viscosity=0.02), (2.0, 1.0, 1.0))
    assert gpu.set_batched_momentum(True)
    initialize(gpu, cpu, (0.0, 0.0, 0.0), 0.0)
    for s in (gpu, cpu):
        s.iterate(20)
    assert_same_state(gpu, cpu, True)
    assert_same_report(gpu, cpu)
    wd = gpu.field("wall_distance")
    assert np.array_equal(wd.view(np.uint32), cpu.field("wall_distance").view(np.uint32))
    assert 0 < wd.min() and wd.max() <= 0.5  # the duct's half width
    u = gpu.field("u")
    # Equal cells across a unit section: their mean streamwise velocity is
    # the inlet's, while the core speeds up between the no-slip walls.
    assert abs(u.mean() - 1.0) < 1e-3 and u.max() > 1.5


def test_open_conditions_are_validated(device):
    c = through_flow(1)
    c["xmax"] = PressureOutlet(float("inf"))
    with pytest.raises(kutta_amd.KuttaError):
        kutta_amd.SteadySolver(device, 4, 2, 1, c)
    with pytest.raises(kutta_amd.KuttaError):
        kutta_amd.SteadySolver(device, 4, 2, 1, through_flow(1), extent=(1.0, -1.0, 1.0))


def test_profiled_inlets_couette_matches_the_oracle(device):
    """Inlets with the linear profile u = y (a function of the face centroid,
    called back through the C ABI) between a moving lid and a wall at rest
    (tests/fv/test_steady.cpp:49-82): bitwise with the oracle after 100
    iterations, lid forces equal, and the steady profile reached."""
    print("[open] couette with profiled inlets", flush=True)
    c = {name: Symmetry() for name in PATCHES}
    c["ymax"] = Wall((1.0, 4.0, 0.0))
    c["ymin"] = Wall()
    c["xmin"] = c["xmax"] = VelocityInlet(lambda x: (x[1], 0.0, 0.0))
    settings = SolverSettings(viscosity=1.0)
    kinds, velocity, pressure, total = kutta_amd.pack_open_conditions(c)
    profiled, callback = kutta_amd.profile_callback(c)
    extent = (1.0, 1.0, 0.1)
    gpu = kutta_amd.SteadySolver(device, 4, 12, 1, c, settings, extent)
    cpu = oracle.SteadySolver(4, 12, 1, kinds, velocity, settings.packed(), pressure, total, extent, profiled,
                              callback)
    for s in (gpu, cpu):
        s.iterate(100)
    assert_same_state(gpu, cpu, True)
    assert_same_report(gpu, cpu)
    g, o = gpu.forces(["ymax"]), cpu.forces(["ymax"])
    assert g["viscous"][0] == pytest.approx(o["viscous"][0], rel=1e-12)
    assert np.abs(gpu.field("u") - gpu.centroids()[:, 1]).max() < 2e-4


def test_symmetry_box_comes_to_rest_like_the_oracle(device):
    """tests/fv/test_steady.cpp:9-20 on the device: a one-cell symmetry box
    started in motion is at rest after 100 iterations, bitwise the oracle."""
    c = {name: Symmetry() for name in PATCHES}
    gpu, cpu = make(device, (1, 1, 1), c, SolverSettings(viscosity=0.1), (1.0, 1.0, 1.0))
    initialize(gpu, cpu, (1.0, 0.4, -0.2), 0.0)
    for s in (gpu, cpu):
        s.iterate(100)
    assert_same_state(gpu, cpu, True)
    assert gpu.report()["maximum_velocity"] < 1e-5


@pytest.mark.parametrize("direction", [None, (1.0, 0.2, 0.0)])
def test_mass_flow_inlet_matches_the_oracle(device, direction):
    """A mass-flow inlet (0.3 kg/s, density 1.5) feeding the duct from rest,
    along the normal or an oblique direction: 20 iterations bitwise."""
    from kutta_amd import MassFlowInlet
    print(f"[open] mass-flow inlet {direction}", flush=True)
    c = duct(PressureOutlet())
    c["xmin"] = MassFlowInlet(0.3, direction)
    gpu, cpu = make(device, (10, 5, 4), c, SolverSettings(density=1.5, viscosity=0.02), (2.0, 1.0, 1.0))
    initialize(gpu, cpu, (0.0, 0.0, 0.0), 0.0)
    for s in (gpu, cpu):
        s.iterate(20)
    assert_same_state(gpu, cpu, True)
    assert_same_report(gpu, cpu)
    assert abs(gpu.field("u").mean() - 0.3 / 1.5) < 1e-3  # equal cells across a unit section
