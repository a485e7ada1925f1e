"""SST k-omega SIMPLEC on the MI355X against the oracle
(fv/steady_solver.cpp:484-542, 638-837; fv/transport.cpp:57-159).

The SST kernels evaluate tanh, log and pow, whose device results differ from
the host libm's by a few ulps, so the transported fields are compared within
the reference's own backend tolerance for this case
(tests/fv/test_rans.cpp:50: |a - b| <= 3e-4 (0.01 + |b|)); the mesh setup,
the fits and the integer model fields are compared bit for bit, and the wall
force within 0.1%."""
import numpy as np
import pytest

import kutta_amd as ka
import oracle
from kutta_amd import meshes

pytestmark = pytest.mark.gpu

WALLS = ("ymin", "ymax", "zmax")
FIELDS = ("u", "v", "w", "p", "k", "omega", "mu_t", "mass")
SETUP = ("volume", "weight", "owner_to_face_x", "wall_distance", "wall_normal_distance", "velocity_lsq_inverse")


def rans_case():
    """tests/fv/test_rans.cpp:11-18."""
    mesh = meshes.box_hex(4, 3, 2, hi=(1.0, 0.4, 0.3))
    cond = {name: ka.Wall() for name in WALLS}
    cond["xmin"] = ka.VelocityInlet((1.0, 0.0, 0.0), ka.TurbulenceValues(0.01, 4.0))
    cond["xmax"] = ka.PressureOutlet(0.0, None, True, ka.IntensityViscosityRatio(0.05, 4.0))
    cond["zmin"] = ka.Symmetry()
    return mesh, cond


def close(a, b, eps=3e-4, scale=0.01):
    return np.abs(a.astype(np.float64) - b) <= eps * (scale + np.maximum(np.abs(a), np.abs(b)))


def assert_fields_close(gpu, cpu, names=FIELDS, eps=3e-4):
    for name in names:
        a, b = gpu.field(name), cpu.field(name)
        assert a.shape == b.shape, name
        ok = close(a, b, eps)
        assert ok.all(), f"{name}: {np.count_nonzero(~ok)} of {a.size} differ, worst " \
                         f"{np.max(np.abs(a.astype(np.float64) - b)):.3e}"


def pair(device, mesh, cond, settings, sst):
    gpu = ka.SteadySolver.from_mesh(device, mesh, cond, settings, sst)
    cpu = oracle.SteadySolver.from_mesh(mesh, ka.pack_mesh_conditions(mesh, cond), settings.packed(),
                                        sst.packed() if sst is not None else None)
    return gpu, cpu


def test_rans_case_against_oracle(device):
    mesh, cond = rans_case()
    settings = ka.SolverSettings(viscosity=0.001, monitor_interval=10)
    gpu, cpu = pair(device, mesh, cond, settings, ka.SstSettings())
    for name in SETUP:
        a, b = gpu.field(name), cpu.field(name)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name
    n = mesh.n_cells
    gpu.initialize_fields((0.8, 0.01, 0.02), 0.01, 0.01, 4.0)
    cpu.initialize_fields(*[np.full(n, v) for v in (0.8, 0.01, 0.02, 0.01, 0.01, 4.0)])
    assert_fields_close(gpu, cpu)
    for name in ("fixed_omega_mask", "selected_wall_face"):
        assert np.array_equal(gpu.field(name).view(np.int32), cpu.field(name).view(np.int32)), name
    before = gpu.field("omega")
    gpu.report(monitor=True)
    assert np.array_equal(gpu.field("omega"), before)
    for total in (3, 20):
        gpu.iterate(total - gpu.info()["iteration"])
        cpu.iterate(total - int(cpu.report()[0]))
        print(f"[sst] rans case after {total} iterations", flush=True)
        assert_fields_close(gpu, cpu)
        g = gpu.report(monitor=True)
        c = dict(zip(ka.REPORT_FIELDS, cpu.report(True)))
        assert g["k_raw"] > 0 and g["omega_raw"] > 0
        assert g["k_raw"] == pytest.approx(c["k_raw"], rel=2e-2)
        assert g["omega_raw"] == pytest.approx(c["omega_raw"], rel=2e-2)
        fg, fc = gpu.forces(list(WALLS)), cpu.forces(list(WALLS))
        assert fg["viscous"][0] > 0
        for part in ("pressure", "viscous"):
            for a, b in zip(fg[part], fc[part]):
                assert a == pytest.approx(b, rel=1e-3, abs=1e-3 * abs(fc["viscous"][0])), part


@pytest.mark.parametrize("pseudo_transient,dt,alpha", [(True, 0.05, 0.75), (False, 0.1, 0.75)])
def test_homogeneous_decay_closed_form(device, pseudo_transient, dt, alpha):
    """One cell, symmetry all round: the pseudo-time implicit Euler step with
    Patankar relaxation of dk/dt = -beta* k w and dw/dt = -beta w^2 (Newton
    linearized), beta* = 0.09, beta = beta2 = 0.0828 (Menter 1994)."""
    mesh = meshes.box_hex(1, 1, 1)
    cond = {name: ka.Symmetry() for name, _, _ in mesh.patches}
    settings = ka.SolverSettings(viscosity=1e-3, pseudo_time_step=dt, monitor_interval=1000)
    sst = ka.SstSettings(k_relaxation=alpha, omega_relaxation=alpha, pseudo_transient=pseudo_transient,
                         linear=(1e-12, 1e-30, 10))
    s = ka.SteadySolver.from_mesh(device, mesh, cond, settings, sst)
    k, w = 0.3, 7.0
    s.initialize_fields((0.0, 0.0, 0.0), 0.0, k, w)
    inv_dt = 1.0 / dt if pseudo_transient else 0.0
    for it in range(12):
        s.iterate(1)
        k = k * (1 - alpha * 0.09 * w / (0.09 * w + inv_dt))
        w = w * (1 - alpha * 0.0828 * w / (2 * 0.0828 * w + inv_dt))
        assert float(s.field("k")[0]) == pytest.approx(k, rel=2e-5), it
        assert float(s.field("omega")[0]) == pytest.approx(w, rel=2e-5), it


def test_batched_momentum_with_wall_functions_is_bitwise_sequential(device):
    """Axis-aligned walls: the wall-function terms of one component read the
    others only through products with zero normal components, so solving
    u, v, w together gives the sequential fields bit for bit."""
    mesh, cond = rans_case()
    settings = ka.SolverSettings(viscosity=0.001, monitor_interval=10)
    a = ka.SteadySolver.from_mesh(device, mesh, cond, settings, ka.SstSettings())
    b = ka.SteadySolver.from_mesh(device, mesh, cond, settings, ka.SstSettings())
    assert a.set_batched_momentum(True)
    assert not b.set_batched_momentum(False)
    for s in (a, b):
        s.initialize_fields((0.8, 0.01, 0.02), 0.01, 0.01, 4.0)
    for _ in range(3):
        a.iterate(4)
        b.iterate(4)
        for name in FIELDS:
            assert np.array_equal(a.field(name).view(np.uint32), b.field(name).view(np.uint32)), name


def test_backward_step_against_oracle(device):
    """A coarse backward-facing step (the C3 geometry, h = 0.1 m, first cell
    4 mm, growth 1.3) at 20 m/s, air, 1% intensity with viscosity ratio 10.

    From the uniform start the first iterations are an impulsive transient
    (speeds reach 15x the inflow) that amplifies any rounding difference, so
    the bound is calibrated on the oracle itself: the GPU's distance from the
    oracle, per field and relative to the field's largest value, must not
    exceed the oracle's own distance from a run whose initial k is perturbed
    by one part in a million (about ten float ulps)."""
    g = meshes.StepGeometry(h=0.1, upstream=2.0, downstream=8.0, channel=3.0, span=1.0)
    mesh = meshes.backward_step(g, first=0.004, growth=1.3, largest=0.04, nz=3)
    turb = ka.IntensityViscosityRatio(0.01, 10.0)
    cond = {"inlet": ka.VelocityInlet((20.0, 0.0, 0.0), turb), "outlet": ka.PressureOutlet(0.0, None, False, turb),
            "wall": ka.Wall(), "symmetry": ka.Symmetry()}
    settings = ka.SolverSettings(density=1.225, viscosity=1.7894e-5, monitor_interval=10)
    gpu, cpu = pair(device, mesh, cond, settings, ka.SstSettings())
    nudged = oracle.SteadySolver.from_mesh(mesh, ka.pack_mesh_conditions(mesh, cond), settings.packed(),
                                           ka.SstSettings().packed())
    n = mesh.n_cells
    k0 = 1.5 * (0.01 * 20.0) ** 2
    w0 = 1.225 * k0 / (1.7894e-5 * 10.0)
    gpu.initialize_fields((20.0, 0.0, 0.0), 0.0, k0, w0)
    for solver, k in ((cpu, k0), (nudged, k0 * (1 + 1e-6))):
        solver.initialize_fields(np.full(n, 20.0), np.zeros(n), np.zeros(n), np.zeros(n), np.full(n, k),
                                 np.full(n, w0))
    for total in (5, 20):
        gpu.iterate(total - gpu.info()["iteration"])
        for solver in (cpu, nudged):
            solver.iterate(total - int(solver.report()[0]))
        for name in ("u", "v", "p", "k", "omega"):
            a, b = gpu.field(name).astype(np.float64), cpu.field(name).astype(np.float64)
            c = nudged.field(name).astype(np.float64)
            scale = np.max(np.abs(b))
            worst, sensitivity = np.max(np.abs(a - b)) / scale, np.max(np.abs(c - b)) / scale
            print(f"[sst] step case, {n} cells, iteration {total}, {name}: gpu {worst:.2e}, "
                  f"oracle perturbed 1e-6: {sensitivity:.2e} (of max|b|)", flush=True)
            assert worst <= max(sensitivity, 1e-6), (name, total, worst, sensitivity)
