"""Pins the oracle's SST k-omega SIMPLEC step (oracle/simplec.cpp, restating
fv/steady_solver.cpp:484-542, 638-837 and fv/transport.cpp:57-159):

* the reference's own SST case (tests/fv/test_rans.cpp:10-51): a 4x3x2 hex
  channel with a k-omega inlet, a total-pressure outlet with an intensity and
  viscosity-ratio backflow condition, a symmetry plane and walls; its
  assertions (finite fields, positive k and omega residuals, positive wall
  drag, a monitor that leaves omega unchanged, an initial flux that does not
  depend on the relaxation settings) hold here;
* closed forms of homogeneous decay on one cell, where the strain, the
  gradients and the face fluxes vanish: the pseudo-time implicit Euler step
  with Patankar under-relaxation (relaxation alpha, step dt) of
  dk/dt = -beta* k omega, with the sink k * beta* omega implicit, and of
  domega/dt = -beta omega^2, linearized by Newton about the old omega, k
  before omega in each iteration. beta* = 0.09 and beta = beta2 = 0.0828
  (Menter 1994: F1 = 0 far from any wall) make these
  k' = k (1 - alpha beta* w / (beta* w + 1/dt)) and
  w' = w (1 - alpha beta w / (2 beta w + 1/dt)); without the pseudo-time
  term k' = (1 - alpha) k and w' = (1 - alpha/2) w.
"""
import numpy as np
import pytest

import kutta_amd as ka
import oracle
from kutta_amd import meshes

WALLS = ("ymin", "ymax", "zmax")


def rans_case():
    """tests/fv/test_rans.cpp:11-18."""
    mesh = meshes.box_hex(4, 3, 2, hi=(1.0, 0.4, 0.3))
    cond = {name: ka.Wall() for name in WALLS}
    cond["xmin"] = ka.VelocityInlet((1.0, 0.0, 0.0), ka.TurbulenceValues(0.01, 4.0))
    cond["xmax"] = ka.PressureOutlet(0.0, None, True, ka.IntensityViscosityRatio(0.05, 4.0))
    cond["zmin"] = ka.Symmetry()
    return mesh, cond


def make(mesh, cond, settings, sst):
    return oracle.SteadySolver.from_mesh(mesh, ka.pack_mesh_conditions(mesh, cond), settings.packed(),
                                         sst.packed() if sst is not None else None)


def rans_fields(n):
    return (np.full(n, 0.8), np.full(n, 0.01), np.full(n, 0.02), np.full(n, 0.01), np.full(n, 0.01), np.full(n, 4.0))


def report(solver, monitor=False):
    return dict(zip(ka.REPORT_FIELDS, solver.report(monitor)))


def test_rans_case_assertions():
    mesh, cond = rans_case()
    settings = ka.SolverSettings(viscosity=0.001, monitor_interval=10)
    s = make(mesh, cond, settings, ka.SstSettings())
    s.initialize_fields(*rans_fields(mesh.n_cells))
    flux = s.field("mass")
    other = make(mesh, cond, ka.SolverSettings(viscosity=0.001, monitor_interval=10, momentum_relaxation=0.9,
                                               pressure_relaxation=0.2, pseudo_time_scale=13), ka.SstSettings())
    other.initialize_fields(*rans_fields(mesh.n_cells))
    assert np.array_equal(other.field("mass"), flux)
    before = s.field("omega")
    s.report(monitor=True)
    assert np.array_equal(s.field("omega"), before)  # monitoring fixes no omega rows in the state
    s.iterate(3)
    r = report(s, monitor=True)
    assert np.isfinite(r["k_scaled"]) and np.isfinite(r["omega_scaled"])
    assert r["k_raw"] > 0 and r["omega_raw"] > 0
    assert r["k_iterations"] > 0 and r["omega_iterations"] > 0
    f = s.forces(list(WALLS))
    assert f["viscous"][0] > 0
    for name in ("u", "v", "w", "p", "k", "omega", "mu_t"):
        assert np.all(np.isfinite(s.field(name))), name
    assert np.all(s.field("k") > 0) and np.all(s.field("omega") > 0)


def test_rans_case_rejects_bad_turbulence_input():
    mesh, cond = rans_case()
    settings = ka.SolverSettings(viscosity=0.001)
    missing = dict(cond)
    missing["xmin"] = ka.VelocityInlet((1.0, 0.0, 0.0))
    with pytest.raises(oracle.OracleError, match="missing turbulence specification"):
        make(mesh, missing, settings, ka.SstSettings())
    s = make(mesh, cond, settings, ka.SstSettings())
    u, v, w, p, k, om = rans_fields(mesh.n_cells)
    k[0] = 1e100
    with pytest.raises(oracle.OracleError, match="exceeds float storage"):
        s.initialize_fields(u, v, w, p, k, om)
    with pytest.raises(oracle.OracleError, match="initialize_fields"):
        s.initialize(u, v, w, p)


def one_cell(pseudo_transient, dt, alpha):
    mesh = meshes.box_hex(1, 1, 1)
    cond = {name: ka.Symmetry() for name, _, _ in mesh.patches}
    settings = ka.SolverSettings(viscosity=1e-3, pseudo_time_step=dt, monitor_interval=1000)
    sst = ka.SstSettings(k_relaxation=alpha, omega_relaxation=alpha, pseudo_transient=pseudo_transient,
                         linear=(1e-12, 1e-30, 10))
    return make(mesh, cond, settings, sst)


@pytest.mark.parametrize("pseudo_transient,dt,alpha", [(True, 0.05, 0.75), (True, 2.0, 0.5), (False, 0.1, 0.75)])
def test_homogeneous_decay_closed_form(pseudo_transient, dt, alpha):
    s = one_cell(pseudo_transient, dt, alpha)
    k, w = 0.3, 7.0
    s.initialize_fields([0.0], [0.0], [0.0], [0.0], [k], [w])
    beta_star, beta = 0.09, 0.0828
    for it in range(12):
        s.iterate(1)
        inv_dt = 1.0 / dt if pseudo_transient else 0.0
        w_before = w
        k = k * (1 - alpha * beta_star * w / (beta_star * w + inv_dt))
        w = w * (1 - alpha * beta * w / (2 * beta * w + inv_dt))
        got_k, got_w = float(s.field("k")[0]), float(s.field("omega")[0])
        assert got_k == pytest.approx(k, rel=2e-5), it
        assert got_w == pytest.approx(w, rel=2e-5), it
    assert np.all(s.field("u") == 0)
    # The model last ran between the k and omega solves (new k, old omega):
    # mu_t = rho a1 k / max(a1 omega, S F2) = k / omega at zero strain.
    assert float(s.field("mu_t")[0]) == pytest.approx(k / w_before, rel=2e-5)


def test_backward_step_mesh_closed_forms():
    """The step generator: cell and patch counts and the total volume follow
    from the node lines; the oracle's mesh validation (closure, outward
    boundary normals) accepts it."""
    g = meshes.StepGeometry(h=0.1, upstream=2.0, downstream=6.0, channel=2.0, span=1.0)
    m = meshes.backward_step(g, first=0.004, growth=1.3, largest=0.05, nz=3)
    x = np.unique(np.round(m.points[:, 0], 12))
    y = np.unique(np.round(m.points[:, 1], 12))
    nx, ny, nz = len(x) - 1, len(y) - 1, 3
    nxu = int(np.count_nonzero(x < 0))
    nyl = int(np.count_nonzero(y < 0))
    assert m.n_cells == (nx * ny - nxu * nyl) * nz
    sizes = {name: size for name, _, size in m.patches}
    assert sizes["inlet"] == (ny - nyl) * nz and sizes["outlet"] == ny * nz
    assert sizes["symmetry"] == 2 * m.n_cells // nz
    assert sizes["wall"] == (2 * nx + nyl) * nz  # floors, top wall, step face
    cond = {"inlet": ka.VelocityInlet((1.0, 0.0, 0.0), ka.TurbulenceValues(1e-3, 10.0)),
            "outlet": ka.PressureOutlet(0.0, None, False, ka.TurbulenceValues(1e-3, 10.0)),
            "wall": ka.Wall(), "symmetry": ka.Symmetry()}
    s = make(m, cond, ka.SolverSettings(viscosity=1e-3), ka.SstSettings())
    volume = s.field("volume").astype(np.float64).sum()
    expected = (g.upstream * g.h * g.channel * g.h + g.downstream * g.h * (g.channel + 1) * g.h) * g.span * g.h
    assert volume == pytest.approx(expected, rel=1e-6)
