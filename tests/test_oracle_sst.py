"""The oracle's SstModel::update (turbulence/sst.cpp:101-158, restated in
oracle/simplec.cpp over kernels pinned bit for bit against the reference's
own bodies) checked by what the model fixes in closed form: the rates of a
pure shear, the a1 limit on the eddy viscosity, the wall rows and zero wall k
diffusivity, and the ambient sustain terms."""

import numpy as np

import kutta_amd
import oracle
from kutta_amd import PATCHES, SolverSettings, SstOptions, Symmetry, Wall
from sst_inputs import sst_inputs


def duct(n=(8, 5, 4)):
    c = {name: Wall() for name in PATCHES}
    c["zmin"] = c["zmax"] = Symmetry()
    kinds, velocity, pressure, total = kutta_amd.pack_open_conditions(c)
    s = oracle.SteadySolver(*n, kinds, velocity, SolverSettings().packed(), pressure, total, (2.0, 1.0, 1.0))
    cells = n[0] * n[1] * n[2]
    boundary = 2 * (n[1] * n[2] + n[0] * n[2] + n[0] * n[1])
    return s, cells, boundary


def test_pure_shear_rates_and_viscosity_limit():
    """du/dy = 3 alone: |S| = |Omega| = 3 exactly; the eddy viscosity obeys
    mu_t <= rho k / omega (a1 limiter, alpha* = 1 without low-Re damping)."""
    s, n, nb = duct()
    x = sst_inputs(n, nb, seed=1)
    for k in range(3, 12):
        x[k] = np.zeros(n, np.float32)
    x[4] = np.full(n, 3.0, np.float32)
    s.sst_update(SstOptions().packed(), x)
    assert np.all(s.field("strain") == 3.0) and np.all(s.field("vorticity") == 3.0)
    mu_t = s.field("mu_t").astype(np.float64)
    bound = x[12].astype(np.float64) * x[14] / x[15]
    assert np.all(mu_t >= 0) and np.all(mu_t <= bound * (1 + 1e-5))
    f1, f2 = s.field("f1"), s.field("f2")
    assert np.all((0 <= f1) & (f1 <= 1)) and np.all((0 <= f2) & (f2 <= 1))


def test_wall_rows_and_wall_diffusivity():
    """Cells touching a wall are fixed-omega rows (mask 1, the selected face is
    one of their wall faces), the others are free; at wall faces the k
    diffusivity is zero and the omega diffusivity is the molecular viscosity."""
    s, n, nb = duct()
    x = sst_inputs(n, nb, seed=2)
    s.sst_update(SstOptions().packed(), x)
    mask = s.field("fixed_omega_mask").view(np.int32)
    selected = s.field("selected_wall_face").view(np.int32)
    d = s.field("wall_distance")
    assert set(np.unique(mask)) == {0, 1}
    assert np.all((mask == 1) == (selected >= 0))
    # Cells on a wall sit half a cell from it: nearer than any interior cell.
    assert d[mask == 1].max() < d[mask == 0].min()
    wall_faces = np.flatnonzero(s.field("wall_shear_coefficient") != 0)
    assert set(selected[selected >= 0]) <= set(range(nb))
    fk, fw = s.field("face_diffusivity_k"), s.field("face_diffusivity_omega")
    internal = len(fk) - nb
    assert len(wall_faces) > 0 and np.all(fk[internal + wall_faces] == 0)
    assert np.all(fw[internal + wall_faces] > 0)


def test_ambient_sustain_terms():
    """SstOptions::ambient adds 0.09 rho omega_a k_a to the k source and
    beta(F1) rho omega_a^2 to the omega source (turbulence/sst.kern.h:255-259),
    bit for bit on top of the plain sources."""
    s, n, nb = duct()
    x = sst_inputs(n, nb, seed=3)
    s.sst_update(SstOptions().packed(), x)
    plain = s.field("k_source").copy()
    ka, wa = 1e-3, 20.0
    s.sst_update(SstOptions(ambient=(ka, wa)).packed(), x)
    rho = x[12]
    expect = plain + np.float32(0.09) * rho * np.float32(wa) * np.float32(ka)
    assert np.array_equal(s.field("k_source").view(np.uint32), expect.astype(np.float32).view(np.uint32))
    assert np.all(s.field("omega_source") > 0)


def test_invalid_sst_options_raise():
    s, n, nb = duct()
    bad = SstOptions(kato_launder=True, vorticity_production=True).packed()
    try:
        s.sst_update(bad, sst_inputs(n, nb))
    except oracle.OracleError:
        return
    raise AssertionError("mutually exclusive production options accepted")
