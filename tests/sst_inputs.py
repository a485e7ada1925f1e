"""Seeded SstModel::update inputs for a box solver (shared by the oracle and
device tests): every cell gets a velocity, velocity gradients, density,
viscosity spanning the viscous sublayer to the log layer, k and omega over
several decades and their gradients; walls are at rest."""
import numpy as np


def sst_inputs(cells, boundary, seed=0):
    rng = np.random.default_rng(seed)
    n = cells
    f = lambda a: np.ascontiguousarray(a, np.float32)  # noqa: E731
    velocity = [f(1 + 0.3 * rng.standard_normal(n)), f(0.3 * rng.standard_normal(n)), f(0.3 * rng.standard_normal(n))]
    gradients = [f(5 * rng.standard_normal(n)) for _ in range(9)]
    density = f(1.2 * (1 + 0.1 * rng.random(n)))
    viscosity = f(10 ** rng.uniform(-5, -1, n))
    k = f(10 ** rng.uniform(-5, -1, n))
    omega = f(10 ** rng.uniform(0, 4, n))
    grad_k = [f(0.1 * rng.standard_normal(n)) for _ in range(3)]
    grad_omega = [f(100 * rng.standard_normal(n)) for _ in range(3)]
    wall = [f(np.zeros(boundary)) for _ in range(3)]
    return velocity + gradients + [density, viscosity, k, omega] + grad_k + grad_omega + wall
