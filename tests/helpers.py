"""Shared generators for the parity tests: seeded CSR operators and vectors."""
import numpy as np

import oracle


def csr_arrays(a):
    return a.row_ptr, a.col, a.diag, a.val.astype(np.float32)


def random_operator(seed=1, nx=9, ny=8, nz=7, neumann=False):
    """A random-coefficient 7-point operator (tests/linalg/matrices.hpp:17-62,
    random=true) with seeded perturbations of the values so that rows differ."""
    a = oracle.poisson(nx, ny, nz, 1.0, random=True, neumann=neumann)
    return a


def random_vector(n, seed, scale=1.0, dtype=np.float32):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(n) * scale).astype(dtype)


def mixed_magnitude_vector(n, seed):
    """Values spanning ~12 decades with both signs: exercises compensation."""
    rng = np.random.default_rng(seed)
    mant = rng.uniform(-1, 1, n)
    expo = rng.integers(-6, 6, n)
    return (mant * 10.0 ** expo).astype(np.float32)
