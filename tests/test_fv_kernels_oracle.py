"""Pins the oracle's finite-volume kernel restatements (oracle/fv_kernels.cpp)
bit for bit against the reference's own bodies (src/kutta/kernels/fv_*.kern.h,
compiled into oracle/_ref), on seeded inputs over real box meshes."""
import numpy as np
import pytest

import oracle
from kutta_amd.fv_specs import SPECS, FvMesh, build_args, same_bits, signature_of

needs_ref = pytest.mark.skipif(not oracle.ref_built("ref_kernels"),
                               reason="oracle/_ref not built (reference tree absent)")


@pytest.fixture(scope="module", params=[False, True], ids=["hex", "tet"])
def mesh(request):
    return FvMesh(oracle.pressure_problem(6, 5, 4, tet=request.param, amplitude=0.2))


def test_oracle_covers_every_spec():
    assert sorted(oracle.fv_names()) == sorted(SPECS)
    for name, (_, spec) in SPECS.items():
        assert oracle.fv_signature(name) == signature_of(spec), name


@needs_ref
@pytest.mark.parametrize("name", sorted(SPECS))
def test_fv_kernel_bitwise_vs_reference(mesh, name):
    domain, spec = SPECS[name]
    n = mesh.domain(domain)
    for seed in (1, 2, 3, 4):
        args, outputs = build_args(mesh, spec, seed)
        mine = [a.copy() if isinstance(a, np.ndarray) else a for a in args]
        theirs = [a.copy() if isinstance(a, np.ndarray) else a for a in args]
        oracle.fv_each(name, n, *mine)
        oracle.ref_fv_each(name, n, *theirs)
        for k in outputs:
            assert same_bits(mine[k], theirs[k]), (name, seed, k)
        # The outputs changed, so the comparison is not vacuous.
        assert any(not np.array_equal(mine[k], args[k]) for k in outputs), name
