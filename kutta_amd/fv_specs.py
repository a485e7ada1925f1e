"""Parameter roles of the finite-volume kernels, and synthetic inputs built
from them, for the parity tests and the per-kernel benchmark.

Each kernel is listed with its iteration domain and one role per parameter, in
the reference's parameter order (src/kutta/kernels/fv_*.kern.h, sst.kern.h).
build_args() makes every argument from a real box mesh (connectivity) and
seeded random values (everything else), so the oracle, the reference body and
the device kernel all see identical bytes.

Roles:
  owner neighbour kind offsets faces sign   mesh connectivity / boundary kinds 0-4
  F C B        float per face / cell / boundary face, standard normal
  F+ C+ B+     float, uniform in [0.5, 2] (volumes, diagonals, coefficients)
  B01          float 0 or 1 per boundary face (selection masks)
  oF oC oB     float output (pre-filled with normal noise: some are updated in place)
  oBi          int output per boundary face
  adjacent     per cell-face entry: the neighbour cell, or -(b+1) for boundary face b
  E            float per cell-face entry (least-squares row weights)
  C6 oC6       six float planes per cell (the symmetric 3x3 least-squares inverse)
  subset S6    distinct cells (a third of them, shuffled) and six planes of values each
  internal cells count   the number of internal faces / cells / subset cells (scalars)
  f f+ one     float scalar (normal, uniform [0.5, 2]) / int scalar 1
  eps          limiter eps2: negative (Barth-Jespersen) for odd seeds, positive
               (Venkatakrishnan) for even seeds
  entry rows cols diag   the pressure stencil's cell-face entry map and CSR arrays
  ref          a pinned cell (int scalar): the middle cell
  bmask spec   int per boundary face: 0 or 1 / a turbulence specification 0-2
  component    int scalar: the velocity component (or another 0-2 mode), seed % 3
  Cu           float per cell, uniform in [0, 1] (blending factors)
  oCi          int output per cell
  N oN         float per matrix nonzero (input / output)
  mask         int 0 or 1 per cell
  flag flag2   int switches: seed % 2 and (seed // 2) % 2, so consecutive seeds
               cover both sides of each
"""
from __future__ import annotations

import numpy as np

SPECS = {
    # kernels/fv_flow.kern.h
    "fv_velocity_boundary": ("boundary", "owner kind F F F C C C B B B F oB oB oB oBi internal f+"),
    "fv_pressure_boundary": ("boundary", "owner kind F F F C C C C B oB internal"),
    "fv_pressure_faces": ("faces", "owner neighbour F F F F F F C C C C B oF oF internal"),
    "fv_momentum_source": ("cells", "offsets faces sign F F F F C+ C C C oC oC oC"),
    "fv_momentum_faces": ("faces", "owner neighbour kind F F F F F F C C C C C C C C C C C C C+ "
                                   "B B B oF oF oF internal"),
    "fv_convective_outflow": ("cells", "offsets faces sign F oC"),
    "fv_peclet_side_velocity": ("internal", "owner neighbour F F F F F F C C C C C+ C+ C+ oF"),
    "fv_rhie_chow": ("faces", "owner neighbour kind F F F F+ C+ C+ C+ F F F F oF oF internal f+"),
    "fv_correct_flux": ("faces", "owner neighbour C F oF internal"),
    "fv_correct_velocity": ("cells", "C+ C+ C+ C C C C oC oC oC oC f+"),
    "fv_monitor_cell": ("cells", "offsets faces sign F C C C C C C C oC oC oC oC oC oC"),
    "fv_speed": ("cells", "C C C oC"),
    "fv_continuity": ("cells", "offsets faces sign F oC"),
    "fv_wall_pressure_force": ("boundary", "B01 F F F F oB oB oB internal f"),
    "fv_wall_viscous_force": ("boundary", "owner B01 F F F F F F C C C C C C C C C C C C B B B "
                                          "oB oB oB internal f+"),
    "fv_fluent_momentum_scale": ("cells", "C+ C+ C+ C+ C+ oC oC oC f+ f+"),
    "fv_physical_time_terms": ("cells", "C+ C C oC oC f+ f f f one"),
    "fv_face_time_defect": ("faces", "F F F F F F F oF f+"),
    "fv_face_time_flux": ("faces", "owner neighbour kind C+ C+ F F oF internal f+ f f"),
    "fv_cell_mass_error": ("cells", "C C+ oC f+"),
    # kernels/fv_gradient.kern.h
    "fv_gradient": ("cells", "offsets adjacent E E E B B B C6 C B oC oC oC cells"),
    "fv_paired_gradient": ("cells", "offsets adjacent E E E B B B C6 C C B B oC oC oC oC oC oC cells"),
    "fv_range_shift": ("cells", "C oC oC f+"),
    "fv_limiter": ("cells", "offsets adjacent faces sign F F F F F F C B C C C oC eps"),
    "fv_paired_limiter": ("cells", "offsets adjacent faces sign F F F F F F C C B kind C C C oC eps"),
    "fv_paired_gradient_limiter": ("cells", "offsets adjacent faces sign F F F F F F C C B B "
                                            "C C C C C C oC eps"),
    "fv_vector_gradient": ("cells", "offsets adjacent E E E B B B C6 C C C B B B "
                                    "oC oC oC oC oC oC oC oC oC cells"),
    "fv_scale_gradient": ("cells", "C C C C oC oC oC"),
    "fv_limiter_scale": ("cells", "offsets adjacent faces sign F F F F F F C B C C C oC oC oC oC eps"),
    "fv_paired_limiter_scale": ("cells", "offsets adjacent faces sign F F F F F F C C B kind "
                                         "C C C oC oC oC oC eps"),
    "fv_paired_gradient_limiter_scale": ("cells", "offsets adjacent faces sign F F F F F F C C B B "
                                                  "C C C C C C oC oC oC oC eps"),
    "fv_jump_gradient": ("cells", "offsets faces adjacent sign E E E C6 F F F F F F C B C C C "
                                  "oC oC oC cells"),
    "fv_face_jump": ("faces", "owner neighbour F F F F F F C B C C C oF oF internal"),
    "fv_gather_jump_gradient": ("cells", "offsets faces sign E E E C6 F F C C C oC oC oC cells"),
    "fv_paired_jump_gradient": ("cells", "offsets faces adjacent sign E E E C6 F F F F F F C C B B "
                                         "C C C C C C oC oC oC cells"),
    "fv_lsq_restore_inverse": ("subset", "subset S6 oC6 count cells"),
    # kernels/fv_transport.kern.h
    "fv_diffusion_correction_bound": ("faces", "owner neighbour kind C B F+ oF internal"),
    "fv_transport_faces": ("faces", "owner neighbour F F F F F F F F F F C C C C F F+ kind oF "
                                    "internal flag flag2"),
    "fv_transport_assemble": ("cells", "offsets faces sign adjacent entry rows diag F+ F F+ kind B F "
                                       "C C+ C C+ C C oN oC oC oC oC oC internal f+ f+ f+ flag"),
    "fv_fixed_rows": ("cells", "rows diag mask C oN oC"),
    "fv_floor": ("cells", "oC f"),
    "fv_scalar_steady_measure": ("cells", "C C C mask oC oC flag"),
    "fv_scalar_set_fixed": ("cells", "mask C oC"),
    "fv_scalar_monitor": ("cells", "C C C mask C+ oC oC oC f+ f+"),
    # kernels/fv_paired.kern.h
    "fv_compensated_pressure_jump": ("faces", "owner neighbour F F F F F F C C C C B oF internal"),
    "fv_paired_pressure_jump": ("faces", "owner neighbour F F F F F F C C C C C C C C B B oF internal"),
    "fv_compensated_rhie_chow": ("faces", "owner neighbour kind F F F F+ C+ C+ F F F F F F F oF oF "
                                          "internal f+"),
    "fv_paired_pressure_boundary_low": ("boundary", "owner kind F F F C C C C C C C C B oB internal"),
    "fv_paired_pressure_faces": ("faces", "owner neighbour kind mask F F F F F F C C C C C C C C B B C+ "
                                          "oF oF internal flag"),
    "fv_paired_momentum_residual": ("cells", "rows cols N diag C C C C C C+ C C oC"),
    "fv_compensated_momentum_residual": ("cells", "rows cols N diag C C C C C C+ C C oC"),
    "fv_paired_accumulate": ("cells", "C oC oC f+"),
    "fv_compensated_pressure_force": ("cells", "offsets faces sign F F F C+ F F oC oC oC"),
    "fv_paired_velocity_boundary": ("boundary", "owner kind F F F C C C F F oB oB oB oBi internal f+"),
    "fv_paired_advection": ("faces", "owner neighbour kind F F F F F F C B C C C C F F oF internal flag"),
    "fv_paired_face_residual": ("cells", "offsets faces sign F F F C oC flag"),
    "fv_paired_flux": ("faces", "owner neighbour kind F F F F F F F F F C+ C+ C+ F+ C C C C C C C F "
                                "oF oF internal f+ f+"),
    "fv_paired_flux_offset": ("faces", "F F F F oF oF f+"),
    "fv_paired_continuity": ("cells", "offsets faces sign F F oC oC"),
    "fv_paired_pressure_rhs": ("cells", "C oC ref"),
    "fv_paired_correct_flux": ("faces", "owner neighbour F F F C F oF oF internal f+"),
    "fv_paired_correct_velocity": ("cells", "C+ C+ C+ C C C C oC oC oC oC oC oC oC oC f+ f+ f+"),
    "fv_paired_value": ("cells", "C C oC"),
    "fv_paired_vector_values": ("cells", "C C C C C C oC oC oC"),
    "fv_compensated_correct_flux": ("faces", "owner neighbour F F F C F oF oF internal f+"),
    "fv_compensated_correct_velocity": ("cells", "C+ C+ C+ C C C C oC oC oC oC oC oC oC oC f+ f+ f+"),
    # kernels/fv_rans.kern.h
    "fv_backflow_pressure": ("boundary", "owner bmask F C C C oB internal f+"),
    "fv_effective_viscosity": ("faces", "owner neighbour kind F C+ oF internal f+"),
    "fv_transpose_stress": ("faces", "owner neighbour kind F F+ F F F C C C C C C C C C oF oF oF internal"),
    "fv_stress_source": ("cells", "offsets faces sign C+ F F F oC oC oC"),
    "fv_boundary_momentum": ("cells", "offsets faces kind F F F B+ F+ F+ C C C B B B oC oC internal "
                                      "component flag"),
    "fv_momentum_diagonal": ("cells", "C C C C C C oC oC"),
    "fv_sst_force": ("boundary", "B01 F F F B B B oB oB oB internal"),
    "fv_turbulence_boundary": ("boundary", "owner kind spec B+ B+ F B B B C+ C+ B+ oB oB oBi internal f+ f+"),
    # kernels/sst.kern.h
    "sst_rates": ("cells", "C C C C C C C C C oC oC"),
    "sst_blending": ("cells", "C+ C+ C+ C+ C+ C C C C C C oC oC oC f+ f+ f+"),
    "sst_viscosity": ("cells", "C+ C+ C+ C+ C+ C+ Cu Cu oC oC oC flag flag2 f+ f+ f+"),
    "sst_blending_viscosity": ("cells", "C+ C+ C+ C+ C+ C C C C C C C+ C+ oC oC oC oC oC oC f+ f+ f+ "
                                        "flag flag2 f+"),
    "sst_wall": ("boundary", "owner bmask internal F F F C C C B B B B+ C+ C+ C+ C+ "
                             "oB oB oB oB oB oB oB oB oB oB f+ f+"),
    "sst_wall_omega": ("boundary", "owner bmask internal B+ C+ C+ B+ oB oB"),
    "sst_wall_cells": ("cells", "offsets faces bmask B+ B B internal oCi oCi oC oC"),
    "sst_sources": ("cells", "C+ C+ C+ C+ C+ C+ Cu C C+ mask C+ oC oC oC oC oC flag flag2 component "
                             "flag flag2 flag f+ flag2 f+ f+ f+ f"),
    "sst_face_diffusion": ("faces", "owner neighbour F bmask C+ C+ C+ internal oF oF"),
    "sst_sustain": ("cells", "C+ Cu oC oC f+ f+"),
    # kernels/fv_fluent_schemes.kern.h
    "fv_simple_inverse": ("cells", "diag N oC"),
    "fv_standard_pressure_faces": ("faces", "owner neighbour kind C C+ oF oB internal"),
    "fv_cell_pressure_boundary": ("boundary", "owner kind C oB internal"),
    "fv_force_to_shear": ("boundary", "B01 F+ F+ F+ B B B oB oB oB internal"),
    "fv_poor_cell_pressure_faces": ("faces", "owner neighbour kind mask F F F F F F C C C C oF oB internal"),
    "fv_poor_pressure_interfaces": ("internal", "owner neighbour mask C oF"),
}

# Kernels whose bodies call transcendental functions (tanh, log, pow): the
# device's math library and the host's round differently in the last bits,
# so the device comparison for these uses a relative tolerance.
TRANSCENDENTAL = {"sst_blending", "sst_blending_viscosity", "sst_wall", "sst_wall_omega"}


class FvMesh:
    """Connectivity of a renumbered, perturbed box mesh: any mapping with the
    arrays owner, neighbour, cell_face_offsets, cell_faces, cell_face_sign,
    adjacent, matrix_entry, row_ptr, diag and col (kutta_amd.MeshSetup or a
    pressure-problem bundle)."""

    def __init__(self, bundle):
        self.owner = np.ascontiguousarray(bundle["owner"], np.int32)
        self.neighbour = np.ascontiguousarray(bundle["neighbour"], np.int32)
        self.offsets = np.ascontiguousarray(bundle["cell_face_offsets"], np.int32)
        self.faces = np.ascontiguousarray(bundle["cell_faces"], np.int32)
        self.sign = np.ascontiguousarray(bundle["cell_face_sign"], np.float32)
        self.adjacent = np.ascontiguousarray(bundle["adjacent"], np.int32)
        self.entry = np.ascontiguousarray(bundle["matrix_entry"], np.int32)
        self.rows = np.ascontiguousarray(bundle["row_ptr"], np.int32)
        self.diag = np.ascontiguousarray(bundle["diag"], np.int32)
        self.cols = np.ascontiguousarray(bundle["col"], np.int32)
        self.nnz = int(self.rows[-1])
        self.cells = self.offsets.size - 1
        self.entries = self.faces.size
        self.subset = self.cells // 3
        self.n_faces = self.owner.size
        self.internal = self.neighbour.size
        self.boundary = self.n_faces - self.internal

    def domain(self, name):
        return {"faces": self.n_faces, "internal": self.internal, "boundary": self.boundary,
                "cells": self.cells, "subset": self.subset}[name]


def build_args(mesh: FvMesh, spec: str, seed: int):
    """Arguments for one kernel call and the indices of its output arrays."""
    rng = np.random.default_rng(seed)
    size = {"F": mesh.n_faces, "C": mesh.cells, "B": mesh.boundary, "E": mesh.entries,
            "C6": 6 * mesh.cells, "S6": 6 * mesh.subset, "N": mesh.nnz}
    args, outputs = [], []
    for role in spec.split():
        if role == "owner":
            args.append(mesh.owner.copy())
        elif role == "neighbour":
            args.append(mesh.neighbour.copy())
        elif role == "kind":
            args.append(rng.integers(0, 5, mesh.boundary).astype(np.int32))
        elif role == "offsets":
            args.append(mesh.offsets.copy())
        elif role == "faces":
            args.append(mesh.faces.copy())
        elif role == "sign":
            args.append(mesh.sign.copy())
        elif role == "adjacent":
            args.append(mesh.adjacent.copy())
        elif role == "entry":
            args.append(mesh.entry.copy())
        elif role == "rows":
            args.append(mesh.rows.copy())
        elif role == "diag":
            args.append(mesh.diag.copy())
        elif role == "cols":
            args.append(mesh.cols.copy())
        elif role == "mask":
            args.append(rng.integers(0, 2, mesh.cells).astype(np.int32))
        elif role == "bmask":
            args.append(rng.integers(0, 2, mesh.boundary).astype(np.int32))
        elif role == "spec":
            args.append(rng.integers(0, 3, mesh.boundary).astype(np.int32))
        elif role == "subset":
            args.append(rng.permutation(mesh.cells)[:mesh.subset].astype(np.int32))
        elif role in ("F", "C", "B", "E", "C6", "S6", "N"):
            args.append(rng.standard_normal(size[role]).astype(np.float32))
        elif role == "Cu":
            args.append(rng.uniform(0.0, 1.0, mesh.cells).astype(np.float32))
        elif role == "oCi":
            outputs.append(len(args))
            args.append(np.full(mesh.cells, -7, np.int32))
        elif role in ("F+", "C+", "B+"):
            args.append(rng.uniform(0.5, 2.0, size[role[0]]).astype(np.float32))
        elif role == "B01":
            args.append(rng.integers(0, 2, mesh.boundary).astype(np.float32))
        elif role in ("oF", "oC", "oB", "oC6", "oN"):
            outputs.append(len(args))
            args.append(rng.standard_normal(size[role[1:]]).astype(np.float32))
        elif role == "oBi":
            outputs.append(len(args))
            args.append(np.full(mesh.boundary, -7, np.int32))
        elif role == "internal":
            args.append(int(mesh.internal))
        elif role == "cells":
            args.append(int(mesh.cells))
        elif role == "count":
            args.append(int(mesh.subset))
        elif role == "ref":
            args.append(int(mesh.cells // 2))
        elif role == "component":
            args.append(seed % 3)
        elif role == "eps":
            args.append(-1.0 if seed % 2 else float(np.float32(rng.uniform(0.01, 0.5))))
        elif role == "f":
            args.append(float(np.float32(rng.standard_normal())))
        elif role == "f+":
            args.append(float(np.float32(rng.uniform(0.5, 2.0))))
        elif role == "one":
            args.append(1)
        elif role == "flag":
            args.append(seed % 2)
        elif role == "flag2":
            args.append((seed // 2) % 2)
        else:
            raise ValueError(f"unknown role {role}")
    return args, outputs


def signature_of(spec: str) -> str:
    """The parameter kinds a spec implies ('b' buffer, 'f' float, 'i' int)."""
    out = []
    for role in spec.split():
        if role in ("internal", "one", "cells", "count", "flag", "flag2", "ref", "component"):
            out.append("i")
        elif role in ("f", "f+", "eps"):
            out.append("f")
        else:
            out.append("b")
    return "".join(out)


def same_bits(a: np.ndarray, b: np.ndarray) -> bool:
    """Bitwise equality, except that any two NaNs compare equal (a NaN's sign
    and payload depend on the hardware that produced it)."""
    if a.dtype != b.dtype or a.shape != b.shape:
        return False
    if a.dtype == np.float32:
        both_nan = np.isnan(a) & np.isnan(b)
        return bool(np.all(both_nan | (a.view(np.uint32) == b.view(np.uint32))))
    return bool(np.array_equal(a, b))
