"""kutta_amd — MI355X (gfx950) HIP backend for Kutta's SIMPLEC pressure path.

Python view of the two native libraries built in-tree by ``kutta_amd.build``:

* ``libkutta_hip.so`` — the device ABI (include/kutta_hip.h): buffers, ordered
  launches of the hand-written HIP kernels, FP64 reductions, graphs.
* ``libkutta_amd.so`` — C++ host code (include/kutta_amd.h): aggregation AMG,
  the device-resident Krylov solvers, mesh and pressure-operator setup.

Names mirror the reference's C++ API (``Device``, ``make_amg_solver``,
``SolveControl``, ``SolveStats``, ``AmgOptions`` …). There is no fallback: if
the native libraries are missing or no HIP device is present, construction
raises.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_hip = None
_host = None


class KuttaError(RuntimeError):
    """A failure reported by the native libraries (kutta::Error)."""


class NonfiniteError(KuttaError):
    """A diverged / non-finite solve (kutta::NonfiniteError)."""


def _load(name: str) -> C.CDLL:
    path = _HERE / name
    if not path.exists():
        raise KuttaError(f"{path} is not built; run `python -m kutta_amd.build`")
    return C.CDLL(str(path), mode=C.RTLD_GLOBAL)


def hip_lib() -> C.CDLL:
    """The device ABI library, with prototypes declared."""
    global _hip
    if _hip is None:
        lib = _load("libkutta_hip.so")
        vp, i32, sz = C.c_void_p, C.c_int, C.c_size_t
        lib.kh_last_error.restype = C.c_char_p
        lib.kh_device_count.restype = i32
        lib.kh_create.argtypes = [i32, C.POINTER(vp)]
        lib.kh_destroy.argtypes = [vp]
        lib.kh_name.argtypes = [vp]
        lib.kh_name.restype = C.c_char_p
        lib.kh_alloc.argtypes = [vp, sz, C.POINTER(vp)]
        lib.kh_free.argtypes = [vp, vp]
        lib.kh_upload.argtypes = [vp, vp, vp, sz]
        lib.kh_download.argtypes = [vp, vp, vp, sz]
        lib.kh_copy.argtypes = [vp, vp, vp, sz]
        lib.kh_fill32.argtypes = [vp, vp, C.c_uint32, sz]
        lib.kh_synchronize.argtypes = [vp]
        lib.kh_launch.argtypes = [vp, C.c_uint32, sz, C.POINTER(KhArg), sz]
        lib.kh_kernel_name.argtypes = [C.c_uint32]
        lib.kh_kernel_name.restype = C.c_char_p
        lib.kh_reduce_host.argtypes = [vp, C.POINTER(KhReduction), sz, C.POINTER(C.c_double)]
        lib.kh_reduce_into.argtypes = [vp, i32, vp, vp, vp, vp, sz, vp, sz]
        lib.kh_reduce_into_static.argtypes = [vp, i32, vp, vp, vp, vp, sz, vp, sz]
        lib.kh_graph_begin.argtypes = [vp]
        lib.kh_graph_end.argtypes = [vp, C.POINTER(vp)]
        lib.kh_graph_launch.argtypes = [vp, vp]
        lib.kh_graph_destroy.argtypes = [vp]
        lib.kh_profile_enable.argtypes = [vp, C.c_uint32, i32]
        lib.kh_profile_read.argtypes = [vp, C.c_uint32, C.POINTER(C.c_long), C.POINTER(C.c_double)]
        lib.kh_profile_reset.argtypes = [vp]
        lib.kh_time_launches.argtypes = [vp, C.c_uint32, sz, C.POINTER(KhArg), sz, C.c_int,
                                         C.POINTER(C.c_double)]
        lib.kh_logical_bytes.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_long)]
        lib.kh_resident_partition.argtypes = [vp, vp, C.c_int32, vp, C.POINTER(C.c_int32)]
        _hip = lib
    return _hip


def host_lib() -> C.CDLL:
    """The host library (solvers, mesh), with prototypes declared."""
    global _host
    if _host is None:
        hip_lib()
        lib = _load("libkutta_amd.so")
        vp, i32, dbl = C.c_void_p, C.c_int, C.c_double
        pi32 = C.POINTER(C.c_int32)
        lib.ka_last_error.restype = C.c_char_p
        lib.ka_device_create.argtypes = [i32, C.POINTER(vp)]
        lib.ka_device_destroy.argtypes = [vp]
        lib.ka_device_raw.argtypes = [vp]
        lib.ka_device_raw.restype = vp
        lib.ka_csr_create.argtypes = [i32, vp, vp, vp, vp, C.POINTER(vp)]
        lib.ka_csr_destroy.argtypes = [vp]
        lib.ka_csr_rows.argtypes = [vp]
        lib.ka_csr_nnz.argtypes = [vp]
        lib.ka_csr_copy.argtypes = [vp, vp, vp, vp, vp]
        lib.ka_matrix_upload.argtypes = [vp, vp, C.POINTER(vp)]
        lib.ka_matrix_destroy.argtypes = [vp]
        for f in ("ka_matrix_values", "ka_matrix_row_ptr", "ka_matrix_col", "ka_matrix_diag"):
            getattr(lib, f).argtypes = [vp]
            getattr(lib, f).restype = vp
        lib.ka_solver_create.argtypes = [vp, i32, vp, pi32, dbl, C.POINTER(vp)]
        lib.ka_solver_destroy.argtypes = [vp]
        lib.ka_solver_refresh.argtypes = [vp, vp]
        lib.ka_solver_solve.argtypes = [vp, vp, vp, vp, dbl, dbl, i32, i32, C.POINTER(dbl)]
        lib.ka_solver_describe.argtypes = [vp, C.c_char_p, i32]
        lib.ka_solver_plan.argtypes = [vp, C.c_char_p, i32]
        lib.ka_solver_set_plan.argtypes = [vp, i32, i32, i32, i32]
        lib.ka_simplec_pressure_levels.argtypes = [vp, vp, vp, i32]
        lib.ka_simplec_pressure_plan.argtypes = [vp, C.c_char_p, i32]
        lib.ka_solver_solve_batched.argtypes = [vp, vp, vp, vp, i32, dbl, dbl, i32, i32, C.POINTER(dbl)]
        lib.ka_solver_levels.argtypes = [vp, pi32, pi32, i32]
        lib.ka_solver_transfer.argtypes = [vp, i32, i32, pi32]
        lib.ka_solver_transfer.restype = C.c_long
        lib.ka_solver_level_values.argtypes = [vp, i32, C.POINTER(C.c_float)]
        lib.ka_solver_level_values.restype = C.c_long
        lib.ka_solver_set_fused.argtypes = [vp, i32]
        lib.ka_solver_host_waits.argtypes = [vp]
        lib.ka_solver_set_tail.argtypes = [vp, i32, pi32]
        lib.ka_solver_set_restart.argtypes = [vp, i32]
        lib.ka_solver_set_direct.argtypes = [vp, i32, pi32]
        lib.ka_solver_set_dense_tolerance.argtypes = [vp, dbl]
        lib.ka_solver_dense_refreshes.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                                  C.POINTER(dbl)]
        lib.ka_solver_set_coarse_bf16.argtypes = [vp, i32]
        lib.ka_solver_set_graphs.argtypes = [vp, i32]
        lib.ka_solver_autotune_tail.argtypes = [vp, vp, i32, pi32]
        lib.ka_solver_set_learned.argtypes = [vp, i32, i32, C.c_uint64, i32]
        lib.ka_solver_learned_mode.argtypes = [vp, i32, dbl, dbl]
        lib.ka_solver_learned_targets.argtypes = [vp, dbl, pi32]
        lib.ka_solver_learned_train.argtypes = [vp, i32, C.c_float, i32, vp, vp]
        lib.ka_solver_learned_evaluate.argtypes = [vp, i32, i32, C.POINTER(C.c_double)]
        lib.ka_solver_learned_report.argtypes = [vp, C.POINTER(dbl)]
        lib.ka_solver_learned_parameters.argtypes = [vp, vp, C.c_int64, i32]
        lib.ka_aggregate.argtypes = [vp, pi32, dbl, C.POINTER(vp)]
        lib.ka_coarsening_array.argtypes = [vp, i32, pi32]
        lib.ka_coarsening_array.restype = C.c_long
        lib.ka_coarsening_values.argtypes = [vp, C.POINTER(C.c_double)]
        lib.ka_coarsening_values.restype = C.c_long
        lib.ka_coarsening_destroy.argtypes = [vp]
        lib.ka_simplec_create.argtypes = [vp, i32, i32, i32, pi32, C.POINTER(dbl), C.POINTER(dbl),
                                          C.POINTER(vp)]
        lib.ka_simplec_create_open.argtypes = [vp, i32, i32, i32, C.POINTER(dbl), pi32, C.POINTER(dbl),
                                               C.POINTER(dbl), pi32, C.POINTER(dbl), C.POINTER(vp)]
        lib.ka_simplec_initialize.argtypes = [vp, vp, vp, vp, vp]
        lib.ka_simplec_create_mesh.argtypes = [vp, i32, vp, i32, vp, vp, vp, i32, vp, i32, i32, C.c_char_p, vp, vp,
                                               vp, vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(vp)]
        lib.ka_simplec_initialize_fields.argtypes = [vp] * 7
        lib.ka_simplec_create_profiled.argtypes = [vp, i32, i32, i32, C.POINTER(dbl), pi32, C.POINTER(dbl),
                                                   C.POINTER(dbl), pi32, pi32, BOUNDARY_PROFILE, vp,
                                                   C.POINTER(dbl), C.POINTER(vp)]
        lib.ka_simplec_sst_update.argtypes = [vp, C.POINTER(dbl), C.POINTER(vp)]
        lib.ka_simplec_destroy.argtypes = [vp]
        lib.ka_simplec_iterate.argtypes = [vp, i32]
        lib.ka_simplec_report.argtypes = [vp, i32, C.POINTER(dbl)]
        lib.ka_simplec_field.argtypes = [vp, C.c_char_p, vp]
        lib.ka_simplec_field.restype = C.c_long
        lib.ka_simplec_info.argtypes = [vp, pi32]
        lib.ka_simplec_forces.argtypes = [vp, C.c_char_p, dbl, C.POINTER(dbl)]
        lib.ka_simplec_force_coefficient.argtypes = [C.POINTER(dbl), C.POINTER(dbl), dbl, dbl, dbl, dbl,
                                                     C.POINTER(dbl)]
        lib.ka_simplec_centroids.argtypes = [vp, vp]
        lib.ka_simplec_set_batched_momentum.argtypes = [vp, i32, pi32]
        lib.ka_simplec_set_graphs.argtypes = [vp, i32]
        lib.ka_simplec_set_pressure_options.argtypes = [vp, vp, C.c_double]
        lib.ka_learned_parameter_count.argtypes = [i32]
        lib.ka_learned_parameter_count.restype = C.c_int64
        lib.ka_learned_create.argtypes = [vp, i32, C.c_uint64, C.POINTER(vp)]
        lib.ka_learned_destroy.argtypes = [vp]
        lib.ka_learned_set_graph.argtypes = [vp, vp]
        lib.ka_learned_set_parameters.argtypes = [vp, vp, C.c_int64]
        lib.ka_learned_parameters.argtypes = [vp, vp, C.c_int64]
        lib.ka_learned_infer.argtypes = [vp, vp, vp]
        lib.ka_learned_gradient.argtypes = [vp, vp, vp, vp, vp, C.POINTER(dbl)]
        lib.ka_learned_train.argtypes = [vp, vp, vp, i32, i32, C.c_float, vp, vp]
        _host = lib
    return _host


def _check_hip(status: int) -> None:
    if status != 0:
        raise KuttaError(hip_lib().kh_last_error().decode())


def _check_host(status: int) -> None:
    if status == 2:
        raise NonfiniteError(host_lib().ka_last_error().decode())
    if status != 0:
        raise KuttaError(host_lib().ka_last_error().decode())


# ----------------------------------------------------------------- device ABI
class KhArg(C.Structure):
    class _V(C.Union):
        _fields_ = [("ptr", C.c_void_p), ("f32", C.c_float), ("i32", C.c_int32),
                    ("u32", C.c_uint32), ("bits", C.c_uint64)]
    _fields_ = [("kind", C.c_uint32), ("size", C.c_uint32), ("value", _V)]


class KhReduction(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("pad", C.c_uint32), ("x", C.c_void_p),
                ("y", C.c_void_p), ("n", C.c_size_t)]


class Kernel(enum.IntEnum):
    """Kernel ids of include/kutta_hip.h (reference kernel names)."""
    csr_spmv = 1
    csr_spmv_guarded = 2
    csr_residual = 3
    csr_residual_guarded = 4
    krylov_diagonal = 5
    krylov_multiply = 6
    krylov_multiply_guarded = 7
    krylov_smooth = 8
    krylov_shift = 9
    krylov_direction = 10
    krylov_step = 11
    krylov_bicg_direction = 12
    krylov_difference = 13
    krylov_bicg_step = 14
    amg_coarse_solve = 15
    amg_smooth = 16
    amg_smooth_zero = 17
    amg_restrict = 18
    amg_prolong = 19
    amg_galerkin = 20
    amg_m_matrix_diagonal = 21
    amg_restrict_guarded = 22
    amg_prolong_guarded = 23
    axpy = 24
    xpby = 25
    axpbyz = 26
    scale_copy = 27
    mul = 28
    solver_limits = 29
    solver_reset = 30
    solver_pcg_alpha = 31
    solver_pcg_beta = 32
    solver_bicg_beta = 33
    solver_bicg_alpha = 34
    solver_bicg_s = 35
    solver_bicg_omega = 36
    solver_check = 37
    solver_step = 38
    solver_direction = 39
    solver_bicg_direction = 40
    solver_bicg_difference = 41
    solver_bicg_step = 42
    solver_project = 43
    fv_pressure_assemble = 100
    fv_velocity_boundary = 101
    fv_pressure_boundary = 102
    fv_pressure_faces = 103
    fv_momentum_source = 104
    fv_momentum_faces = 105
    fv_convective_outflow = 106
    fv_peclet_side_velocity = 107
    fv_rhie_chow = 108
    fv_correct_flux = 109
    fv_correct_velocity = 110
    fv_monitor_cell = 111
    fv_speed = 112
    fv_continuity = 113
    fv_wall_pressure_force = 114
    fv_wall_viscous_force = 115
    fv_fluent_momentum_scale = 116
    fv_physical_time_terms = 117
    fv_face_time_defect = 118
    fv_face_time_flux = 119
    fv_cell_mass_error = 120
    fv_gradient = 121
    fv_paired_gradient = 122
    fv_range_shift = 123
    fv_limiter = 124
    fv_paired_limiter = 125
    fv_paired_gradient_limiter = 126
    fv_vector_gradient = 127
    fv_scale_gradient = 128
    fv_limiter_scale = 129
    fv_paired_limiter_scale = 130
    fv_paired_gradient_limiter_scale = 131
    fv_jump_gradient = 132
    fv_face_jump = 133
    fv_gather_jump_gradient = 134
    fv_paired_jump_gradient = 135
    fv_lsq_restore_inverse = 136
    fv_diffusion_correction_bound = 137
    fv_transport_faces = 138
    fv_transport_assemble = 139
    fv_fixed_rows = 140
    fv_floor = 141
    fv_scalar_steady_measure = 142
    fv_scalar_set_fixed = 143
    fv_scalar_monitor = 144
    fv_compensated_pressure_jump = 145
    fv_paired_pressure_jump = 146
    fv_compensated_rhie_chow = 147
    fv_paired_pressure_boundary_low = 148
    fv_paired_pressure_faces = 149
    fv_paired_momentum_residual = 150
    fv_compensated_momentum_residual = 151
    fv_paired_accumulate = 152
    fv_compensated_pressure_force = 153
    fv_paired_velocity_boundary = 154
    fv_paired_advection = 155
    fv_paired_face_residual = 156
    fv_paired_flux = 157
    fv_paired_flux_offset = 158
    fv_paired_continuity = 159
    fv_paired_pressure_rhs = 160
    fv_paired_correct_flux = 161
    fv_paired_correct_velocity = 162
    fv_paired_value = 163
    fv_paired_vector_values = 164
    fv_compensated_correct_flux = 165
    fv_compensated_correct_velocity = 166
    fv_backflow_pressure = 167
    fv_effective_viscosity = 168
    fv_transpose_stress = 169
    fv_stress_source = 170
    fv_boundary_momentum = 171
    fv_momentum_diagonal = 172
    fv_sst_force = 173
    fv_turbulence_boundary = 174
    sst_rates = 175
    sst_blending = 176
    sst_viscosity = 177
    sst_blending_viscosity = 178
    sst_wall = 179
    sst_wall_omega = 180
    sst_wall_cells = 181
    sst_sources = 182
    sst_face_diffusion = 183
    sst_sustain = 184
    fv_simple_inverse = 185
    fv_standard_pressure_faces = 186
    fv_cell_pressure_boundary = 187
    fv_force_to_shear = 188
    fv_poor_cell_pressure_faces = 189
    fv_poor_pressure_interfaces = 190
    fused_spmv_dot_alpha = 200
    fused_step_dot_check = 201
    fused_residual_restrict = 202
    fused_prolong_smooth = 203
    fused_smooth_dot_beta = 204
    fused_smooth_sum = 205
    fused_project_dot_beta = 206
    fused_smooth_from_zero = 207
    fused_coarse_tail = 208
    fused_dot_bicg_beta = 209
    fused_bicg_direction = 210
    fused_spmv_shadow_alpha = 211
    fused_bicg_difference_s = 212
    fused_spmv_dot2_omega = 213
    fused_bicg_step_check_beta = 214
    gmres_start = 215
    gmres_arnoldi = 216
    gmres_update = 217
    dense_invert = 218
    dense_apply = 219
    learned_graph = 220
    learned_norm = 221
    learned_input = 222
    learned_layer = 223
    learned_head = 224
    learned_head_loss = 225
    learned_back = 226
    learned_weight_grad = 227
    learned_back_input = 228
    learned_update = 229
    learned_restrict = 230
    learned_prolong = 231
    batch_bicg_direction = 232
    batch_spmv_shadow_alpha = 233
    batch_bicg_difference_s = 234
    batch_spmv_dot2_omega = 235
    batch_bicg_step_check_beta = 236
    to_bf16 = 237
    bf16_amg_smooth = 238
    bf16_smooth_from_zero = 239
    bf16_residual_restrict = 240
    bf16_prolong_smooth = 241
    wall_distance = 242
    fused_coarse_sweeps = 243
    fused_residual_restrict_wide = 244
    amg_inner_alpha = 245
    amg_inner_first = 246
    amg_inner_check = 247
    amg_inner_beta = 248
    amg_inner_step = 249
    amg_smooth_wide = 250
    inner_first_step = 251
    inner_first_update = 252
    inner_beta_step = 253
    inner_second_step = 254
    amg_smooth_rowwave = 255
    coarse_sweeps_resident = 256
    csr_residual_rowwave = 257
    prolong_smooth_rowwave = 258


class Buffer:
    """A device array (kutta::Buffer<T>): float32, int32 or uint32 (or other
    fixed-size dtypes for backend-private storage)."""

    def __init__(self, device: "Device", n: int, dtype=np.float32):
        self.device = device
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        ptr = C.c_void_p()
        _check_hip(hip_lib().kh_alloc(device.kh, self.n * self.dtype.itemsize, C.byref(ptr)))
        self.ptr = ptr.value

    @property
    def nbytes(self) -> int:
        return self.n * self.dtype.itemsize

    def upload(self, host) -> "Buffer":
        a = np.ascontiguousarray(host, dtype=self.dtype)
        if a.size != self.n:
            raise KuttaError("upload size mismatch")
        _check_hip(hip_lib().kh_upload(self.device.kh, self.ptr, a.ctypes.data, self.nbytes))
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        _check_hip(hip_lib().kh_download(self.device.kh, self.ptr, out.ctypes.data, self.nbytes))
        return out

    def fill(self, value) -> None:
        word = np.array([value], dtype=self.dtype).view(np.uint32)[0] if self.dtype.itemsize == 4 else None
        if word is None:
            raise KuttaError("fill supports 32-bit element types only")
        _check_hip(hip_lib().kh_fill32(self.device.kh, self.ptr, int(word), self.n))

    def free(self) -> None:
        if self.ptr:
            hip_lib().kh_free(self.device.kh, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.device._alive:
                self.free()
        except Exception:
            pass


def _arg(a) -> KhArg:
    k = KhArg()
    if isinstance(a, Buffer):
        k.kind, k.size, k.value.ptr = 0, 0, a.ptr
    elif isinstance(a, (float, np.floating)):
        k.kind, k.size, k.value.f32 = 1, 4, float(a)
    elif isinstance(a, (int, np.integer)):
        k.kind, k.size, k.value.i32 = 1, 4, int(a)
    else:
        raise TypeError(f"unsupported kernel argument {a!r}")
    return k


class Device:
    """One MI355X device and its in-order stream (kutta::Device)."""

    def __init__(self, ordinal: int = 0):
        lib = host_lib()
        h = C.c_void_p()
        _check_host(lib.ka_device_create(ordinal, C.byref(h)))
        self.handle = h.value
        self.kh = lib.ka_device_raw(self.handle)
        self._alive = True

    @staticmethod
    def count() -> int:
        return hip_lib().kh_device_count()

    @property
    def name(self) -> str:
        return hip_lib().kh_name(self.kh).decode()

    def close(self) -> None:
        if self._alive:
            self._alive = False
            host_lib().ka_device_destroy(self.handle)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # buffers
    def alloc(self, n: int, dtype=np.float32) -> Buffer:
        return Buffer(self, n, dtype)

    def upload(self, host, dtype=None) -> Buffer:
        a = np.asarray(host) if dtype is None else np.asarray(host, dtype=dtype)
        return Buffer(self, a.size, a.dtype).upload(a)

    def copy(self, dst: Buffer, src: Buffer) -> None:
        if dst.nbytes != src.nbytes:
            raise KuttaError("copy size mismatch")
        _check_hip(hip_lib().kh_copy(self.kh, dst.ptr, src.ptr, src.nbytes))

    def synchronize(self) -> None:
        _check_hip(hip_lib().kh_synchronize(self.kh))

    def launch(self, kernel, n: int, *args) -> None:
        arr = (KhArg * max(len(args), 1))(*[_arg(a) for a in args])
        _check_hip(hip_lib().kh_launch(self.kh, int(kernel), int(n), arr, len(args)))

    # reductions (device.hpp:157-202)
    def reduce_many(self, requests) -> list[float]:
        reqs = (KhReduction * max(len(requests), 1))()
        for i, (kind, x, y, n) in enumerate(requests):
            reqs[i].kind = {"sum": 0, "dot": 1, "max_abs": 2}[kind]
            reqs[i].x = x.ptr
            reqs[i].y = (y or x).ptr
            reqs[i].n = n
        out = (C.c_double * max(len(requests), 1))()
        _check_hip(hip_lib().kh_reduce_host(self.kh, reqs, len(requests), out))
        return [out[i] for i in range(len(requests))]

    def sum(self, x: Buffer, n: int | None = None) -> float:
        return self.reduce_many([("sum", x, None, x.n if n is None else n)])[0]

    def dot(self, x: Buffer, y: Buffer, n: int | None = None) -> float:
        return self.reduce_many([("dot", x, y, x.n if n is None else n)])[0]

    def max_abs(self, x: Buffer, n: int | None = None) -> float:
        return self.reduce_many([("max_abs", x, None, x.n if n is None else n)])[0]

    def reduce_into(self, mode: str, x, y, u, v, n, dst: Buffer, offset=0) -> None:
        m = {"sum": 0, "dot": 1, "dot2": 2}[mode]
        _check_hip(hip_lib().kh_reduce_into(self.kh, m, x.ptr, (y or x).ptr, (u or x).ptr,
                                            (v or x).ptr, n, dst.ptr, offset))

    def sum_into(self, x, n, dst, offset=0):
        self.reduce_into("sum", x, x, x, x, n, dst, offset)

    def dot_into(self, x, y, n, dst, offset=0):
        self.reduce_into("dot", x, y, x, y, n, dst, offset)

    def dot2_into(self, x, y, u, v, n, dst, offset=0):
        self.reduce_into("dot2", x, y, u, v, n, dst, offset)

    # profiling
    def profile(self, kernel, enable: bool = True) -> None:
        _check_hip(hip_lib().kh_profile_enable(self.kh, int(kernel), 1 if enable else 0))

    def profile_filter(self, kernel, n: int) -> None:
        lib = hip_lib()
        lib.kh_profile_filter.argtypes = [C.c_void_p, C.c_uint32, C.c_size_t]
        _check_hip(lib.kh_profile_filter(self.kh, int(kernel), int(n)))

    def logical_bytes(self, enable: bool) -> tuple[float, int]:
        """(bytes, counted launches) of the logical-bytes model since the last
        call (kh_logical_bytes); resets them and turns counting on or off."""
        b, n = C.c_double(), C.c_long()
        _check_hip(hip_lib().kh_logical_bytes(self.kh, int(enable), C.byref(b), C.byref(n)))
        return b.value, n.value

    def time_launches(self, kernel, n: int, *args, repeats: int = 20) -> float:
        """Milliseconds of `repeats` back-to-back launches, bracketed by one
        pair of HIP events on the device stream."""
        arr = (KhArg * max(len(args), 1))(*[_arg(a) for a in args])
        ms = C.c_double()
        _check_hip(hip_lib().kh_time_launches(self.kh, int(kernel), int(n), arr, len(args), int(repeats),
                                              C.byref(ms)))
        return ms.value

    def profile_read(self, kernel) -> tuple[int, float]:
        n, ms = C.c_long(), C.c_double()
        _check_hip(hip_lib().kh_profile_read(self.kh, int(kernel), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def profile_reset(self) -> None:
        _check_hip(hip_lib().kh_profile_reset(self.kh))


# ----------------------------------------------------------------- linalg API
class SolverKind(enum.IntEnum):
    """linalg/solver.hpp:237-243"""
    Jacobi = 0
    PcgJacobi = 1
    BicgstabJacobi = 2
    AmgPcg = 3
    AmgBicgstab = 4
    FgmresJacobi = 5   # flexible GMRES(m): no reference counterpart
    AmgFgmres = 6


class NullSpace(enum.IntEnum):
    NONE = 0
    Constant = 1


class OperatorStructure(enum.IntEnum):
    General = 0
    SymmetricMMatrix = 1


class AmgCycle(enum.IntEnum):
    """linalg/amg.hpp:6"""
    V = 0
    K = 1


@dataclass
class AmgOptions:
    """linalg/amg.hpp:8-20. k_depth and k_min_rows (no reference
    counterpart) keep the K-cycle's inner corrections to coarse levels down
    to k_depth and of at least k_min_rows rows; 0 keeps them wherever the
    reference uses them."""
    pre_sweeps: int = 5
    post_sweeps: int = 5
    coarse_sweeps: int = 32
    coarse_rows: int = 24
    max_levels: int = 24
    aggregation_passes: int = 3
    strength_threshold: float = 0.25
    nullspace: NullSpace = NullSpace.NONE
    structure: OperatorStructure = OperatorStructure.General
    cycle: AmgCycle = AmgCycle.V
    k_depth: int = 0
    k_min_rows: int = 0

    def packed(self):
        return (C.c_int32 * 11)(self.pre_sweeps, self.post_sweeps, self.coarse_sweeps,
                                self.coarse_rows, self.max_levels, self.aggregation_passes,
                                int(self.nullspace), int(self.structure), int(self.cycle), self.k_depth,
                                self.k_min_rows)


@dataclass
class SolveControl:
    """linalg/solver.hpp:223-228"""
    rel_tol: float = 1e-2
    abs_tol: float = 1e-12
    max_iter: int = 200
    residual_check_interval: int = 8


@dataclass
class SolveStats:
    """linalg/solver.hpp:230-235"""
    iterations: int = 0
    initial_residual: float = 0.0
    final_residual: float = 0.0
    converged: bool = False


class CsrMatrixHost:
    """Host CSR (linalg/csr.hpp:16-28)."""

    def __init__(self, row_ptr, col, diag, val):
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.diag = np.ascontiguousarray(diag, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.n_rows = len(self.row_ptr) - 1
        h = C.c_void_p()
        _check_host(host_lib().ka_csr_create(self.n_rows, self.row_ptr.ctypes.data,
                                             self.col.ctypes.data, self.diag.ctypes.data,
                                             self.val.ctypes.data, C.byref(h)))
        self.handle = h.value

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1])

    def __del__(self):
        try:
            host_lib().ka_csr_destroy(self.handle)
        except Exception:
            pass


class CsrMatrix:
    """Device CSR (linalg/csr.hpp:31-38); `val` changes between solves."""

    def __init__(self, device: Device, host: CsrMatrixHost):
        h = C.c_void_p()
        _check_host(host_lib().ka_matrix_upload(device.handle, host.handle, C.byref(h)))
        self.device = device
        self.handle = h.value
        self.n_rows = host.n_rows
        self.nnz = host.nnz

    def _view(self, fn, n, dtype):
        b = Buffer.__new__(Buffer)
        b.device, b.dtype, b.n = self.device, np.dtype(dtype), n
        b.ptr = getattr(host_lib(), fn)(self.handle)
        b.free = lambda: None  # owned by the matrix
        return b

    @property
    def val(self) -> Buffer:
        return self._view("ka_matrix_values", self.nnz, np.float32)

    @property
    def row_ptr(self) -> Buffer:
        return self._view("ka_matrix_row_ptr", self.n_rows + 1, np.int32)

    @property
    def col(self) -> Buffer:
        return self._view("ka_matrix_col", self.nnz, np.int32)

    @property
    def diag(self) -> Buffer:
        return self._view("ka_matrix_diag", self.n_rows, np.int32)

    def __del__(self):
        try:
            if self.device._alive:
                host_lib().ka_matrix_destroy(self.handle)
        except Exception:
            pass


def upload(device: Device, host: CsrMatrixHost) -> CsrMatrix:
    """linalg/csr.hpp:40-41"""
    return CsrMatrix(device, host)


class LinearSolver:
    """A device-resident Krylov solver (linalg/solver.hpp:279-286)."""

    def __init__(self, device: Device, kind: SolverKind, a: CsrMatrixHost, options: AmgOptions):
        h = C.c_void_p()
        _check_host(host_lib().ka_solver_create(device.handle, int(kind), a.handle, options.packed(),
                                                options.strength_threshold, C.byref(h)))
        self.device = device
        self.handle = h.value
        self.n_rows = a.n_rows

    def refresh(self, a: CsrMatrix) -> None:
        _check_host(host_lib().ka_solver_refresh(self.handle, a.handle))

    def solve(self, a: CsrMatrix, x: Buffer, b: Buffer, control: SolveControl | None = None) -> SolveStats:
        c = control or SolveControl()
        stats = (C.c_double * 4)()
        _check_host(host_lib().ka_solver_solve(self.handle, a.handle, x.ptr, b.ptr, c.rel_tol, c.abs_tol,
                                               c.max_iter, c.residual_check_interval, stats))
        return SolveStats(int(stats[0]), stats[1], stats[2], bool(stats[3]))

    def set_direct(self, rows: int) -> int:
        """Mixed-precision AMG: solve the first coarse level of at most `rows`
        rows by an FP64 dense inverse (computed at the next refresh); 0 off.
        Returns that level (0: none)."""
        level = C.c_int32()
        _check_host(host_lib().ka_solver_set_direct(self.handle, int(rows), C.byref(level)))
        return level.value

    def set_dense_tolerance(self, tolerance: float) -> None:
        """Lazy refresh of the dense level's inverse: recompute it only when
        the level's values moved more than `tolerance` (relative l2) from
        those it was computed from; 0 recomputes at every refresh."""
        _check_host(host_lib().ka_solver_set_dense_tolerance(self.handle, float(tolerance)))

    def dense_refreshes(self) -> tuple[int, int, float]:
        """(inversions done, inverses kept, last relative distance)."""
        i, k, d = C.c_int64(), C.c_int64(), C.c_double()
        _check_host(host_lib().ka_solver_dense_refreshes(self.handle, C.byref(i), C.byref(k), C.byref(d)))
        return i.value, k.value, d.value

    def set_learned(self, level: int = 2, layers: int = 3, seed: int = 0, capacity: int = 256) -> None:
        """Learned coarse correction stage (AMG FGMRES only): a network on
        coarse level `level` with room for `capacity` harvested pairs."""
        _check_host(host_lib().ka_solver_set_learned(self.handle, int(level), int(layers), int(seed),
                                                     int(capacity)))

    def set_learned_mode(self, mode: int, reduction: float = 0.5, max_relative: float = 0.5) -> None:
        """0 off, 1 harvest, 2 apply: not applied when the last training epoch
        left more than `max_relative` of the targets' mean square unexplained,
        dropped within a solve when the residual falls by less than
        `reduction` per step."""
        _check_host(host_lib().ka_solver_learned_mode(self.handle, int(mode), float(reduction),
                                                      float(max_relative)))

    def learned_targets(self, rel_tol: float = 1e-6) -> int:
        k = C.c_int32()
        _check_host(host_lib().ka_solver_learned_targets(self.handle, float(rel_tol), C.byref(k)))
        return k.value

    def learned_train(self, steps: int, lr: float = 1e-3, pairs: int = 0):
        """(loss, unexplained fraction) before each Adam step, cycling over the
        first `pairs` harvested pairs (0: all)."""
        out = np.empty(max(int(steps), 1), np.float64)
        rel = np.empty(max(int(steps), 1), np.float64)
        _check_host(host_lib().ka_solver_learned_train(self.handle, int(steps), float(lr), int(pairs),
                                                       out.ctypes.data, rel.ctypes.data))
        return out[:int(steps)], rel[:int(steps)]

    def learned_evaluate(self, first: int, count: int) -> float:
        """Mean unexplained fraction on pairs [first, first + count), untrained."""
        v = C.c_double()
        _check_host(host_lib().ka_solver_learned_evaluate(self.handle, int(first), int(count), C.byref(v)))
        return v.value

    def learned_report(self) -> dict:
        out = (C.c_double * 11)()
        _check_host(host_lib().ka_solver_learned_report(self.handle, out))
        keys = ["level", "rows", "mode", "harvested", "targets", "applications", "fallbacks", "solves"]
        rep = {k: int(out[i]) for i, k in enumerate(keys)}
        rep["last_loss"] = out[8]
        rep["relative"] = out[9]
        rep["gated"] = int(out[10])
        return rep

    def learned_parameters(self, layers: int = 3) -> np.ndarray:
        out = np.empty(LearnedNetwork.parameter_count(layers), np.float32)
        _check_host(host_lib().ka_solver_learned_parameters(self.handle, out.ctypes.data, out.size, 0))
        return out

    def set_learned_parameters(self, theta) -> None:
        t = np.ascontiguousarray(theta, dtype=np.float32)
        _check_host(host_lib().ka_solver_learned_parameters(self.handle, t.ctypes.data, t.size, 1))

    def autotune_tail(self, a, repeats: int = 5) -> int:
        """Time V-cycles for each admissible coarse-tail level (and none) on
        the refreshed operator; keep and return the fastest (identical
        iterates whichever)."""
        level = C.c_int32()
        _check_host(host_lib().ka_solver_autotune_tail(self.handle, a.handle, int(repeats), C.byref(level)))
        return level.value

    def set_graphs(self, on: bool) -> None:
        """Replay each fused PCG iteration from a captured HIP graph (default
        on; identical iterates)."""
        _check_host(host_lib().ka_solver_set_graphs(self.handle, int(bool(on))))

    def set_coarse_bf16(self, on: bool) -> None:
        """Mixed-precision AMG: bfloat16 copies of the coarse levels' values
        (made at the next refresh) for their smoothing and transfers."""
        _check_host(host_lib().ka_solver_set_coarse_bf16(self.handle, int(bool(on))))

    def set_restart(self, m: int) -> None:
        """Restart length of the FGMRES kinds (1..63; default 30)."""
        _check_host(host_lib().ka_solver_set_restart(self.handle, int(m)))

    def describe(self) -> str:
        buf = C.create_string_buffer(512)
        host_lib().ka_solver_describe(self.handle, buf, 512)
        return buf.value.decode()

    def set_plan(self, restrict_mode: int = -1, coarse_sweeps: int = -1, prolong_mode: int = -1,
                 sweep_mode: int = -1) -> None:
        """Forces the V-cycle forms (see plan()); -1 keeps a setting.
        restrict_mode: 0 fused, 1 wide, 2 split, 3 split with a wavefront per
        residual row; coarse_sweeps: 0 a launch per coarsest-level sweep, 1
        one workgroup, 2 one workgroup reading the operator once;
        prolong_mode: 0 folded, 1 split, 2 folded with a wavefront per row;
        sweep_mode: 0 thread per row, 1 wavefront per row."""
        _check_host(host_lib().ka_solver_set_plan(self.handle, int(restrict_mode), int(coarse_sweeps),
                                                  int(prolong_mode), int(sweep_mode)))

    def plan(self) -> str:
        """The V-cycle forms the autotuner picked (same iterates for all)."""
        buf = C.create_string_buffer(2048)
        host_lib().ka_solver_plan(self.handle, buf, 2048)
        return buf.value.decode()

    def levels(self) -> list[tuple[int, int]]:
        rows, nnz = (C.c_int32 * 64)(), (C.c_int32 * 64)()
        k = host_lib().ka_solver_levels(self.handle, rows, nnz, 64)
        return [(rows[i], nnz[i]) for i in range(k)]

    def transfer(self, level: int, which: str) -> np.ndarray | None:
        w = ["aggregate", "member_offsets", "members", "entry_offsets", "entries"].index(which)
        n = host_lib().ka_solver_transfer(self.handle, level, w, None)
        if n < 0:
            return None
        out = np.empty(n, dtype=np.int32)
        host_lib().ka_solver_transfer(self.handle, level, w, out.ctypes.data_as(C.POINTER(C.c_int32)))
        return out

    def level_values(self, level: int) -> np.ndarray | None:
        n = host_lib().ka_solver_level_values(self.handle, level, None)
        if n < 0:
            return None
        out = np.empty(n, dtype=np.float32)
        host_lib().ka_solver_level_values(self.handle, level, out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def set_fused(self, fused: bool) -> None:
        _check_host(host_lib().ka_solver_set_fused(self.handle, 1 if fused else 0))

    @property
    def host_waits(self) -> int:
        return host_lib().ka_solver_host_waits(self.handle)

    def set_tail(self, first: int = -1) -> int:
        """First level of the one-workgroup coarse-grid tail: -1 automatic,
        0 off, k >= 1 level k. Returns the level in use (0: none)."""
        level = C.c_int32(0)
        _check_host(host_lib().ka_solver_set_tail(self.handle, int(first), C.byref(level)))
        return level.value

    def __del__(self):
        try:
            if self.device._alive:
                host_lib().ka_solver_destroy(self.handle)
        except Exception:
            pass


_ARRAY_KINDS = {0: np.int32, 1: np.float32, 2: np.float64}


def _named_array(fn, handle, name: str) -> np.ndarray:
    fn.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_void_p]
    fn.restype = C.c_long
    for kind, dtype in _ARRAY_KINDS.items():
        n = fn(handle, name.encode(), kind, None)
        if n >= 0:
            out = np.empty(n, dtype)
            fn(handle, name.encode(), kind, out.ctypes.data)
            return out
    raise KeyError(name)


C2_SEED = 20261202
# Largest coarsest level smoothed in one launch (KH_COARSE_SWEEPS_MAX_ROWS).
COARSE_SWEEPS_MAX_ROWS = 16384

# Limits of the register-resident coarsest-level sweeps (KH_RESIDENT_* in
# kutta_hip.h).
RESIDENT_THREADS = 512
RESIDENT_MAX_ROWS = 4096


def resident_partition(row_ptr, cols=None):
    """kh_resident_partition: the row partition of KH_COARSE_SWEEPS_RESIDENT
    for a CSR pattern, as (table int32[2 * RESIDENT_THREADS]: first row and
    end row per thread, cap: entries per thread of the parallel phase), or
    None when the level does not fit. Host only."""
    rp = np.ascontiguousarray(row_ptr, np.int32)
    cl = np.ascontiguousarray(cols if cols is not None else np.zeros(1), np.int32)
    table = np.zeros(2 * RESIDENT_THREADS, np.int32)
    cap = C.c_int32(0)
    if hip_lib().kh_resident_partition(rp.ctypes.data, cl.ctypes.data, len(rp) - 1, table.ctypes.data,
                                       C.byref(cap)) != 0:
        return None
    return table, cap.value


# Limits of the one-workgroup coarse-grid tail (KH_TAIL_* in kutta_hip.h).
TAIL_MAX_ROWS = 8192
TAIL_MAX_LEVELS = 16
TAIL_LDS_FLOATS = 24576


class MeshSetup:
    """Host mesh setup of a box on [0,1]^3: make_box_tet/hex, optional seeded
    perturbation, RCM, FP64 geometry and FV stencil (no device needed)."""

    def __init__(self, nx, ny, nz, tet=True, amplitude=0.2, seed=C2_SEED):
        lib = host_lib()
        lib.ka_mesh_setup_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                             C.c_ulonglong, C.POINTER(C.c_void_p)]
        lib.ka_mesh_setup_destroy.argtypes = [C.c_void_p]
        h = C.c_void_p()
        _check_host(lib.ka_mesh_setup_create(nx, ny, nz, int(tet), amplitude, seed, C.byref(h)))
        self.handle = h.value

    def __getitem__(self, name: str) -> np.ndarray:
        return _named_array(host_lib().ka_mesh_setup_array, self.handle, name)

    def __del__(self):
        try:
            host_lib().ka_mesh_setup_destroy(self.handle)
        except Exception:
            pass


class PressureProblem:
    """The SIMPLEC pressure-correction operator of a box mesh, assembled on the
    device by fv_pressure_assemble (zero flux, diffusion coefficients). Usable
    as the matrix of a solver; ``host`` is its host CSR for solver setup."""

    def __init__(self, device: Device, nx, ny, nz, tet=True, amplitude=0.2, seed=C2_SEED,
                 reference=-2):
        lib = host_lib()
        lib.ka_pressure_problem_create.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                                   C.c_double, C.c_ulonglong, C.c_int,
                                                   C.POINTER(C.c_void_p)]
        lib.ka_pressure_problem_destroy.argtypes = [C.c_void_p]
        lib.ka_pressure_problem_csr.argtypes = [C.c_void_p]
        lib.ka_pressure_problem_csr.restype = C.c_void_p
        lib.ka_pressure_problem_assemble.argtypes = [C.c_void_p, C.c_int]
        lib.ka_pressure_problem_info.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        h = C.c_void_p()
        _check_host(lib.ka_pressure_problem_create(device.handle, nx, ny, nz, int(tet), amplitude,
                                                   seed, reference, C.byref(h)))
        self.device = device
        self.handle = h.value
        info = np.zeros(7, np.int32)
        secs = np.zeros(6, np.float64)
        _check_host(lib.ka_pressure_problem_info(self.handle, info.ctypes.data, secs.ctypes.data))
        (self.n_cells, self.n_faces, self.n_internal_faces, self.nnz, self.n_entries,
         self.rounds, self.reference) = (int(v) for v in info)
        self.n_rows = self.n_cells
        self.setup_seconds = dict(zip(["box", "perturb", "renumber", "geometry", "stencil",
                                       "upload_assemble"], secs.tolist()))
        self.host = _BorrowedCsr(lib.ka_pressure_problem_csr(self.handle), self.n_rows, self.nnz)

    def __getitem__(self, name: str) -> np.ndarray:
        return _named_array(host_lib().ka_pressure_problem_array, self.handle, name)

    def assemble(self, reference: int = -2) -> None:
        _check_host(host_lib().ka_pressure_problem_assemble(self.handle, reference))

    def _view(self, fn, n, dtype):
        b = Buffer.__new__(Buffer)
        b.device, b.dtype, b.n = self.device, np.dtype(dtype), n
        b.ptr = getattr(host_lib(), fn)(self.handle)
        b.free = lambda: None  # owned by the problem
        return b

    @property
    def val(self) -> Buffer:
        return self._view("ka_matrix_values", self.nnz, np.float32)

    @property
    def row_ptr(self) -> Buffer:
        return self._view("ka_matrix_row_ptr", self.n_rows + 1, np.int32)

    @property
    def col(self) -> Buffer:
        return self._view("ka_matrix_col", self.nnz, np.int32)

    @property
    def diag(self) -> Buffer:
        return self._view("ka_matrix_diag", self.n_rows, np.int32)

    def __del__(self):
        try:
            if self.device._alive:
                host_lib().ka_pressure_problem_destroy(self.handle)
        except Exception:
            pass


class _BorrowedCsr:
    """A host CSR owned by another native object (same interface as CsrMatrixHost
    for solver construction)."""

    def __init__(self, handle, n_rows, nnz):
        self.handle, self.n_rows, self.nnz = handle, n_rows, nnz


def aggregate(a: CsrMatrixHost, options: AmgOptions | None = None) -> dict:
    """One level of host aggregation (linalg/amg.cpp:160-174): aggregate map,
    member lists, Galerkin entry maps and the coarse CSR operator."""
    o = options or AmgOptions()
    h = C.c_void_p()
    _check_host(host_lib().ka_aggregate(a.handle, o.packed(), o.strength_threshold, C.byref(h)))
    lib = host_lib()
    out = {}
    try:
        names = ["aggregate", "member_offsets", "members", "entry_offsets", "entries",
                 "row_ptr", "col", "diag"]
        for k, name in enumerate(names):
            n = lib.ka_coarsening_array(h, k, None)
            arr = np.empty(n, np.int32)
            lib.ka_coarsening_array(h, k, arr.ctypes.data_as(C.POINTER(C.c_int32)))
            out[name] = arr
        n = lib.ka_coarsening_values(h, None)
        val = np.empty(n, np.float64)
        lib.ka_coarsening_values(h, val.ctypes.data_as(C.POINTER(C.c_double)))
        out["val"] = val
    finally:
        lib.ka_coarsening_destroy(h)
    return out


def solve_batched(solvers, matrices, xs, bs, control: "SolveControl | None" = None) -> list:
    """Several fused BiCGStab-Jacobi solves at once (no reference
    counterpart): each x and SolveStats is bitwise its own solve's."""
    c = control or SolveControl()
    k = len(solvers)
    arr = lambda items: (C.c_void_p * k)(*items)
    stats = (C.c_double * (4 * k))()
    _check_host(host_lib().ka_solver_solve_batched(arr([s.handle for s in solvers]), arr([m.handle for m in matrices]),
                                                   arr([x.ptr for x in xs]), arr([b.ptr for b in bs]), k,
                                                   c.rel_tol, c.abs_tol, c.max_iter, c.residual_check_interval,
                                                   stats))
    return [SolveStats(int(stats[4 * i]), stats[4 * i + 1], stats[4 * i + 2], bool(stats[4 * i + 3]))
            for i in range(k)]


def make_amg_solver(kind: SolverKind, device: Device, a: CsrMatrixHost,
                    options: AmgOptions | None = None) -> LinearSolver:
    """linalg/amg.hpp:22-27 (only AmgPcg / AmgBicgstab)."""
    if kind not in (SolverKind.AmgPcg, SolverKind.AmgBicgstab, SolverKind.AmgFgmres):
        raise KuttaError("AMG: expected an AMG solver kind")
    return LinearSolver(device, kind, a, options or AmgOptions())


def make_solver(kind: SolverKind, device: Device, a: CsrMatrixHost,
                nullspace: NullSpace = NullSpace.NONE,
                structure: OperatorStructure = OperatorStructure.General) -> LinearSolver:
    """linalg/solver.hpp:290-293; AMG kinds use default AmgOptions."""
    return LinearSolver(device, kind, a, AmgOptions(nullspace=nullspace, structure=structure))


# ---------------------------------------------------------------- SIMPLEC
PATCHES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


@dataclass
class Wall:
    """fv/steady_solver.hpp:23-27 (a moving or stationary no-slip wall)."""
    velocity: tuple = (0.0, 0.0, 0.0)


@dataclass
class Symmetry:
    """fv/steady_solver.hpp:65."""


@dataclass
class TurbulenceValues:
    """fv/steady_solver.hpp:28: inflow k and omega."""
    k: float = 0.0
    omega: float = 0.0


@dataclass
class IntensityLength:
    """fv/steady_solver.hpp:29: intensity (a fraction) and length scale."""
    intensity: float = 0.0
    length: float = 0.0


@dataclass
class IntensityViscosityRatio:
    """fv/steady_solver.hpp:30: intensity and turbulent viscosity ratio."""
    intensity: float = 0.0
    ratio: float = 0.0


@dataclass
class VelocityInlet:
    """fv/steady_solver.hpp:38-42; the velocity is a tuple or a function of
    the face centroid; `turbulence` (TurbulenceValues, IntensityLength or
    IntensityViscosityRatio) is required under SST."""
    velocity: object = (0.0, 0.0, 0.0)
    turbulence: object = None


@dataclass
class MassFlowInlet:
    """fv/steady_solver.hpp:60-65: kg/s into the domain over the whole
    patch, along `direction` or each face's inward normal (None)."""
    mass_flow: float = 0.0
    direction: tuple | None = None
    turbulence: object = None


@dataclass
class PressureOutlet:
    """fv/steady_solver.hpp:43-51: static `pressure`; backflow along the face
    normal when `backflow_velocity` is None (NormalBackflow, the default),
    else that velocity (a tuple or a function of the face centroid);
    `backflow_total` selects BackflowPressure::Total."""
    pressure: float = 0.0
    backflow_velocity: object = None
    backflow_total: bool = False
    turbulence: object = None


@dataclass
class SolverSettings:
    """fv/steady_solver.hpp:193-271: the members the MI355X SIMPLEC honours
    (the others keep the reference defaults)."""
    density: float = 1.0
    viscosity: float = 0.01
    momentum_relaxation: float = 0.7
    pressure_relaxation: float = 1.0
    pseudo_time_scale: float = 1.0
    monitor_interval: int = 20
    paired_state: bool = True
    paired_state_slack_ulps: float = 1.0
    reference_point: tuple = (0.0, 0.0, 0.0)
    pseudo_time_step: float = 0.0  # fixed pseudo time step; 0 is automatic

    def packed(self):
        """settings[12] (the box constructors read the first 11)."""
        return (C.c_double * 12)(self.density, self.viscosity, self.momentum_relaxation,
                                 self.pressure_relaxation, self.pseudo_time_scale, self.monitor_interval,
                                 1.0 if self.paired_state else 0.0, self.paired_state_slack_ulps,
                                 *self.reference_point, self.pseudo_time_step)


def pack_conditions(conditions: dict):
    """One condition per box patch in generator order (xmin ... zmax)."""
    kinds, wall = (C.c_int32 * 6)(), (C.c_double * 18)()
    for p, name in enumerate(PATCHES):
        c = conditions[name]
        if isinstance(c, Wall):
            kinds[p] = 0
            wall[3 * p:3 * p + 3] = [float(v) for v in c.velocity]
        elif isinstance(c, Symmetry):
            kinds[p] = 3
        else:
            raise KuttaError(f"SIMPLEC: unsupported condition for {name}")
    return kinds, wall


@dataclass
class SstOptions:
    """turbulence/sst.hpp:46-78 (SstOptions; no intermittency coupling)."""
    ambient: tuple | None = None  # (k, omega)
    production_limiter: bool = True
    production_clip: float = 10.0
    limit_omega_production: bool = True
    omega_production_molecular_ratio: float = 0.0
    low_re: bool = False
    gamma_menter2003: bool = False
    vorticity_limiter: bool = False
    kato_launder: bool = False
    vorticity_production: bool = False
    wall_production: bool = True
    cdkw_floor: float = 1e-10
    limit_cross_diffusion: bool = True
    cross_diffusion_clip: float = 1e6
    omega_floor: float = 1e-20
    k_floor: float = 1e-14
    max_viscosity_ratio: float = 1e5

    def packed(self):
        a = self.ambient
        return (C.c_double * 19)(1.0 if a else 0.0, a[0] if a else 0.0, a[1] if a else 0.0,
                                 float(self.production_limiter), self.production_clip,
                                 float(self.limit_omega_production), self.omega_production_molecular_ratio,
                                 float(self.low_re), float(self.gamma_menter2003), float(self.vorticity_limiter),
                                 float(self.kato_launder), float(self.vorticity_production),
                                 float(self.wall_production), self.cdkw_floor, float(self.limit_cross_diffusion),
                                 self.cross_diffusion_clip, self.omega_floor, self.k_floor, self.max_viscosity_ratio)


@dataclass
class SstSettings:
    """fv/steady_solver.hpp:102-118 (no transition): the model options, the
    k and omega convection schemes (True: second-order upwind), relaxation,
    the linear solve control (rel_tol, abs_tol, max_iter), the steady
    tolerance, compensated transport and pseudo-transient scalars."""
    model: SstOptions = field(default_factory=SstOptions)
    k_second_order: bool = True
    omega_second_order: bool = True
    k_relaxation: float = 0.75
    omega_relaxation: float = 0.75
    linear: tuple = (1e-3, 1e-20, 100)
    steady_tolerance: float = 1e-6
    compensated_transport: bool = True
    pseudo_transient: bool = False

    def packed(self):
        return (C.c_double * 29)(*self.model.packed(), float(self.k_second_order), float(self.omega_second_order),
                                 self.k_relaxation, self.omega_relaxation, *[float(v) for v in self.linear],
                                 self.steady_tolerance, float(self.compensated_transport),
                                 float(self.pseudo_transient))


def _turbulence_code(t):
    if t is None:
        return -1, 0.0, 0.0
    if isinstance(t, TurbulenceValues):
        return 0, float(t.k), float(t.omega)
    if isinstance(t, IntensityLength):
        return 1, float(t.intensity), float(t.length)
    if isinstance(t, IntensityViscosityRatio):
        return 2, float(t.intensity), float(t.ratio)
    raise KuttaError(f"FV: unsupported turbulence condition {t!r}")


def pack_mesh_conditions(mesh, conditions: dict) -> dict:
    """Per-patch arrays of ka_simplec_create_mesh for `mesh`'s patches, from
    a dict patch name -> condition (uniform velocities only)."""
    names = [p[0] for p in mesh.patches]
    missing = [n for n in names if n not in conditions]
    if missing or len(conditions) != len(names):
        raise KuttaError(f"FV supply exactly one condition per patch (missing {missing})")
    P = len(names)
    out = {"kinds": (C.c_int32 * P)(), "velocity": (C.c_double * (3 * P))(), "pressure": (C.c_double * P)(),
           "total": (C.c_int32 * P)(), "flags": (C.c_int32 * P)(), "turbulence": (C.c_int32 * P)(),
           "turbulence_ab": (C.c_double * (2 * P))()}
    for p, name in enumerate(names):
        c = conditions[name]
        code, a, b = _turbulence_code(getattr(c, "turbulence", None))
        out["turbulence"][p] = code
        out["turbulence_ab"][2 * p:2 * p + 2] = [a, b]
        vel = None
        if isinstance(c, Wall):
            out["kinds"][p], vel = 0, c.velocity
        elif isinstance(c, VelocityInlet):
            out["kinds"][p], vel = 1, c.velocity
        elif isinstance(c, Symmetry):
            out["kinds"][p] = 3
        elif isinstance(c, MassFlowInlet):
            out["kinds"][p] = 5
            out["pressure"][p] = float(c.mass_flow)
            if c.direction is not None:
                vel = c.direction
                out["flags"][p] = 1
        elif isinstance(c, PressureOutlet):
            out["kinds"][p] = 4 if c.backflow_velocity is None else 2
            vel = c.backflow_velocity
            out["pressure"][p] = float(c.pressure)
            out["total"][p] = 1 if c.backflow_total else 0
        else:
            raise KuttaError(f"SIMPLEC: unsupported condition for {name}")
        if vel is not None:
            if callable(vel):
                raise KuttaError("SIMPLEC: profiles on general meshes are not supported; give a uniform velocity")
            out["velocity"][3 * p:3 * p + 3] = [float(v) for v in vel]
    return out


SST_INPUTS = ("u", "v", "w", "dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz", "dwdx", "dwdy", "dwdz", "density",
              "viscosity", "k", "omega", "dkdx", "dkdy", "dkdz", "domegadx", "domegady", "domegadz",
              "wall_u", "wall_v", "wall_w")
SST_FIELDS = ("mu_t", "k_source", "k_sink", "omega_source", "omega_sink", "strain", "vorticity", "f1", "f2", "cdkw",
              "diffusivity_k", "diffusivity_omega", "production_k", "fixed_omega", "wall_production",
              "wall_tau", "wall_u_star", "wall_u_tau", "wall_y_plus", "wall_y_plus_utau", "wall_shear_x",
              "wall_shear_y", "wall_shear_z", "wall_shear_coefficient", "wall_face_omega",
              "face_diffusivity_k", "face_diffusivity_omega")
SST_INT_FIELDS = ("fixed_omega_mask", "selected_wall_face")


BOUNDARY_PROFILE = C.CFUNCTYPE(None, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)


class _ProfileCallback:
    """The C callback around the patches' velocity functions. A function that
    raises, or returns anything but three numbers, writes NaN, which the
    solver's finiteness check rejects; the first such exception is kept and
    re-raised once the constructor returns (an exception cannot cross the C
    frames)."""

    def __init__(self, functions):
        self.functions = functions
        self.error = None
        self.c_function = BOUNDARY_PROFILE(self._call)

    def _call(self, patch, point, velocity, _user):
        try:
            v = [float(x) for x in self.functions[patch]((point[0], point[1], point[2]))]
            if len(v) != 3:
                raise ValueError(f"boundary profile of patch {PATCHES[patch]} returned {len(v)} components, not 3")
            velocity[0], velocity[1], velocity[2] = v
        except Exception as exc:  # noqa: BLE001 - reported after the C call returns
            if self.error is None:
                self.error = exc
            velocity[0] = velocity[1] = velocity[2] = float("nan")

    def reraise(self):
        if self.error is not None:
            raise self.error


def profile_callback(conditions: dict):
    """(profiled[6], callback) for the patches whose inlet or backflow
    velocity is a function of the face centroid, else (None, None)."""
    functions = {}
    for p, name in enumerate(PATCHES):
        c = conditions[name]
        f = c.velocity if isinstance(c, VelocityInlet) else (
            c.backflow_velocity if isinstance(c, PressureOutlet) else None)
        if callable(f):
            functions[p] = f
    if not functions:
        return None, None
    profiled = (C.c_int32 * 6)(*[1 if p in functions else 0 for p in range(6)])
    return profiled, _ProfileCallback(functions)


def pack_open_conditions(conditions: dict):
    """As pack_conditions with inlets and outlets: (kinds, velocity, pressure,
    total) in the boundary_kind codes 0 wall, 1 velocity inlet, 2 outlet with
    a backflow velocity, 3 symmetry, 4 outlet with normal backflow."""
    kinds, velocity = (C.c_int32 * 6)(), (C.c_double * 18)()
    pressure, total = (C.c_double * 6)(), (C.c_int32 * 6)()
    for p, name in enumerate(PATCHES):
        c = conditions[name]
        if isinstance(c, (Wall, VelocityInlet)):
            kinds[p] = 0 if isinstance(c, Wall) else 1
            if not callable(c.velocity):
                velocity[3 * p:3 * p + 3] = [float(v) for v in c.velocity]
        elif isinstance(c, Symmetry):
            kinds[p] = 3
        elif isinstance(c, MassFlowInlet):
            kinds[p] = 5
            pressure[p] = float(c.mass_flow)
            if c.direction is not None:
                velocity[3 * p:3 * p + 3] = [float(v) for v in c.direction]
        elif isinstance(c, PressureOutlet):
            kinds[p] = 4 if c.backflow_velocity is None else 2
            if c.backflow_velocity is not None and not callable(c.backflow_velocity):
                velocity[3 * p:3 * p + 3] = [float(v) for v in c.backflow_velocity]
            pressure[p] = float(c.pressure)
            total[p] = 1 if c.backflow_total else 0
        else:
            raise KuttaError(f"SIMPLEC: unsupported condition for {name}")
    return kinds, velocity, pressure, total


def initial_fields(centroids: np.ndarray, velocity, pressure):
    """(u, v, w, p) per cell in float64 from constants, functions of a
    centroid, or per-cell arrays."""
    n = len(centroids)

    def each(value, width):
        if callable(value):
            value = [value(tuple(x)) for x in centroids]
        out = np.asarray(value, np.float64)
        if out.ndim == (1 if width else 0):
            out = np.broadcast_to(out, (n, width) if width else (n,))
        if out.shape != ((n, width) if width else (n,)):
            raise KuttaError(f"SIMPLEC: initialization needs one value per cell ({n})")
        return out

    u, p = each(velocity, 3), each(pressure, 0)
    return tuple(np.ascontiguousarray(u[:, k]) for k in range(3)) + (np.ascontiguousarray(p),)


def cavity_conditions(lid: float = 1.0) -> dict:
    """validation/cases/cavity.hpp:113-118: lid (ymax) moving in x, walls at
    rest, symmetry front and back (zmin, zmax)."""
    c = {name: Wall() for name in PATCHES}
    c["ymax"] = Wall((lid, 0.0, 0.0))
    c["zmin"] = c["zmax"] = Symmetry()
    return c


REPORT_FIELDS = ("iteration", "u_raw", "u_scaled", "u_fluent_scaled", "u_fluent_normalization",
                 "v_raw", "v_scaled", "v_fluent_scaled", "v_fluent_normalization",
                 "w_raw", "w_scaled", "w_fluent_scaled", "w_fluent_normalization",
                 "continuity_raw", "continuity_scaled", "continuity_fluent_scaled",
                 "continuity_fluent_normalization", "corrected_mass_imbalance", "maximum_velocity",
                 "pseudo_time_step", "momentum_u_iterations", "momentum_v_iterations",
                 "momentum_w_iterations", "pressure_iterations", "pressure_final_residual",
                 "pressure_converged", "k_raw", "k_scaled", "k_fluent_scaled", "k_fluent_normalization",
                 "omega_raw", "omega_scaled", "omega_fluent_scaled", "omega_fluent_normalization",
                 "k_iterations", "omega_iterations")


class SteadySolver:
    """Steady laminar SIMPLEC on the MI355X (fv/steady_solver.hpp:286-360) over
    make_box_hex(nx, ny, nz) on [0, extent] (the unit cube by default) with
    the given patch conditions."""

    def __init__(self, device: Device, nx: int, ny: int, nz: int, conditions: dict,
                 settings: SolverSettings | None = None, extent=(1.0, 1.0, 1.0)):
        s = settings or SolverSettings()
        kinds, velocity, pressure, total = pack_open_conditions(conditions)
        h = C.c_void_p()
        box = (C.c_double * 3)(*[float(e) for e in extent])
        profiled, callback = profile_callback(conditions)
        if profiled is None:
            _check_host(host_lib().ka_simplec_create_open(device.handle, nx, ny, nz, box, kinds, velocity, pressure,
                                                          total, s.packed(), C.byref(h)))
        else:  # the callback runs inside the constructor only
            status = host_lib().ka_simplec_create_profiled(device.handle, nx, ny, nz, box, kinds, velocity,
                                                           pressure, total, profiled, callback.c_function, None,
                                                           s.packed(), C.byref(h))
            callback.reraise()
            _check_host(status)
        self.device, self.handle, self.settings = device, h.value, s

    @classmethod
    def from_mesh(cls, device: Device, mesh, conditions: dict, settings: SolverSettings | None = None,
                  sst: SstSettings | None = None):
        """A solver on a general mesh (kutta_amd.meshes.PolyMeshArrays),
        conditions by patch name; laminar, or k-omega SST with `sst`."""
        s = settings or SolverSettings()
        packed = pack_mesh_conditions(mesh, conditions)
        arrays = [np.ascontiguousarray(mesh.points, np.float64), np.ascontiguousarray(mesh.face_offsets, np.int32),
                  np.ascontiguousarray(mesh.face_nodes, np.int32), np.ascontiguousarray(mesh.owner, np.int32),
                  np.ascontiguousarray(mesh.neighbour, np.int32)]
        starts = np.array([p[1] for p in mesh.patches], np.int32)
        sizes = np.array([p[2] for p in mesh.patches], np.int32)
        sst_packed = sst.packed() if sst is not None else None
        h = C.c_void_p()
        vp = C.c_void_p
        _check_host(host_lib().ka_simplec_create_mesh(
            device.handle, len(mesh.points), arrays[0].ctypes.data, mesh.n_faces, arrays[1].ctypes.data,
            arrays[2].ctypes.data, arrays[3].ctypes.data, mesh.n_internal,
            arrays[4].ctypes.data if mesh.n_internal else None, mesh.n_cells, len(mesh.patches),
            "\n".join(p[0] for p in mesh.patches).encode(), starts.ctypes.data, sizes.ctypes.data,
            *[C.cast(packed[k], vp) for k in ("kinds", "velocity", "pressure", "total", "flags", "turbulence",
                                               "turbulence_ab")],
            C.cast(s.packed(), vp), C.cast(sst_packed, vp) if sst_packed is not None else None, C.byref(h)))
        self = cls.__new__(cls)
        self.device, self.handle, self.settings, self.sst = device, h.value, s, sst
        return self

    def initialize_fields(self, velocity=(0.0, 0.0, 0.0), pressure=0.0, k=None, omega=None) -> None:
        """SteadySolver::initialize_fields (fv/steady_solver.cpp:1493-1524):
        constants, functions of the cell centroid or per-cell arrays in the
        solver's cell order; k and omega are required under SST."""
        cent = self.centroids()
        u, v, w, p = initial_fields(cent, velocity, pressure)
        extra = []
        for value in (k, omega):
            if value is None:
                extra.append(None)
            else:
                extra.append(initial_fields(cent, (0.0, 0.0, 0.0), value)[3])
        _check_host(host_lib().ka_simplec_initialize_fields(
            self.handle, u.ctypes.data, v.ctypes.data, w.ctypes.data, p.ctypes.data,
            *[None if e is None else e.ctypes.data for e in extra]))

    def initialize(self, velocity=(0.0, 0.0, 0.0), pressure=0.0) -> None:
        """fv/steady_solver.cpp:1480-1491. `velocity` and `pressure` are
        constants, functions of the cell centroid (as the reference's), or
        per-cell arrays in the solver's cell order (that of centroids())."""
        u, v, w, p = initial_fields(self.centroids(), velocity, pressure)
        _check_host(host_lib().ka_simplec_initialize(self.handle, u.ctypes.data, v.ctypes.data, w.ctypes.data,
                                                     p.ctypes.data))

    def close(self) -> None:
        # After the device is gone its buffers cannot be freed through it:
        # drop the handle (the device's teardown released the memory).
        if self.handle:
            if self.device._alive:
                host_lib().ka_simplec_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def iterate(self, count: int = 1) -> None:
        _check_host(host_lib().ka_simplec_iterate(self.handle, count))

    def report(self, monitor: bool = False) -> dict:
        out = (C.c_double * len(REPORT_FIELDS))()
        _check_host(host_lib().ka_simplec_report(self.handle, 1 if monitor else 0, out))
        return dict(zip(REPORT_FIELDS, list(out)))

    def field(self, name: str) -> np.ndarray:
        n = host_lib().ka_simplec_field(self.handle, name.encode(), None)
        if n < 0:
            raise KuttaError(f"SIMPLEC: unknown field {name}")
        out = np.empty(n, np.float32)
        if host_lib().ka_simplec_field(self.handle, name.encode(), out.ctypes.data) < 0:
            raise KuttaError(host_lib().ka_last_error().decode())
        return out

    def forces(self, patches, reference_pressure: float = 0.0) -> dict:
        """Pressure and viscous force on the named wall patches
        (fv/steady_solver.cpp:1711-1736), FP64."""
        out = (C.c_double * 10)()
        _check_host(host_lib().ka_simplec_forces(self.handle, ",".join(patches).encode(),
                                                 float(reference_pressure), out))
        v = list(out)
        return {"pressure": v[0:3], "viscous": v[3:6], "area_vector": v[6:9], "reference_pressure": v[9]}

    @staticmethod
    def force_coefficient(report: dict, direction, density=1.0, velocity=1.0, area=1.0,
                          pressure=0.0) -> float:
        """ForceReport::coefficient (fv/steady_solver.cpp:1780-1784)."""
        packed = (C.c_double * 10)(*report["pressure"], *report["viscous"], *report["area_vector"],
                                   report["reference_pressure"])
        d = (C.c_double * 3)(*direction)
        out = C.c_double()
        _check_host(host_lib().ka_simplec_force_coefficient(packed, d, density, velocity, area, pressure,
                                                            C.byref(out)))
        return out.value

    def info(self) -> dict:
        out = (C.c_int32 * 4)()
        _check_host(host_lib().ka_simplec_info(self.handle, out))
        return {"cells": out[0], "reference": out[1], "host_waits": out[2], "iteration": out[3]}

    def sst_update(self, options: SstOptions, inputs) -> None:
        """SstModel::update (turbulence/sst.cpp:101-158) on this mesh; `inputs`
        are 25 device Buffers in SST_INPUTS order; read the results with
        field(name) for SST_FIELDS and SST_INT_FIELDS (int32 bits)."""
        if len(inputs) != len(SST_INPUTS):
            raise KuttaError(f"SST: {len(SST_INPUTS)} input arrays expected")
        ptrs = (C.c_void_p * len(inputs))(*[b.ptr for b in inputs])
        _check_host(host_lib().ka_simplec_sst_update(self.handle, options.packed(), ptrs))

    def set_graphs(self, on: bool) -> None:
        """HIP graph replay of the pressure PCG iterations (default on)."""
        _check_host(host_lib().ka_simplec_set_graphs(self.handle, int(bool(on))))

    def set_pressure_options(self, options: "AmgOptions") -> None:
        """Rebuilds the pressure AMG with other options (no reference
        counterpart), e.g. the K-cycle; structure SymmetricMMatrix."""
        packed = options.packed()
        _check_host(host_lib().ka_simplec_set_pressure_options(self.handle, C.cast(packed, C.c_void_p),
                                                               float(options.strength_threshold)))

    def set_batched_momentum(self, on: bool) -> bool:
        """Solve u, v, w in one batched BiCGStab (on by default where it is
        bitwise the sequential order); returns whether it is in use."""
        used = C.c_int32()
        _check_host(host_lib().ka_simplec_set_batched_momentum(self.handle, int(bool(on)), C.byref(used)))
        return bool(used.value)

    def pressure_levels(self) -> list:
        rows, nnz = (C.c_int32 * 64)(), (C.c_int32 * 64)()
        count = host_lib().ka_simplec_pressure_levels(self.handle, rows, nnz, 64)
        if count < 0:
            raise KuttaError(host_lib().ka_last_error().decode())
        return [(rows[i], nnz[i]) for i in range(min(count, 64))]

    def pressure_plan(self) -> str:
        buf = C.create_string_buffer(2048)
        if host_lib().ka_simplec_pressure_plan(self.handle, buf, 2048) < 0:
            raise KuttaError(host_lib().ka_last_error().decode())
        return buf.value.decode()

    def centroids(self) -> np.ndarray:
        n = self.info()["cells"]
        out = np.empty((n, 3), np.float64)
        _check_host(host_lib().ka_simplec_centroids(self.handle, out.ctypes.data))
        return out


# ------------------------------------------------- learned coarse correction
class LearnedNetwork:
    """Message-passing network over a coarse operator's graph (no reference
    counterpart): hidden width 64, `layers` layers, bf16 MFMA layers with FP32
    accumulation, trained with Adam on the device. Vectors are device Buffers
    of the graph's node count."""

    HIDDEN = 64

    def __init__(self, device: Device, layers: int = 3, seed: int = 0):
        h = C.c_void_p()
        _check_host(host_lib().ka_learned_create(device.handle, int(layers), int(seed), C.byref(h)))
        self.device = device
        self.handle = h.value
        self.layers = int(layers)
        self.count = int(host_lib().ka_learned_parameter_count(int(layers)))

    @staticmethod
    def parameter_count(layers: int) -> int:
        return int(host_lib().ka_learned_parameter_count(int(layers)))

    def set_graph(self, a: CsrMatrix) -> None:
        _check_host(host_lib().ka_learned_set_graph(self.handle, a.handle))

    def set_parameters(self, theta) -> None:
        t = np.ascontiguousarray(theta, dtype=np.float32)
        _check_host(host_lib().ka_learned_set_parameters(self.handle, t.ctypes.data, t.size))

    def parameters(self) -> np.ndarray:
        out = np.empty(self.count, np.float32)
        _check_host(host_lib().ka_learned_parameters(self.handle, out.ctypes.data, out.size))
        return out

    def infer(self, r: Buffer, e: Buffer) -> None:
        _check_host(host_lib().ka_learned_infer(self.handle, r.ptr, e.ptr))

    def gradient(self, r: Buffer, target: Buffer, nodes: int):
        """(loss, gradient, normalised outputs y) at the current parameters."""
        grad = np.empty(self.count, np.float32)
        y = np.empty(nodes, np.float32)
        loss = (C.c_double * 2)()
        _check_host(host_lib().ka_learned_gradient(self.handle, r.ptr, target.ptr, grad.ctypes.data,
                                                   y.ctypes.data, loss))
        return loss[0], grad, y

    def train(self, r: Buffer, target: Buffer, pairs: int, steps: int, lr: float = 1e-3,
              relative: bool = False):
        """Losses before each Adam step (and, with relative=True, the
        unexplained fractions as a second array)."""
        losses = np.empty(max(int(steps), 1), np.float64)
        rel = np.empty(max(int(steps), 1), np.float64)
        _check_host(host_lib().ka_learned_train(self.handle, r.ptr, target.ptr, int(pairs), int(steps),
                                                float(lr), losses.ctypes.data, rel.ctypes.data))
        return (losses[:int(steps)], rel[:int(steps)]) if relative else losses[:int(steps)]

    def close(self) -> None:
        if self.handle:
            if self.device._alive:
                host_lib().ka_learned_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
