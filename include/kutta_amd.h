/* kutta_amd.h — C ABI of the MI355X host library (libkutta_amd.so): the
 * HIP-native pressure solver and the mesh/operator setup of Kutta's SIMPLEC
 * pressure-correction path, built on the kutta_hip device ABI.
 *
 * Every int-returning function returns 0 on success, 2 for a non-finite
 * (diverged) input, 1 for any other error; ka_last_error() describes it.
 */
#ifndef KUTTA_AMD_H
#define KUTTA_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "kutta_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

const char* ka_last_error(void);

/* A device (one HIP stream). ka_device_raw exposes the kutta_hip handle for
 * buffer management through kh_alloc/kh_upload/... on the same stream. */
int ka_device_create(int ordinal, void** out);
void ka_device_destroy(void* dev);
kh_device* ka_device_raw(void* dev);

/* Host CSR operator (linalg/csr.hpp:16-28): columns sorted, diagonal present,
 * double values. */
int ka_csr_create(int n_rows, const int32_t* row_ptr, const int32_t* col, const int32_t* diag,
                  const double* val, void** out);
void ka_csr_destroy(void* csr);
int ka_csr_rows(void* csr);
int ka_csr_nnz(void* csr);
int ka_csr_copy(void* csr, int32_t* row_ptr, int32_t* col, int32_t* diag, double* val);

/* Host aggregation of one level (linalg/amg.cpp:160-174), no device needed.
 * options as for ka_solver_create. Arrays (which): 0 aggregate,
 * 1 member_offsets, 2 members, 3 entry_offsets, 4 entries, 5 coarse row_ptr,
 * 6 coarse col, 7 coarse diag. ka_coarsening_values gives the coarse values. */
int ka_aggregate(void* csr, const int32_t* options, double threshold, void** out);
long ka_coarsening_array(void* coarsening, int which, int32_t* out);
long ka_coarsening_values(void* coarsening, double* out);
void ka_coarsening_destroy(void* coarsening);

/* Device CSR (linalg/csr.hpp:31-48): upload(Device&, CsrMatrixHost&). The value
 * array is the part that changes between solves. */
int ka_matrix_upload(void* dev, void* csr, void** out);
void ka_matrix_destroy(void* matrix);
float* ka_matrix_values(void* matrix);
const int32_t* ka_matrix_row_ptr(void* matrix);
const int32_t* ka_matrix_col(void* matrix);
const int32_t* ka_matrix_diag(void* matrix);

/* Solvers (linalg/solver.hpp:237-293, linalg/amg.hpp:8-27).
 * kind: 0 Jacobi, 1 PcgJacobi, 2 BicgstabJacobi, 3 AmgPcg, 4 AmgBicgstab.
 * options[8]: pre_sweeps, post_sweeps, coarse_sweeps, coarse_rows, max_levels,
 *             aggregation_passes, nullspace (0 none, 1 constant),
 *             structure (0 general, 1 symmetric M-matrix). */
int ka_solver_create(void* dev, int kind, void* csr, const int32_t* options, double threshold,
                     void** out);
void ka_solver_destroy(void* solver);
int ka_solver_refresh(void* solver, void* matrix);
/* x and b are device vectors of n_rows floats. stats[4]: iterations,
 * initial residual, final residual, converged. */
int ka_solver_solve(void* solver, void* matrix, float* x, const float* b, double rel_tol,
                    double abs_tol, int max_iter, int interval, double* stats);
int ka_solver_describe(void* solver, char* out, int capacity);
int ka_solver_levels(void* solver, int32_t* rows, int32_t* nnz, int capacity);
/* which: 0 aggregate, 1 member_offsets, 2 members, 3 entry_offsets, 4 entries.
 * Returns the length (copies when out is non-null), or -1. */
long ka_solver_transfer(void* solver, int level, int which, int32_t* out);
long ka_solver_level_values(void* solver, int level, float* out);
/* 1: fused MI355X kernels (default); 0: the reference kernel sequence. */
int ka_solver_set_fused(void* solver, int fused);
int ka_solver_host_waits(void* solver);

#ifdef __cplusplus
}
#endif

#endif /* KUTTA_AMD_H */
