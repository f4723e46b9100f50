/*
 * spike_hip.h -- the HIP backend plugin: the reference's direct-solver handle API
 * (OpenMP/spike_superlu.h:107-135, spike_pardiso.h:87-115, spike_cuda.h:125-171) backed by
 * this library's GPU banded solver, so a driver written against that plugin slot
 * (selected in the reference by -D_*_BACKEND_, spike_algebra.h:30-36) gets GPU banded
 * factorizations and multi-RHS solves by compiling against this header instead.
 *
 * Semantics kept from the reference:
 *   - CreateHandler / Configure(max_nrhs) / Factorize(CSR) / SolveForRHS(nrhs, x, b) /
 *     ShowStatistics / Finalize (frees the handle) / Solve (one-shot);
 *   - CSR int32 0-based, x and b dense column-major with leading dimension n, host memory;
 *   - the factorization of A is kept for any number of later SolveForRHS calls.
 * Changes: values are FP64 (the reference's complex_t under _DATATYPE_D_); the factors live
 * in HBM; failures return negative SPIKE_ERR_* codes (spike_matrix.h) instead of aborting; the
 * handle owns copies of the CSR arrays (the reference keeps the caller's pointers,
 * spike_pardiso.c:93-97, which this backend never needs after Factorize returns).
 *
 * The header is self-contained so that it can take a backend's place in the reference's own
 * translation units (spike_algebra.h:30-36 includes it after spike_datatypes.h): it declares
 * no container types, its scalar typedefs are the reference's (int) and are skipped when
 * spike_datatypes.h or this library's spike_matrix.h already made them, and the handle is an
 * opaque struct.
 */
#ifndef HSPIKE_SPIKE_HIP_H
#define HSPIKE_SPIKE_HIP_H

#if !defined(_SPIKE_DATATYPES_H_) && !defined(HSPIKE_SPIKE_MATRIX_H) && !defined(HSPIKE_SCALAR_TYPES)
#define HSPIKE_SCALAR_TYPES
typedef int integer_t;   /* spike_datatypes.h:86 */
typedef int Error_t;     /* spike_datatypes.h:87 */
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct DirectSolverHander_s DirectSolverHander_t;   /* opaque */

/* spike_superlu.h:107 */
DirectSolverHander_t *directSolver_CreateHandler(void);

/* spike_superlu.h:109: max_nrhs sizes the RHS staging buffers (spike_cuda.c:549-587) */
Error_t directSolver_Configure(DirectSolverHander_t *handler, const integer_t max_nrhs);

/* spike_superlu.h:112-117: banded LU of the n x n CSR matrix on the GPU */
Error_t directSolver_Factorize(DirectSolverHander_t *handler, const integer_t n, const integer_t nnz,
                               integer_t *colind, integer_t *rowptr, double *aij);

/* spike_superlu.h:119-122: x = A^{-1} b for nrhs columns */
Error_t directSolver_SolveForRHS(DirectSolverHander_t *handler, const integer_t nrhs,
                                 double *xij, double *bij);

/* spike_superlu.h:124: prints factor/solve timings and RHS counts to stderr */
Error_t directSolver_ShowStatistics(DirectSolverHander_t *handler);

/* spike_superlu.h:126: releases device memory and the handle */
Error_t directSolver_Finalize(DirectSolverHander_t *handler);

/* spike_superlu.h:128-135: factor + solve + release */
Error_t directSolver_Solve(integer_t n, integer_t nnz, integer_t nrhs, integer_t *colind,
                           integer_t *rowptr, double *aij, double *xij, double *bij);

/* spike_algebra.h:157-164: solver for the reduced system R (CSR, FP64) used by the driver's
 * stage 3 (spike_interfaces.c:245); here any banded CSR matrix is factored on the GPU by a band
 * LU with partial pivoting and refined in FP64 (spike_band_solve_host). */
Error_t directSolver_Host_Solve(integer_t n, integer_t nnz, integer_t nrhs, integer_t *colind,
                                integer_t *rowptr, double *aij, double *xij, double *bij);

/* spike_algebra.c:97-151: prints ||A x - b||, ||b|| and the relative residual of the first
 * five columns to stderr, as the reference does; computed in FP64 on the GPU.  The variant
 * that returns every column's value is spike_residual_host (spike_ext.h). */
Error_t ComputeResidualOfLinearSystem(integer_t *colind, integer_t *rowptr, double *aij,
                                      double *x, double *b, integer_t n, integer_t nrhs);

#ifdef __cplusplus
}
#endif
#endif
