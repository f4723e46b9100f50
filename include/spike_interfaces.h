/*
 * spike_interfaces.h -- top-level SPIKE entry points, same symbols and signatures as the
 * reference's OpenMP/spike_interfaces.h:28-72.
 *
 * Each call solves A X = B for a square n x n CSR matrix A (int32, 0-based) with nrhs
 * right-hand sides; B and X are dense column-major n x nrhs with leading dimension n.
 * The caller owns every array; B and the CSR are read only, X is written.
 *
 * Differences from the reference, all deliberate:
 *   - s/c/z are implemented (the reference's are stubs returning success, spike_interfaces.c:3-39);
 *   - every column of X is correct (the reference's write-back copies column 0 into every
 *     column, spike_matrix.c:1339-1345);
 *   - errors come back as negative codes from spike_matrix.h instead of abort()/exit();
 *   - all numerical work runs on the GPU (HIP, gfx950); host arrays are staged to HBM.
 *
 * Precision: d/z solve in FP64 and refine in FP64 until every column's relative residual
 * ||A x - b|| / ||b|| is <= 1e-10 (or the refinement cap is hit: SPIKE_ERR_NOT_CONVERGED).
 * s/c factor in FP32 with FP64 residuals and corrections, and return the FP32-rounded answer.
 * The environment variable HSPIKE_PARTITIONS=p fixes the partition count (default: autotuned).
 */
#ifndef HSPIKE_SPIKE_INTERFACES_H
#define HSPIKE_SPIKE_INTERFACES_H

#include "spike_matrix.h"

#ifdef __cplusplus
extern "C" {
#endif

/* spike_interfaces.h:28-35 */
Error_t sspike_core_host(const integer_t n, const integer_t nnz, const integer_t nrhs,
                         integer_t *colind, integer_t *rowptr,
                         float *aij, float *xij, float *bij);

/* spike_interfaces.h:37-44 */
Error_t dspike_core_host(const integer_t n, const integer_t nnz, const integer_t nrhs,
                         integer_t *colind, integer_t *rowptr,
                         double *aij, double *xij, double *bij);

/* spike_interfaces.h:46-53 */
Error_t cspike_core_host(const integer_t n, const integer_t nnz, const integer_t nrhs,
                         integer_t *colind, integer_t *rowptr,
                         complex8 *aij, complex8 *xij, complex8 *bij);

/* spike_interfaces.h:55-62 */
Error_t zspike_core_host(const integer_t n, const integer_t nnz, const integer_t nrhs,
                         integer_t *colind, integer_t *rowptr,
                         complex16 *aij, complex16 *xij, complex16 *bij);

/* spike_interfaces.h:65-72: RHS columns are processed in blocks of 200 columns
 * (blockingDistance, spike_analysis.c:51) against factors kept resident in HBM. */
Error_t dspike_core_host_blocking(const integer_t n, const integer_t nnz, const integer_t nrhs,
                                  integer_t *colind, integer_t *rowptr,
                                  double *aij, double *xij, double *bij);

#ifdef __cplusplus
}
#endif
#endif
