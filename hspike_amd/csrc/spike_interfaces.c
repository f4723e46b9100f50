/*
 * spike_interfaces.c -- the reference's top-level entry points (OpenMP/spike_interfaces.h:28-72)
 * on the GPU SPIKE solver.
 *
 * Reference behaviour kept: CSR int32 0-based, B and X column-major n x nrhs with leading
 * dimension n, the caller owns all arrays and the driver never modifies aij or bij.
 * Reference behaviour fixed: s/c/z are implemented (stubs in spike_interfaces.c:3-39); every
 * column of X is written (spike_matrix.c:1339-1345 copies column 0); failures return negative
 * codes instead of abort() (spike_interfaces.c:53-56 and the backends).
 *
 * Environment: HSPIKE_PARTITIONS=<p> fixes the partition count (default: autotuned; the
 * reference uses 4, spike_analysis.c:168); HSPIKE_VERBOSE=1 prints per-solve statistics;
 * HSPIKE_TOL overrides the 1e-10 relative-residual target.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "spike_interfaces.h"
#include "spike_ext.h"

static void opts_from_env(spike_opts_t *o, int precision)
{
    spike_default_opts(o);
    o->precision = precision;
    const char *v = getenv("HSPIKE_PARTITIONS");
    if (v && atoi(v) > 0) o->partitions = atoi(v);
    v = getenv("HSPIKE_VERBOSE");
    if (v && atoi(v) > 0) o->verbose = 1;
    v = getenv("HSPIKE_TOL");
    if (v && atof(v) > 0) o->tol = atof(v);
}

static int check_args(integer_t n, integer_t nnz, integer_t nrhs, const integer_t *colind,
                      const integer_t *rowptr, const void *aij, const void *x, const void *b)
{
    if (n <= 0 || nnz < 0 || nrhs < 0 || !rowptr || (nnz > 0 && (!colind || !aij)) || (nrhs > 0 && (!x || !b)))
        return SPIKE_ERR_INVALID_ARG;
    return matrix_CheckCSR(n, (uLong_t)nnz, colind, rowptr);
}

/* solve with host arrays of double / complex16 values, in column blocks of `colblock` */
static Error_t solve_host(int dtype, int precision, integer_t n, integer_t nnz, integer_t nrhs,
                          const integer_t *colind, const integer_t *rowptr, const void *aij,
                          void *x, const void *b, int colblock)
{
    spike_opts_t o;
    opts_from_env(&o, precision);
    spike_handle_t *h = NULL;
    int e = spike_create(&h, &o);
    if (e) return e;
    e = spike_set_matrix_host(h, dtype, n, nnz, rowptr, colind, aij);
    if (!e) e = spike_factor(h);
    const size_t hs = dtype == SPIKE_DTYPE_Z ? 16 : 8;
    for (int c0 = 0; !e && c0 < nrhs; c0 += colblock) {
        const int m = nrhs - c0 < colblock ? nrhs - c0 : colblock;
        e = spike_solve_host(h, m, (const char *)b + (size_t)c0 * n * hs, n, (char *)x + (size_t)c0 * n * hs, n);
    }
    spike_destroy(h);
    return e;
}

Error_t dspike_core_host(const integer_t n, const integer_t nnz, const integer_t nrhs, integer_t *colind,
                         integer_t *rowptr, double *aij, double *xij, double *bij)
{
    int e = check_args(n, nnz, nrhs, colind, rowptr, aij, xij, bij);
    if (e) return e;
    return solve_host(SPIKE_DTYPE_D, SPIKE_PREC_FP64, n, nnz, nrhs, colind, rowptr, aij, xij, bij, nrhs > 0 ? nrhs : 1);
}

Error_t dspike_core_host_blocking(const integer_t n, const integer_t nnz, const integer_t nrhs, integer_t *colind,
                                  integer_t *rowptr, double *aij, double *xij, double *bij)
{
    int e = check_args(n, nnz, nrhs, colind, rowptr, aij, xij, bij);
    if (e) return e;
    return solve_host(SPIKE_DTYPE_D, SPIKE_PREC_FP64, n, nnz, nrhs, colind, rowptr, aij, xij, bij, 200);
}

Error_t zspike_core_host(const integer_t n, const integer_t nnz, const integer_t nrhs, integer_t *colind,
                         integer_t *rowptr, complex16 *aij, complex16 *xij, complex16 *bij)
{
    int e = check_args(n, nnz, nrhs, colind, rowptr, aij, xij, bij);
    if (e) return e;
    return solve_host(SPIKE_DTYPE_Z, SPIKE_PREC_FP64, n, nnz, nrhs, colind, rowptr, aij, xij, bij, nrhs > 0 ? nrhs : 1);
}

/* FP32 interfaces: FP32 factorization and sweeps, FP64 residual and correction, FP32 result */
Error_t sspike_core_host(const integer_t n, const integer_t nnz, const integer_t nrhs, integer_t *colind,
                         integer_t *rowptr, float *aij, float *xij, float *bij)
{
    int e = check_args(n, nnz, nrhs, colind, rowptr, aij, xij, bij);
    if (e) return e;
    double *a = (double *)malloc(sizeof(double) * (nnz ? nnz : 1));
    double *b = (double *)malloc(sizeof(double) * ((size_t)n * nrhs + 1));
    double *x = (double *)malloc(sizeof(double) * ((size_t)n * nrhs + 1));
    if (!a || !b || !x) { free(a); free(b); free(x); return SPIKE_ERR_OOM; }
    for (integer_t i = 0; i < nnz; ++i) a[i] = aij[i];
    for (size_t i = 0; i < (size_t)n * nrhs; ++i) b[i] = bij[i];
    e = solve_host(SPIKE_DTYPE_D, SPIKE_PREC_MIXED, n, nnz, nrhs, colind, rowptr, a, x, b, nrhs > 0 ? nrhs : 1);
    if (!e || e == SPIKE_ERR_NOT_CONVERGED)
        for (size_t i = 0; i < (size_t)n * nrhs; ++i) xij[i] = (float)x[i];
    free(a); free(b); free(x);
    return e;
}

Error_t cspike_core_host(const integer_t n, const integer_t nnz, const integer_t nrhs, integer_t *colind,
                         integer_t *rowptr, complex8 *aij, complex8 *xij, complex8 *bij)
{
    int e = check_args(n, nnz, nrhs, colind, rowptr, aij, xij, bij);
    if (e) return e;
    complex16 *a = (complex16 *)malloc(sizeof(complex16) * (nnz ? nnz : 1));
    complex16 *b = (complex16 *)malloc(sizeof(complex16) * ((size_t)n * nrhs + 1));
    complex16 *x = (complex16 *)malloc(sizeof(complex16) * ((size_t)n * nrhs + 1));
    if (!a || !b || !x) { free(a); free(b); free(x); return SPIKE_ERR_OOM; }
    for (integer_t i = 0; i < nnz; ++i) { a[i].real = aij[i].real; a[i].imag = aij[i].imag; }
    for (size_t i = 0; i < (size_t)n * nrhs; ++i) { b[i].real = bij[i].real; b[i].imag = bij[i].imag; }
    e = solve_host(SPIKE_DTYPE_Z, SPIKE_PREC_MIXED, n, nnz, nrhs, colind, rowptr, a, x, b, nrhs > 0 ? nrhs : 1);
    if (!e || e == SPIKE_ERR_NOT_CONVERGED)
        for (size_t i = 0; i < (size_t)n * nrhs; ++i) { xij[i].real = (float)x[i].real; xij[i].imag = (float)x[i].imag; }
    free(a); free(b); free(x);
    return e;
}
