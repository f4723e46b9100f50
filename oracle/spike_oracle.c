/*
 * spike_oracle.c -- CPU FP64 oracle for the SPIKE banded-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the package's
 * smoke() entry and bench.py's cpu_baseline leg may load it, and only to check or time the
 * GPU library against it.  Nothing in the product (hspike_amd/) links or calls it.
 *
 * It restates the reference's algorithm (srodrb/HSPIKE, OpenMP/spike_interfaces.c:43-419):
 * same partition boundaries, explicit y/V/W spike solves per partition, the reference's
 * reduced-system CSR pattern and tip layout, and recovery by re-solving each partition.
 * SuperLU_MT / PARDISO (not vendored, not buildable here) are replaced by a banded LU with
 * partial pivoting (LAPACK gbtf2/gbtrs semantics); the reduced system is solved with the
 * same band LU.  Partitions run in parallel with OpenMP (the reference runs them serially).
 *
 * Parity pins (tests/test_oracle.py): Tests/spike/penta_10.readme (R and g, P=2),
 * Tests/solve_small/readme.txt (x), Tests/split/block{1,2}.bin (extraction),
 * Tests/bandwidth/readme.txt (ku=2, kl=4), scipy splu/solve_banded golden vectors,
 * dense brute-force inverses for the spike tips, and the residual closed form.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <complex.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* spike_algebra.c:56-95: ku = max(col - row), kl = max(row - col) over the nonzeros */
void or_bandwidth(int n, const int *rowptr, const int *colind, int *ku, int *kl)
{
    int u = 0, l = 0;
    for (int r = 0; r < n; ++r)
        for (int idx = rowptr[r]; idx < rowptr[r + 1]; ++idx) {
            int d = colind[idx] - r;
            if (d > u) u = d;
            if (-d > l) l = -d;
        }
    *ku = u; *kl = l;
}

/* spike_analysis.c:38-76: rows[i] = i*floor(n/P) for i < P, rows[P] = n (the remainder goes
 * to the last partition); red[i] = sum_{j<i} (ku + kl). */
void or_schedule(int n, int P, int ku, int kl, int *rows, int *red)
{
    int nreg = n / P;
    for (int i = 0; i < P; ++i) { rows[i] = i * nreg; red[i] = i * (ku + kl); }
    rows[P] = n; red[P] = P * (ku + kl);
}

/* spike_matrix.c:1379-1412: nnz of the reduced system R */
int or_reduced_nnz(int P, int ku, int kl)
{
    const int k2 = ku + kl;
    if (P == 1) return k2;
    return k2 * (ku + 1) + (P - 2) * k2 * (k2 + 1) + k2 * (kl + 1);
}

/* spike_matrix.c:563-645: CSR sub-matrix M[r0:rf, c0:cf] with shifted column indices.
 * Pass NULL outputs to count; returns nnz of the block. */
int or_extract_csr(const int *rowptr, const int *colind, const double *aij,
                   int r0, int rf, int c0, int cf, int *orow, int *ocol, double *oaij)
{
    int q = 0;
    if (orow) orow[0] = 0;
    for (int r = r0; r < rf; ++r) {
        for (int idx = rowptr[r]; idx < rowptr[r + 1]; ++idx) {
            int c = colind[idx];
            if (c >= c0 && c < cf) {
                if (ocol) { ocol[q] = c - c0; oaij[q] = aij[idx]; }
                ++q;
            }
        }
        if (orow) orow[r - r0 + 1] = q;
    }
    return q;
}

static double or_abs_d(double x) { return fabs(x); }
static double or_abs_z(double _Complex x) { return cabs(x); }

#define OR_T double
#define OR_S d
#define OR_ABS or_abs_d
#include "spike_oracle_body.h"
#undef OR_T
#undef OR_S
#undef OR_ABS

#define OR_T double _Complex
#define OR_S z
#define OR_ABS or_abs_z
#include "spike_oracle_body.h"
#undef OR_T
#undef OR_S
#undef OR_ABS

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
