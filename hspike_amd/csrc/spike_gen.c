/*
 * spike_gen.c -- deterministic synthetic banded systems for tests and benchmarks.
 *
 * Random numbers: splitmix64 (Steele, Lea, Flood 2014), one stream per row seeded from
 * (seed, row), so generation is parallel and a row's values never depend on the thread count.
 * Off-diagonal entries in the band are present with probability `fill` and uniform on [-1, 1]
 * (complex: real and imaginary parts independent U[-1,1]); the diagonal is
 *   max(d_r * sum_j |a_rj| + diag_add, diag_floor),  d_r ~ U[dom_lo, dom_hi],
 * so rows without off-diagonals still get a nonzero pivot (diag_floor).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "spike_ext.h"

static inline uint64_t sm64_next(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static inline double sm64_u01(uint64_t *s) { return (double)(sm64_next(s) >> 11) * (1.0 / 9007199254740992.0); }

static inline uint64_t row_seed(uint64_t seed, int64_t row)
{
    uint64_t s = seed ^ (0xD1342543DE82EF95ULL * (uint64_t)(row + 1));
    (void)sm64_next(&s);
    return s;
}

/* Generates row r into (cols, vals); returns its nnz.  cols/vals may be NULL (count only);
 * the random stream is consumed identically either way. */
static int gen_row(int dtype, int n, int ku, int kl, double fill, double dom_lo, double dom_hi, double diag_add,
                   double diag_floor, uint64_t seed, int r, int *cols, double *vals)
{
    uint64_t s = row_seed(seed, r);
    const int cz = dtype == SPIKE_DTYPE_Z;
    int q = 0, dpos = -1;
    double sum = 0.0;
    for (int c = r - kl; c <= r + ku; ++c) {
        if (c < 0 || c >= n) continue;
        if (c == r) {
            dpos = q;
            if (cols) cols[q] = c;
            ++q;
            continue;
        }
        const double pres = sm64_u01(&s);
        const double vr = 2.0 * sm64_u01(&s) - 1.0;
        const double vi = cz ? 2.0 * sm64_u01(&s) - 1.0 : 0.0;
        if (fill < 1.0 && pres >= fill) continue;
        if (cols) {
            cols[q] = c;
            if (cz) { vals[2 * q] = vr; vals[2 * q + 1] = vi; }
            else vals[q] = vr;
        }
        sum += cz ? hypot(vr, vi) : fabs(vr);
        ++q;
    }
    const double d = dom_lo + (dom_hi - dom_lo) * sm64_u01(&s);
    double dv = d * sum + diag_add;
    if (dv < diag_floor) dv = diag_floor;
    if (cols && dpos >= 0) {
        if (cz) { vals[2 * dpos] = dv; vals[2 * dpos + 1] = 0.0; }
        else vals[dpos] = dv;
    }
    return q;
}

int spike_gen_banded(int dtype, int n, int ku, int kl, double fill, double dom_lo, double dom_hi, double diag_add,
                     double diag_floor, uint64_t seed, int *rowptr, int *colind, void *aij, int64_t *nnz_out)
{
    if (n <= 0 || ku < 0 || kl < 0 || (dtype != SPIKE_DTYPE_D && dtype != SPIKE_DTYPE_Z)) return SPIKE_ERR_INVALID_ARG;
    int *cnt = (int *)malloc(sizeof(int) * (size_t)n);
    if (!cnt) return SPIKE_ERR_OOM;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n; ++r)
        cnt[r] = gen_row(dtype, n, ku, kl, fill, dom_lo, dom_hi, diag_add, diag_floor, seed, r, NULL, NULL);
    int64_t nnz = 0;
    for (int r = 0; r < n; ++r) nnz += cnt[r];
    if (nnz_out) *nnz_out = nnz;
    if (nnz > INT32_MAX) { free(cnt); return SPIKE_ERR_UNSUPPORTED; }
    if (rowptr && colind && aij) {
        rowptr[0] = 0;
        for (int r = 0; r < n; ++r) rowptr[r + 1] = rowptr[r] + cnt[r];
        const int w = dtype == SPIKE_DTYPE_Z ? 2 : 1;
#pragma omp parallel for schedule(static)
        for (int r = 0; r < n; ++r)
            gen_row(dtype, n, ku, kl, fill, dom_lo, dom_hi, diag_add, diag_floor, seed, r, colind + rowptr[r],
                    (double *)aij + (size_t)w * rowptr[r]);
    }
    free(cnt);
    return SPIKE_SUCCESS;
}

int spike_gen_dense(int dtype, int64_t count, double lo, double hi, uint64_t seed, void *out)
{
    if (count < 0 || !out) return SPIKE_ERR_INVALID_ARG;
    const int w = dtype == SPIKE_DTYPE_Z ? 2 : 1;
    double *o = (double *)out;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < count * w; ++i) {
        uint64_t s = row_seed(seed, i);
        o[i] = lo + (hi - lo) * sm64_u01(&s);
    }
    return SPIKE_SUCCESS;
}
