/*
 * spike_matrix.c -- host-side containers of the SPIKE path (spike_matrix.h).
 *
 * CSR file format of the reference (OpenMP/spike_matrix.c:11-51):
 *   int32 n, int32 nnz, int32 dtype (0 real, 1 complex), aij[nnz], int32 colind[nnz],
 *   int32 rowptr[n+1].
 * Complex files written by the reference's generator (generator_banded.py:134-141) store all
 * real parts, then all imaginary parts.  The reference's loader reads interleaved complex_t
 * (spike_matrix.c:34-35) and so misreads those files; matrix_LoadCSR_z reads the planar layout
 * its writer produces, matrix_LoadCSR_zl takes the layout explicitly (both layouts have the same
 * file size, so it cannot be inferred from the file).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "spike_matrix.h"

static void *xmalloc(size_t n) { return malloc(n ? n : 1); }
static void *xcalloc(size_t n, size_t s) { return calloc(n ? n : 1, s ? s : 1); }

/* ------------------------------------------------------------------ validation */
Error_t matrix_CheckCSR(integer_t n, uLong_t nnz, const integer_t *colind, const integer_t *rowptr)
{
    if (n <= 0 || !rowptr) return SPIKE_ERR_INVALID_ARG;
    if (nnz > 0 && !colind) return SPIKE_ERR_INVALID_ARG;
    if (rowptr[0] != 0 || (uLong_t)rowptr[n] != nnz) return SPIKE_ERR_MALFORMED_CSR;
    /* monotone row pointers from 0 to nnz cover the entries [0, nnz) exactly, so the column
       indices are checked flat; both scans run on all host cores for large matrices */
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad) if (n > 65536)
    for (integer_t r = 0; r < n; ++r) bad |= rowptr[r + 1] < rowptr[r];
    if (bad) return SPIKE_ERR_MALFORMED_CSR;
    const long long m = (long long)nnz;
#pragma omp parallel for schedule(static) reduction(| : bad) if (m > 262144)
    for (long long i = 0; i < m; ++i) bad |= (colind[i] < 0) | (colind[i] >= n);
    return bad ? SPIKE_ERR_MALFORMED_CSR : SPIKE_SUCCESS;
}

/* ------------------------------------------------------------------ file IO */
static int read_header(FILE *f, integer_t hdr[3])
{
    return fread(hdr, sizeof(integer_t), 3, f) == 3 ? 0 : -1;
}

matrix_t *matrix_LoadCSR(const char *filename)
{
    FILE *f = fopen(filename, "rb");
    if (!f) return NULL;
    integer_t hdr[3];
    matrix_t *M = NULL;
    if (read_header(f, hdr) || hdr[0] <= 0 || hdr[1] < 0 || hdr[2] != 0) goto fail;
    M = (matrix_t *)xcalloc(1, sizeof(matrix_t));
    M->n = hdr[0];
    M->nnz = (uLong_t)hdr[1];
    M->aij = (double *)xmalloc(sizeof(double) * M->nnz);
    M->colind = (integer_t *)xmalloc(sizeof(integer_t) * M->nnz);
    M->rowptr = (integer_t *)xmalloc(sizeof(integer_t) * (M->n + 1));
    if (fread(M->aij, sizeof(double), M->nnz, f) != M->nnz) goto fail;
    if (fread(M->colind, sizeof(integer_t), M->nnz, f) != M->nnz) goto fail;
    if (fread(M->rowptr, sizeof(integer_t), (size_t)M->n + 1, f) != (size_t)M->n + 1) goto fail;
    fclose(f);
    if (matrix_CheckCSR(M->n, M->nnz, M->colind, M->rowptr) != SPIKE_SUCCESS) { matrix_Deallocate(M); return NULL; }
    matrix_ComputeBandwidth(M->n, M->colind, M->rowptr, M->aij, &M->ku, &M->kl);
    M->type = _DIAG_BLOCK_;
    return M;
fail:
    fclose(f);
    if (M) matrix_Deallocate(M);
    return NULL;
}

Error_t matrix_SaveCSR(const char *filename, const matrix_t *M)
{
    if (!M || !filename) return SPIKE_ERR_INVALID_ARG;
    FILE *f = fopen(filename, "wb");
    if (!f) return SPIKE_ERR_IO;
    integer_t hdr[3] = {M->n, (integer_t)M->nnz, 0};
    int ok = fwrite(hdr, sizeof(integer_t), 3, f) == 3 &&
             fwrite(M->aij, sizeof(double), M->nnz, f) == M->nnz &&
             fwrite(M->colind, sizeof(integer_t), M->nnz, f) == M->nnz &&
             fwrite(M->rowptr, sizeof(integer_t), (size_t)M->n + 1, f) == (size_t)M->n + 1;
    fclose(f);
    return ok ? SPIKE_SUCCESS : SPIKE_ERR_IO;
}

matrix_z_t *matrix_LoadCSR_zl(const char *filename, int planar)
{
    FILE *f = fopen(filename, "rb");
    if (!f) return NULL;
    integer_t hdr[3];
    matrix_z_t *M = NULL;
    double *raw = NULL;
    if (read_header(f, hdr) || hdr[0] <= 0 || hdr[1] < 0) goto fail;
    M = (matrix_z_t *)xcalloc(1, sizeof(matrix_z_t));
    M->n = hdr[0];
    M->nnz = (uLong_t)hdr[1];
    raw = (double *)xmalloc(sizeof(double) * 2 * M->nnz);
    M->aij = (complex16 *)xmalloc(sizeof(complex16) * M->nnz);
    M->colind = (integer_t *)xmalloc(sizeof(integer_t) * M->nnz);
    M->rowptr = (integer_t *)xmalloc(sizeof(integer_t) * (M->n + 1));
    if (hdr[2] == 0) {             /* real file: imaginary parts zero */
        if (fread(raw, sizeof(double), M->nnz, f) != M->nnz) goto fail;
        for (uLong_t i = 0; i < M->nnz; ++i) { M->aij[i].real = raw[i]; M->aij[i].imag = 0.0; }
    } else if (fread(raw, sizeof(double), 2 * M->nnz, f) != 2 * M->nnz) goto fail;
    if (fread(M->colind, sizeof(integer_t), M->nnz, f) != M->nnz) goto fail;
    if (fread(M->rowptr, sizeof(integer_t), (size_t)M->n + 1, f) != (size_t)M->n + 1) goto fail;
    fclose(f);
    f = NULL;
    if (matrix_CheckCSR(M->n, M->nnz, M->colind, M->rowptr) != SPIKE_SUCCESS) goto fail;
    if (hdr[2] != 0) {
        for (uLong_t i = 0; i < M->nnz; ++i) {
            M->aij[i].real = planar ? raw[i] : raw[2 * i];
            M->aij[i].imag = planar ? raw[M->nnz + i] : raw[2 * i + 1];
        }
        M->planar_file = planar ? 1 : 0;
    }
    free(raw);
    matrix_ComputeBandwidth(M->n, M->colind, M->rowptr, M->aij, &M->ku, &M->kl);
    return M;
fail:
    if (f) fclose(f);
    free(raw);
    if (M) matrix_z_Deallocate(M);
    return NULL;
}

Error_t matrix_Deallocate(matrix_t *M)
{
    if (!M) return SPIKE_SUCCESS;
    free(M->colind); free(M->rowptr); free(M->aij); free(M);
    return SPIKE_SUCCESS;
}

Error_t matrix_z_Deallocate(matrix_z_t *M)
{
    if (!M) return SPIKE_SUCCESS;
    free(M->colind); free(M->rowptr); free(M->aij); free(M);
    return SPIKE_SUCCESS;
}

/* ------------------------------------------------------------------ construction */
matrix_t *matrix_CreateEmptyMatrix(integer_t n, uLong_t nnz)
{
    if (n < 0) return NULL;
    matrix_t *M = (matrix_t *)xcalloc(1, sizeof(matrix_t));
    M->n = n;
    M->nnz = nnz;
    M->colind = (integer_t *)xcalloc(nnz, sizeof(integer_t));
    M->rowptr = (integer_t *)xcalloc((size_t)n + 1, sizeof(integer_t));
    M->aij = (double *)xcalloc(nnz, sizeof(double));
    return M;
}

/* Wraps the caller's arrays without copying; free with free() on the struct only. */
matrix_t *matrix_CreateFromComponents(integer_t n, uLong_t nnz, integer_t *colind, integer_t *rowptr, double *aij)
{
    matrix_t *A = (matrix_t *)xcalloc(1, sizeof(matrix_t));
    A->n = n; A->nnz = nnz; A->colind = colind; A->rowptr = rowptr; A->aij = aij;
    matrix_ComputeBandwidth(n, colind, rowptr, aij, &A->ku, &A->kl);
    return A;
}

Bool_t matrix_AreEqual(const matrix_t *A, const matrix_t *B)
{
    if (!A || !B || A->n != B->n || A->nnz != B->nnz) return 0;
    if (memcmp(A->rowptr, B->rowptr, sizeof(integer_t) * ((size_t)A->n + 1))) return 0;
    if (memcmp(A->colind, B->colind, sizeof(integer_t) * A->nnz)) return 0;
    for (uLong_t i = 0; i < A->nnz; ++i)
        if (A->aij[i] != B->aij[i]) return 0;
    return 1;
}

/* ------------------------------------------------------------------ bandwidth / extraction */
Error_t matrix_ComputeBandwidth(integer_t n, const integer_t *colind, const integer_t *rowptr, const double *aij,
                                integer_t *ku, integer_t *kl)
{
    (void)aij;
    if (n < 0 || !rowptr || !ku || !kl || (n > 0 && rowptr[n] > 0 && !colind)) return SPIKE_ERR_INVALID_ARG;
    integer_t u = 0, l = 0;
    for (integer_t r = 0; r < n; ++r)
        for (integer_t idx = rowptr[r]; idx < rowptr[r + 1]; ++idx) {
            integer_t d = colind[idx] - r;
            if (d > u) u = d;
            if (-d > l) l = -d;
        }
    *ku = u;
    *kl = l;
    return SPIKE_SUCCESS;
}

matrix_t *matrix_ExtractMatrix(const matrix_t *M, integer_t r0, integer_t rf, integer_t c0, integer_t cf, matrixtype_t type)
{
    if (!M || r0 < 0 || rf > M->n || rf < r0) return NULL;
    uLong_t cnt = 0;
    for (integer_t r = r0; r < rf; ++r)
        for (integer_t idx = M->rowptr[r]; idx < M->rowptr[r + 1]; ++idx)
            cnt += (M->colind[idx] >= c0 && M->colind[idx] < cf);
    matrix_t *B = matrix_CreateEmptyMatrix(rf - r0, cnt);
    uLong_t q = 0;
    for (integer_t r = r0; r < rf; ++r) {
        for (integer_t idx = M->rowptr[r]; idx < M->rowptr[r + 1]; ++idx) {
            integer_t c = M->colind[idx];
            if (c >= c0 && c < cf) { B->colind[q] = c - c0; B->aij[q] = M->aij[idx]; ++q; }
        }
        B->rowptr[r - r0 + 1] = (integer_t)q;
    }
    B->type = type;
    matrix_ComputeBandwidth(B->n, B->colind, B->rowptr, B->aij, &B->ku, &B->kl);
    return B;
}

block_t *block_CreateEmptyBlock(integer_t n, integer_t m, integer_t ku, integer_t kl, blocktype_t type, blocksection_t section)
{
    if (n < 0 || m < 0) return NULL;
    block_t *B = (block_t *)xcalloc(1, sizeof(block_t));
    B->n = n; B->m = m; B->ku = ku; B->kl = kl; B->type = type; B->section = section;
    B->aij = (double *)xcalloc((size_t)n * m, sizeof(double));
    return B;
}

block_t *matrix_ExtractBlock(const matrix_t *M, integer_t r0, integer_t rf, integer_t c0, integer_t cf, blocktype_t type)
{
    if (!M || r0 < 0 || rf > M->n || rf < r0 || cf < c0) return NULL;
    block_t *B = block_CreateEmptyBlock(rf - r0, cf - c0, M->ku, M->kl, type, _WHOLE_SECTION_);
    for (integer_t r = r0; r < rf; ++r)
        for (integer_t idx = M->rowptr[r]; idx < M->rowptr[r + 1]; ++idx) {
            integer_t c = M->colind[idx];
            if (c >= c0 && c < cf) B->aij[(size_t)(r - r0) + (size_t)(c - c0) * B->n] = M->aij[idx];
        }
    return B;
}

Error_t block_Deallocate(block_t *B)
{
    if (B) { free(B->aij); free(B); }
    return SPIKE_SUCCESS;
}

/* ------------------------------------------------------------------ schedule */
sm_schedule_t *spike_solve_analysis_p(const matrix_t *A, integer_t nrhs, integer_t p)
{
    if (!A || A->n <= 0) return NULL;
    if (p <= 0) p = 4;                               /* spike_analysis.c:168 */
    if (p > A->n) p = A->n;
    sm_schedule_t *S = (sm_schedule_t *)xcalloc(1, sizeof(sm_schedule_t));
    const integer_t nreg = A->n / p;
    S->p = p;
    S->max_n = A->n - (p - 1) * nreg;
    S->max_m = A->ku > A->kl ? A->ku : A->kl;
    S->max_nrhs = S->max_m > nrhs ? S->max_m : nrhs;
    S->blockingDistance = 200;                       /* spike_analysis.c:51 */
    S->n = (integer_t *)xcalloc((size_t)p + 1, sizeof(integer_t));
    S->r = (integer_t *)xcalloc((size_t)p + 1, sizeof(integer_t));
    S->ku = (integer_t *)xcalloc((size_t)p, sizeof(integer_t));
    S->kl = (integer_t *)xcalloc((size_t)p, sizeof(integer_t));
    for (integer_t i = 0; i < p; ++i) {
        S->ku[i] = A->ku;
        S->kl[i] = A->kl;
        S->n[i] = i * nreg;
        S->r[i] = i * (A->ku + A->kl);
    }
    S->n[p] = A->n;
    S->r[p] = p * (A->ku + A->kl);
    return S;
}

sm_schedule_t *spike_solve_analysis(matrix_t *A, const integer_t nrhs)
{
    if (!A || A->n <= 0) return NULL;
    if (matrix_ComputeBandwidth(A->n, A->colind, A->rowptr, A->aij, &A->ku, &A->kl) != SPIKE_SUCCESS) return NULL;
    const char *v = getenv("HSPIKE_PARTITIONS");
    return spike_solve_analysis_p(A, nrhs, (v && atoi(v) > 0) ? atoi(v) : 4);
}

void schedule_Destroy(sm_schedule_t *S)
{
    if (!S) return;
    free(S->n); free(S->r); free(S->ku); free(S->kl); free(S);
}

/* ------------------------------------------------------------------ reduced system */
/* Pattern of R (spike_matrix.c:657-740): partition p owns rows [nr_p, nr_{p+1}) with
 * nr_p = sum_{q<p}(ku_q + kl_q); each of its rows holds, in column order, kl W entries at
 * columns [nr_p - kl, nr_p) when p > 0, a unit diagonal, and ku V entries at columns
 * [nr_{p+1}, nr_{p+1} + ku) when p < P-1.  The first ku rows take the top tips, the last kl
 * the bottom tips. */
matrix_t *matrix_CreateEmptyReducedSystem(integer_t P, const integer_t *n, const integer_t *ku, const integer_t *kl)
{
    (void)n;
    if (P <= 0 || !ku || !kl) return NULL;
    integer_t rows = 0;
    uLong_t nnz = 0;
    for (integer_t p = 0; p < P; ++p) {
        const integer_t k2 = ku[p] + kl[p];
        rows += k2;
        nnz += (uLong_t)k2 * (1 + (p > 0 ? kl[p] : 0) + (p < P - 1 ? ku[p] : 0));
    }
    matrix_t *R = matrix_CreateEmptyMatrix(rows, nnz);
    uLong_t q = 0;
    integer_t row = 0, nrp = 0;
    for (integer_t p = 0; p < P; ++p) {
        const integer_t k2 = ku[p] + kl[p], nrp1 = nrp + k2;
        for (integer_t t = 0; t < k2; ++t, ++row) {
            if (p > 0)
                for (integer_t c = nrp - kl[p]; c < nrp; ++c) R->colind[q++] = c;
            R->colind[q] = row;
            R->aij[q++] = 1.0;
            if (p < P - 1)
                for (integer_t c = nrp1; c < nrp1 + ku[p]; ++c) R->colind[q++] = c;
            R->rowptr[row + 1] = (integer_t)q;
        }
        nrp = nrp1;
    }
    R->type = _DIAG_BLOCK_;
    matrix_ComputeBandwidth(R->n, R->colind, R->rowptr, R->aij, &R->ku, &R->kl);
    return R;
}

/* Scatter row-major tips into R's values (spike_matrix.c:1414-1495 for one partition at a time);
 * the tip arrays here are column-major per partition: Vt[p][r + c*ku], Vb[p][r + c*kl], ... */
Error_t matrix_AssembleReducedSystem(matrix_t *R, integer_t P, integer_t ku, integer_t kl,
                                     const double *Vt, const double *Vb, const double *Wt, const double *Wb)
{
    if (!R || P <= 0) return SPIKE_ERR_INVALID_ARG;
    const integer_t k2 = ku + kl;
    uLong_t q = 0;
    for (integer_t p = 0; p < P; ++p)
        for (integer_t t = 0; t < k2; ++t) {
            const int top = t < ku;
            const integer_t tr = top ? t : t - ku;
            if (p > 0)
                for (integer_t c = 0; c < kl; ++c)
                    R->aij[q++] = top ? Wt[(size_t)p * ku * kl + tr + (size_t)c * ku] : Wb[(size_t)p * kl * kl + tr + (size_t)c * kl];
            R->aij[q++] = 1.0;
            if (p < P - 1)
                for (integer_t c = 0; c < ku; ++c)
                    R->aij[q++] = top ? Vt[(size_t)p * ku * ku + tr + (size_t)c * ku] : Vb[(size_t)p * kl * ku + tr + (size_t)c * kl];
        }
    return q == R->nnz ? SPIKE_SUCCESS : SPIKE_ERR_INVALID_ARG;
}

/* g rows: top tips at nr_p, bottom tips at nr_p + ku (spike_matrix.c:1272-1297) */
Error_t block_AssembleReducedRHS(double *g, integer_t P, integer_t ku, integer_t kl, integer_t nrhs,
                                 const double *yt, const double *yb)
{
    if (!g || P <= 0) return SPIKE_ERR_INVALID_ARG;
    const integer_t k2 = ku + kl, Nr = P * k2;
    for (integer_t c = 0; c < nrhs; ++c)
        for (integer_t p = 0; p < P; ++p) {
            for (integer_t r = 0; r < ku; ++r) g[(size_t)c * Nr + p * k2 + r] = yt[(size_t)p * ku * nrhs + r + (size_t)c * ku];
            for (integer_t r = 0; r < kl; ++r) g[(size_t)c * Nr + p * k2 + ku + r] = yb[(size_t)p * kl * nrhs + r + (size_t)c * kl];
        }
    return SPIKE_SUCCESS;
}

matrix_z_t *matrix_LoadCSR_z(const char *filename) { return matrix_LoadCSR_zl(filename, 1); }
