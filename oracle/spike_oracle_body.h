/*
 * Type-generic body of the CPU oracle; included by spike_oracle.c once per value type with
 *   OR_T   value type (double or double _Complex)
 *   OR_S   name suffix (d or z)
 *   OR_ABS modulus function
 * Test infrastructure only (see spike_oracle.c).
 */
#define OR_CAT2(a, b) a##_##b
#define OR_CAT(a, b) OR_CAT2(a, b)
#define OR_FN(name) OR_CAT(name, OR_S)

/* Band storage of LAPACK's gbtrf: A(i,j) at ab[(kv + i - j) + j*ldab], kv = ku + kl,
 * ldab = 2*kl + ku + 1; rows [0, kl) of each column are fill-in space for pivoting. */
typedef struct {
    int    n, kl, ku, kv, ldab;
    OR_T  *ab;
    int   *ipiv;
    int    info;          /* 0, or 1 + first column with an exactly zero pivot */
} OR_FN(or_band);

#define OR_AB(F, i, j) ((F)->ab[(size_t)((F)->kv + (i) - (j)) + (size_t)(j) * (F)->ldab])

/* Load rows/cols [r0, rf) of the CSR matrix into band storage (zero elsewhere). */
static int OR_FN(or_band_load)(OR_FN(or_band) *F, const int *rowptr, const int *colind, const OR_T *aij,
                               int r0, int rf, int kl, int ku)
{
    F->n = rf - r0; F->kl = kl; F->ku = ku; F->kv = kl + ku; F->ldab = 2 * kl + ku + 1; F->info = 0;
    F->ab = (OR_T *)calloc((size_t)F->ldab * (size_t)(F->n > 0 ? F->n : 1), sizeof(OR_T));
    F->ipiv = (int *)calloc((size_t)(F->n > 0 ? F->n : 1), sizeof(int));
    if (!F->ab || !F->ipiv) return -1;
    for (int r = r0; r < rf; ++r)
        for (int idx = rowptr[r]; idx < rowptr[r + 1]; ++idx) {
            int c = colind[idx];
            if (c < r0 || c >= rf) continue;
            int i = r - r0, j = c - r0;
            if (i - j > kl || j - i > ku) return -2;   /* outside the declared band */
            OR_AB(F, i, j) += aij[idx];
        }
    return 0;
}

static void OR_FN(or_band_free)(OR_FN(or_band) *F)
{
    free(F->ab); free(F->ipiv); F->ab = NULL; F->ipiv = NULL;
}

/* LU with partial (row) pivoting of a band matrix -- the algorithm of LAPACK's unblocked
 * gbtf2: at column j pick the largest |A(i,j)| for i in [j, j+kl], swap it into row j over
 * the columns the swap can reach, scale the multipliers and apply the rank-1 update inside
 * the band (U's bandwidth grows to ku + kl).  This stands in for SuperLU_MT's pdgstrf with
 * diag_pivot_thresh u = 1.0 (spike_superlu.c:76), which is partial pivoting as well. */
static int OR_FN(or_band_factor)(OR_FN(or_band) *F)
{
    const int n = F->n, kl = F->kl, kv = F->kv;
    int ju = 0;
    for (int j = 0; j < n; ++j) {
        int km = kl < n - 1 - j ? kl : n - 1 - j;
        int jp = 0; double best = -1.0;
        for (int t = 0; t <= km; ++t) {
            double a = OR_ABS(OR_AB(F, j + t, j));
            if (a > best) { best = a; jp = t; }
        }
        F->ipiv[j] = j + jp;
        if (best == 0.0) { if (!F->info) F->info = j + 1; continue; }
        int jlast = j + F->ku + jp; if (jlast > n - 1) jlast = n - 1;
        if (jlast > ju) ju = jlast;
        if (jp) {
            for (int c = j; c <= ju; ++c) {
                OR_T t = OR_AB(F, j, c); OR_AB(F, j, c) = OR_AB(F, j + jp, c); OR_AB(F, j + jp, c) = t;
            }
        }
        OR_T piv = OR_AB(F, j, j);
        for (int r = j + 1; r <= j + km; ++r) OR_AB(F, r, j) /= piv;
        for (int c = j + 1; c <= ju; ++c) {
            OR_T u = OR_AB(F, j, c);
            if (u == 0) continue;
            for (int r = j + 1; r <= j + km; ++r) OR_AB(F, r, c) -= OR_AB(F, r, j) * u;
        }
        (void)kv;
    }
    return F->info;
}

/* Solve A X = B in place for m columns of B (column-major, leading dimension ldb):
 * forward with the unit-lower factor and the recorded interchanges, then backward with U
 * (LAPACK gbtrs, no transpose). */
static void OR_FN(or_band_solve)(const OR_FN(or_band) *F, int m, OR_T *B, int ldb)
{
    const int n = F->n, kl = F->kl, kv = F->kv;
    for (int c = 0; c < m; ++c) {
        OR_T *b = B + (size_t)c * ldb;
        for (int j = 0; j < n - 1; ++j) {
            int lm = kl < n - 1 - j ? kl : n - 1 - j;
            int l = F->ipiv[j];
            if (l != j) { OR_T t = b[l]; b[l] = b[j]; b[j] = t; }
            OR_T bj = b[j];
            if (bj != 0) for (int r = j + 1; r <= j + lm; ++r) b[r] -= OR_AB(F, r, j) * bj;
        }
        for (int j = n - 1; j >= 0; --j) {
            b[j] /= OR_AB(F, j, j);
            OR_T bj = b[j];
            int i0 = j - kv > 0 ? j - kv : 0;
            if (bj != 0) for (int i = i0; i < j; ++i) b[i] -= OR_AB(F, i, j) * bj;
        }
    }
}

/* Dense column-major block M[r0:rf, c0:cf] of a CSR matrix (spike_matrix.c:421-458). */
static void OR_FN(or_extract_dense)(const int *rowptr, const int *colind, const OR_T *aij,
                                    int r0, int rf, int c0, int cf, OR_T *out)
{
    const int nr = rf - r0, nc = cf - c0;
    memset(out, 0, sizeof(OR_T) * (size_t)nr * (size_t)(nc > 0 ? nc : 0));
    for (int r = r0; r < rf; ++r)
        for (int idx = rowptr[r]; idx < rowptr[r + 1]; ++idx) {
            int c = colind[idx];
            if (c >= c0 && c < cf) out[(size_t)(r - r0) + (size_t)(c - c0) * nr] = aij[idx];
        }
}

/* Explicit SPIKE solve of A X = B (spike_interfaces.c:43-419) with every intermediate:
 *  stage 1  per partition: banded LU of A_i = A[r0:rf, r0:rf]                (:113-128)
 *  stage 2  y_i = A_i^{-1} f_i, V_i = A_i^{-1} B_i, W_i = A_i^{-1} C_i        (:135-231)
 *           with B_i = A[r0:rf, rf:rf+ku], C_i = A[r0:rf, r0-kl:r0]
 *           tips: top = first ku rows, bottom = last kl rows (spike_matrix.c:1013-1050)
 *  stage 3  reduced system R (reference CSR pattern) and RHS g, solved by banded LU (:243-248)
 *  stage 4  f_i[ni-ku:] -= A[rf-ku:rf, rf:rf+ku] xr[r(p+1) : +ku];
 *           f_i[:kl]    -= A[r0:r0+kl, r0-kl:r0] xr[r(p)-kl : r(p)];  x_i = A_i^{-1} f_i (:259-396)
 *           every column of x_i is written (the reference copies column 0, spike_matrix.c:1339-1345)
 * Partition bounds: rows[p] = p*floor(n/P), rows[P] = n (spike_analysis.c:38-76).
 * Tip arrays (may be NULL) are per-partition concatenations, column-major:
 *  Vt[P][ku*ku], Vb[P][kl*ku], Wt[P][ku*kl], Wb[P][kl*kl], yt[P][ku*nrhs], yb[P][kl*nrhs].
 * Reduced outputs (may be NULL): R in CSR (Rrowptr[Nr+1], Rcolind/Raij[nnzR]), g and xr (Nr x nrhs).
 * Returns 0, -2 bad band, -3 singular partition, -4 singular reduced system, -1 alloc. */
int OR_FN(or_spike_solve)(int n, const int *rowptr, const int *colind, const OR_T *aij,
                          int P, int nrhs, const OR_T *B, OR_T *X,
                          OR_T *Vt, OR_T *Vb, OR_T *Wt, OR_T *Wb, OR_T *yt, OR_T *yb,
                          int *Rrowptr, int *Rcolind, OR_T *Raij, OR_T *g, OR_T *xr, int nthreads)
{
    int ku, kl;
    or_bandwidth(n, rowptr, colind, &ku, &kl);
    if (P < 1) P = 1;
    int *rows = (int *)malloc(sizeof(int) * (size_t)(P + 1));
    int *rr = (int *)malloc(sizeof(int) * (size_t)(P + 1));
    or_schedule(n, P, ku, kl, rows, rr);
    const int k2 = ku + kl, Nr = P * k2;
    OR_FN(or_band) *F = (OR_FN(or_band) *)calloc((size_t)P, sizeof(OR_FN(or_band)));
    OR_T *T_Vt = (OR_T *)calloc((size_t)P * ku * ku + 1, sizeof(OR_T));
    OR_T *T_Vb = (OR_T *)calloc((size_t)P * kl * ku + 1, sizeof(OR_T));
    OR_T *T_Wt = (OR_T *)calloc((size_t)P * ku * kl + 1, sizeof(OR_T));
    OR_T *T_Wb = (OR_T *)calloc((size_t)P * kl * kl + 1, sizeof(OR_T));
    OR_T *T_yt = (OR_T *)calloc((size_t)P * ku * nrhs + 1, sizeof(OR_T));
    OR_T *T_yb = (OR_T *)calloc((size_t)P * kl * nrhs + 1, sizeof(OR_T));
    int status = 0;
    if (!rows || !rr || !F || !T_Vt || !T_Vb || !T_Wt || !T_Wb || !T_yt || !T_yb) { status = -1; goto done; }

#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    /* stages 1 and 2, independent per partition */
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < P; ++p) {
        const int r0 = rows[p], rf = rows[p + 1], ni = rf - r0;
        OR_FN(or_band) *Fp = &F[p];
        int e = OR_FN(or_band_load)(Fp, rowptr, colind, aij, r0, rf, kl, ku);
        if (e) { status = (e == -1) ? -1 : -2; continue; }
        if (OR_FN(or_band_factor)(Fp)) { status = -3; continue; }
        /* y_i */
        OR_T *fy = (OR_T *)malloc(sizeof(OR_T) * (size_t)ni * (nrhs > 0 ? nrhs : 1));
        for (int c = 0; c < nrhs; ++c) memcpy(fy + (size_t)c * ni, B + (size_t)c * n + r0, sizeof(OR_T) * ni);
        OR_FN(or_band_solve)(Fp, nrhs, fy, ni);
        for (int c = 0; c < nrhs; ++c) {
            for (int r = 0; r < ku; ++r) T_yt[(size_t)p * ku * nrhs + r + (size_t)c * ku] = fy[(size_t)c * ni + r];
            for (int r = 0; r < kl; ++r) T_yb[(size_t)p * kl * nrhs + r + (size_t)c * kl] = fy[(size_t)c * ni + ni - kl + r];
        }
        free(fy);
        if (p < P - 1 && ku > 0) {            /* V_i = A_i^{-1} A[r0:rf, rf:rf+ku] */
            OR_T *V = (OR_T *)malloc(sizeof(OR_T) * (size_t)ni * ku);
            OR_FN(or_extract_dense)(rowptr, colind, aij, r0, rf, rf, rf + ku, V);
            OR_FN(or_band_solve)(Fp, ku, V, ni);
            for (int c = 0; c < ku; ++c) {
                for (int r = 0; r < ku; ++r) T_Vt[(size_t)p * ku * ku + r + (size_t)c * ku] = V[(size_t)c * ni + r];
                for (int r = 0; r < kl; ++r) T_Vb[(size_t)p * kl * ku + r + (size_t)c * kl] = V[(size_t)c * ni + ni - kl + r];
            }
            free(V);
        }
        if (p > 0 && kl > 0) {                /* W_i = A_i^{-1} A[r0:rf, r0-kl:r0] */
            OR_T *W = (OR_T *)malloc(sizeof(OR_T) * (size_t)ni * kl);
            OR_FN(or_extract_dense)(rowptr, colind, aij, r0, rf, r0 - kl, r0, W);
            OR_FN(or_band_solve)(Fp, kl, W, ni);
            for (int c = 0; c < kl; ++c) {
                for (int r = 0; r < ku; ++r) T_Wt[(size_t)p * ku * kl + r + (size_t)c * ku] = W[(size_t)c * ni + r];
                for (int r = 0; r < kl; ++r) T_Wb[(size_t)p * kl * kl + r + (size_t)c * kl] = W[(size_t)c * ni + ni - kl + r];
            }
            free(W);
        }
    }
    if (status) goto done;

    /* stage 3: reduced system in the reference pattern (spike_matrix.c:657-740, 1414-1495) */
    {
        int nnzR = or_reduced_nnz(P, ku, kl);
        int *Rp = (int *)malloc(sizeof(int) * (size_t)(Nr + 1));
        int *Rc = (int *)malloc(sizeof(int) * (size_t)(nnzR > 0 ? nnzR : 1));
        OR_T *Ra = (OR_T *)malloc(sizeof(OR_T) * (size_t)(nnzR > 0 ? nnzR : 1));
        OR_T *G = (OR_T *)calloc((size_t)Nr * (nrhs > 0 ? nrhs : 1), sizeof(OR_T));
        int q = 0; Rp[0] = 0;
        for (int p = 0; p < P; ++p) {
            const int nrp = p * k2, nrp1 = (p + 1) * k2;
            for (int t = 0; t < k2; ++t) {
                const int row = nrp + t, top = t < ku, tr = top ? t : t - ku;
                if (p > 0)
                    for (int c = 0; c < kl; ++c) {
                        Rc[q] = nrp - kl + c;
                        Ra[q++] = top ? T_Wt[(size_t)p * ku * kl + tr + (size_t)c * ku]
                                      : T_Wb[(size_t)p * kl * kl + tr + (size_t)c * kl];
                    }
                Rc[q] = row; Ra[q++] = 1.0;
                if (p < P - 1)
                    for (int c = 0; c < ku; ++c) {
                        Rc[q] = nrp1 + c;
                        Ra[q++] = top ? T_Vt[(size_t)p * ku * ku + tr + (size_t)c * ku]
                                      : T_Vb[(size_t)p * kl * ku + tr + (size_t)c * kl];
                    }
                Rp[row + 1] = q;
            }
            for (int c = 0; c < nrhs; ++c) {
                for (int r = 0; r < ku; ++r) G[(size_t)c * Nr + nrp + r] = T_yt[(size_t)p * ku * nrhs + r + (size_t)c * ku];
                for (int r = 0; r < kl; ++r) G[(size_t)c * Nr + nrp + ku + r] = T_yb[(size_t)p * kl * nrhs + r + (size_t)c * kl];
            }
        }
        if (Rrowptr) memcpy(Rrowptr, Rp, sizeof(int) * (size_t)(Nr + 1));
        if (Rcolind) memcpy(Rcolind, Rc, sizeof(int) * (size_t)q);
        if (Raij) memcpy(Raij, Ra, sizeof(OR_T) * (size_t)q);
        if (g) memcpy(g, G, sizeof(OR_T) * (size_t)Nr * nrhs);
        /* R is banded: solve it with the same partial-pivot band LU (PARDISO in the reference) */
        int rku, rkl;
        or_bandwidth(Nr, Rp, Rc, &rku, &rkl);
        OR_FN(or_band) FR;
        memset(&FR, 0, sizeof(FR));
        int e = OR_FN(or_band_load)(&FR, Rp, Rc, Ra, 0, Nr, rkl, rku);
        if (!e && OR_FN(or_band_factor)(&FR)) e = -4;
        if (!e) OR_FN(or_band_solve)(&FR, nrhs, G, Nr);     /* G now holds xr */
        OR_FN(or_band_free)(&FR);
        if (e) { status = (e == -4) ? -4 : -1; free(Rp); free(Rc); free(Ra); free(G); goto done; }
        if (xr) memcpy(xr, G, sizeof(OR_T) * (size_t)Nr * nrhs);

        /* stage 4: recovery (spike_interfaces.c:259-396) */
#pragma omp parallel for schedule(dynamic, 1)
        for (int p = 0; p < P; ++p) {
            const int r0 = rows[p], rf = rows[p + 1], ni = rf - r0;
            OR_T *f = (OR_T *)malloc(sizeof(OR_T) * (size_t)ni * (nrhs > 0 ? nrhs : 1));
            for (int c = 0; c < nrhs; ++c) memcpy(f + (size_t)c * ni, B + (size_t)c * n + r0, sizeof(OR_T) * ni);
            if (p < P - 1 && ku > 0) {
                OR_T *Bb = (OR_T *)malloc(sizeof(OR_T) * (size_t)ku * ku);
                OR_FN(or_extract_dense)(rowptr, colind, aij, rf - ku, rf, rf, rf + ku, Bb);
                for (int c = 0; c < nrhs; ++c)
                    for (int r = 0; r < ku; ++r) {
                        OR_T s = 0;
                        for (int t = 0; t < ku; ++t) s += Bb[r + (size_t)t * ku] * G[(size_t)c * Nr + (p + 1) * k2 + t];
                        f[(size_t)c * ni + ni - ku + r] -= s;
                    }
                free(Bb);
            }
            if (p > 0 && kl > 0) {
                OR_T *Ct = (OR_T *)malloc(sizeof(OR_T) * (size_t)kl * kl);
                OR_FN(or_extract_dense)(rowptr, colind, aij, r0, r0 + kl, r0 - kl, r0, Ct);
                for (int c = 0; c < nrhs; ++c)
                    for (int r = 0; r < kl; ++r) {
                        OR_T s = 0;
                        for (int t = 0; t < kl; ++t) s += Ct[r + (size_t)t * kl] * G[(size_t)c * Nr + p * k2 - kl + t];
                        f[(size_t)c * ni + r] -= s;
                    }
                free(Ct);
            }
            OR_FN(or_band_solve)(&F[p], nrhs, f, ni);
            for (int c = 0; c < nrhs; ++c) memcpy(X + (size_t)c * n + r0, f + (size_t)c * ni, sizeof(OR_T) * ni);
            free(f);
        }
        free(Rp); free(Rc); free(Ra); free(G);
    }
    if (Vt) memcpy(Vt, T_Vt, sizeof(OR_T) * (size_t)P * ku * ku);
    if (Vb) memcpy(Vb, T_Vb, sizeof(OR_T) * (size_t)P * kl * ku);
    if (Wt) memcpy(Wt, T_Wt, sizeof(OR_T) * (size_t)P * ku * kl);
    if (Wb) memcpy(Wb, T_Wb, sizeof(OR_T) * (size_t)P * kl * kl);
    if (yt) memcpy(yt, T_yt, sizeof(OR_T) * (size_t)P * ku * nrhs);
    if (yb) memcpy(yb, T_yb, sizeof(OR_T) * (size_t)P * kl * nrhs);
done:
    if (F) for (int p = 0; p < P; ++p) OR_FN(or_band_free)(&F[p]);
    free(F); free(rows); free(rr);
    free(T_Vt); free(T_Vb); free(T_Wt); free(T_Wb); free(T_yt); free(T_yb);
    return status;
}

/* Direct banded solve of the whole system (no partitioning): the reference's
 * directSolver_Solve path (spike_superlu.c:325-526) with the band LU above. */
int OR_FN(or_band_direct_solve)(int n, const int *rowptr, const int *colind, const OR_T *aij,
                                int nrhs, const OR_T *B, OR_T *X)
{
    int ku, kl;
    or_bandwidth(n, rowptr, colind, &ku, &kl);
    OR_FN(or_band) F;
    memset(&F, 0, sizeof(F));
    int e = OR_FN(or_band_load)(&F, rowptr, colind, aij, 0, n, kl, ku);
    if (e) { OR_FN(or_band_free)(&F); return e; }
    if (OR_FN(or_band_factor)(&F)) { OR_FN(or_band_free)(&F); return -3; }
    memcpy(X, B, sizeof(OR_T) * (size_t)n * nrhs);
    OR_FN(or_band_solve)(&F, nrhs, X, n);
    OR_FN(or_band_free)(&F);
    return 0;
}

/* FP64 relative residual ||A x - b||_2 / ||b||_2 of every column (spike_algebra.c:97-151,
 * all columns instead of the first five). */
void OR_FN(or_residual)(int n, const int *rowptr, const int *colind, const OR_T *aij,
                        int nrhs, const OR_T *X, const OR_T *B, double *rel)
{
    for (int c = 0; c < nrhs; ++c) {
        double rn = 0.0, bn = 0.0;
        const OR_T *x = X + (size_t)c * n, *b = B + (size_t)c * n;
        for (int r = 0; r < n; ++r) {
            OR_T s = 0;
            for (int idx = rowptr[r]; idx < rowptr[r + 1]; ++idx) s += aij[idx] * x[colind[idx]];
            double d = OR_ABS(s - b[r]), bb = OR_ABS(b[r]);
            rn += d * d; bn += bb * bb;
        }
        rel[c] = bn > 0 ? sqrt(rn) / sqrt(bn) : sqrt(rn);
    }
}

#undef OR_AB
#undef OR_FN
#undef OR_CAT
#undef OR_CAT2
