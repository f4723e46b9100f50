/*
 * spike_matrix.h -- data structures of the SPIKE banded-solve path (MI355X build).
 *
 * Layout-compatible with the reference's containers:
 *   matrix_t      <- OpenMP/spike_matrix.h:39-53   (CSR, int32 0-based indices)
 *   block_t       <- OpenMP/spike_matrix.h:62-72   (dense column-major panel)
 *   sm_schedule_t <- OpenMP/spike_analysis.h:38-53 (partition schedule)
 *   scalar types  <- OpenMP/spike_datatypes.h:86-91, spike_interfaces.h:24-25
 *
 * The reference selects one precision per build with _DATATYPE_{S,D,C,Z}_; this library
 * serves every precision from one build, so the containers hold FP64 values (the path's
 * oracle precision) and the complex entry points take interleaved complex16 arrays.
 *
 * Error codes: the reference returns 0/1 and aborts on every failure; this library never
 * aborts and returns SPIKE_SUCCESS (0) or one of the negative SPIKE_ERR_* codes below.
 */
#ifndef HSPIKE_SPIKE_MATRIX_H
#define HSPIKE_SPIKE_MATRIX_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef HSPIKE_SCALAR_TYPES
#define HSPIKE_SCALAR_TYPES
typedef int      integer_t;   /* spike_datatypes.h:86 */
typedef int      Error_t;     /* spike_datatypes.h:87 */
#endif
typedef int      Bool_t;      /* spike_datatypes.h:88 */
typedef uint64_t uLong_t;     /* spike_datatypes.h:91 */

#ifndef HSPIKE_COMPLEX_TYPES
#define HSPIKE_COMPLEX_TYPES
typedef struct { float  real; float  imag; } complex8;   /* spike_interfaces.h:24 */
typedef struct { double real; double imag; } complex16;  /* spike_interfaces.h:25 */
#endif

#define SPIKE_SUCCESS              0
#define SPIKE_ERR_INVALID_ARG     (-1)  /* n/nnz/nrhs out of range, NULL pointer          */
#define SPIKE_ERR_MALFORMED_CSR   (-2)  /* rowptr not monotone, colind out of range        */
#define SPIKE_ERR_OOM             (-3)  /* device or host allocation failed / over budget  */
#define SPIKE_ERR_SINGULAR        (-4)  /* zero or non-finite pivot in a diagonal block    */
#define SPIKE_ERR_NOT_CONVERGED   (-5)  /* refinement / reduced solve missed its tolerance */
#define SPIKE_ERR_UNSUPPORTED     (-6)  /* bandwidth or shape beyond this build's kernels  */
#define SPIKE_ERR_DEVICE          (-7)  /* HIP runtime error or no GPU / no HIP library    */
#define SPIKE_ERR_IO              (-8)  /* file open / read / write failure                */
#define SPIKE_ERR_CANCELLED       (-9)  /* serving request cancelled or expired            */

typedef enum { _C_BLOCK_, _DIAG_BLOCK_, _B_BLOCK_ } matrixtype_t;             /* spike_matrix.h:35 */
typedef enum { _V_BLOCK_, _W_BLOCK_, _RHS_BLOCK_ } blocktype_t;               /* spike_matrix.h:59 */
typedef enum { _TOP_SECTION_, _CENTRAL_SECTION_, _BOTTOM_SECTION_, _WHOLE_SECTION_ } blocksection_t;

/* CSR matrix: spike_matrix.h:39-53 (aij is FP64 here; complex CSR uses matrix_z_t) */
typedef struct {
    integer_t    n;
    uLong_t      nnz;
    integer_t    ku;        /* upper bandwidth: max(col - row) */
    integer_t    kl;        /* lower bandwidth: max(row - col) */
    matrixtype_t type;
    integer_t   *colind;
    integer_t   *rowptr;
    double      *aij;
} matrix_t;

/* complex128 CSR with interleaved values (the layout spike_matrix.c:34-35 reads) */
typedef struct {
    integer_t    n;
    uLong_t      nnz;
    integer_t    ku;
    integer_t    kl;
    matrixtype_t type;
    integer_t   *colind;
    integer_t   *rowptr;
    complex16   *aij;
    integer_t    planar_file;   /* 1 when the file was read as planar re/im arrays (generator_banded.py:134-141) */
} matrix_z_t;

/* dense column-major panel: spike_matrix.h:62-72, aij[row + col*n] */
typedef struct {
    blocktype_t    type;
    blocksection_t section;
    integer_t      n;
    integer_t      m;
    integer_t      ku;
    integer_t      kl;
    double        *aij;
} block_t;

/* partition schedule: spike_analysis.h:38-53 */
typedef struct {
    integer_t  p;                /* number of partitions                              */
    integer_t  max_n;            /* largest partition                                 */
    integer_t  max_m;            /* max(ku, kl)                                       */
    integer_t  max_nrhs;         /* max(ku, kl, nrhs)                                 */
    integer_t  blockingDistance; /* RHS column block used by the _blocking driver     */
    integer_t *n;                /* [p+1] first row of each partition, n[p] = N       */
    integer_t *r;                /* [p+1] first row of each partition in R            */
    integer_t *ku;               /* [p]                                               */
    integer_t *kl;               /* [p]                                               */
} sm_schedule_t;

/* ---- CSR file IO (format: int32 n, int32 nnz, int32 dtype, aij[nnz], colind[nnz], rowptr[n+1]) */
matrix_t   *matrix_LoadCSR(const char *filename);                          /* spike_matrix.c:11-51 */
Error_t     matrix_SaveCSR(const char *filename, const matrix_t *M);       /* spike_matrix.c:56-85 */
matrix_z_t *matrix_LoadCSR_z(const char *filename);                        /* complex, planar values (generator_banded.py:134-141) */
matrix_z_t *matrix_LoadCSR_zl(const char *filename, int planar);           /* complex, planar (1) or interleaved (0) values */
Error_t     matrix_Deallocate(matrix_t *M);                                /* spike_matrix.c:265-273 */
Error_t     matrix_z_Deallocate(matrix_z_t *M);

/* ---- construction / comparison */
matrix_t   *matrix_CreateEmptyMatrix(integer_t n, uLong_t nnz);            /* spike_matrix.c:160-190 */
matrix_t   *matrix_CreateFromComponents(integer_t n, uLong_t nnz, integer_t *colind,
                                        integer_t *rowptr, double *aij);   /* spike_matrix.c:243-262 (wraps, no copy) */
Bool_t      matrix_AreEqual(const matrix_t *A, const matrix_t *B);         /* spike_matrix.c:277-323 (exact) */
Error_t     matrix_CheckCSR(integer_t n, uLong_t nnz, const integer_t *colind, const integer_t *rowptr);

/* ---- bandwidth and extraction */
/* spike_algebra.h:114-119 / spike_algebra.c:56-95: aij is accepted (reference signature) and not
 * read; ku = max(col - row), kl = max(row - col) over the stored entries */
Error_t     matrix_ComputeBandwidth(integer_t n, const integer_t *colind, const integer_t *rowptr,
                                    const double *aij, integer_t *ku, integer_t *kl);
matrix_t   *matrix_ExtractMatrix(const matrix_t *M, integer_t r0, integer_t rf,
                                 integer_t c0, integer_t cf, matrixtype_t type); /* spike_matrix.c:563-645 */
block_t    *matrix_ExtractBlock(const matrix_t *M, integer_t r0, integer_t rf,
                                integer_t c0, integer_t cf, blocktype_t type);   /* spike_matrix.c:421-458 */
block_t    *block_CreateEmptyBlock(integer_t n, integer_t m, integer_t ku, integer_t kl,
                                   blocktype_t type, blocksection_t section);    /* spike_matrix.c:745-775 */
Error_t     block_Deallocate(block_t *B);

/* ---- schedule: same partition boundaries as spike_analysis.c:38-76.
 * spike_solve_analysis (spike_analysis.h:65, spike_analysis.c:3-83) computes A's bandwidth into
 * A->ku / A->kl as the reference does and uses p = 4 (spike_analysis.c:168), or the value of
 * the HSPIKE_PARTITIONS environment variable when it is set; spike_solve_analysis_p takes a
 * caller-chosen p (p <= 0 selects 4) and reads A->ku / A->kl as they are. */
sm_schedule_t *spike_solve_analysis(matrix_t *A, const integer_t nrhs);
sm_schedule_t *spike_solve_analysis_p(const matrix_t *A, integer_t nrhs, integer_t p);
void           schedule_Destroy(sm_schedule_t *S);      /* spike_analysis.c:85-95 */

/* ---- reduced system in the reference's CSR pattern (spike_matrix.c:657-740, 1013-1120,
 *      1254-1332, 1379-1611).  Tips are column-major:
 *        Vt[ku x ku], Vb[kl x ku], Wt[ku x kl], Wb[kl x kl], yt[ku x nrhs], yb[kl x nrhs]
 *      per partition, each array a concatenation over partitions (V absent for p = P-1, W for p = 0
 *      but slots are still reserved).  Returns R (N_r x N_r CSR) and writes g (N_r x nrhs). */
/* spike_matrix.h:242-245: n (the partition row bounds, P+1 entries) is accepted and not read,
 * as in the reference; the pattern depends on ku[p], kl[p] only */
matrix_t   *matrix_CreateEmptyReducedSystem(integer_t P, const integer_t *n, const integer_t *ku, const integer_t *kl);
Error_t     matrix_AssembleReducedSystem(matrix_t *R, integer_t P, integer_t ku, integer_t kl,
                                         const double *Vt, const double *Vb,
                                         const double *Wt, const double *Wb);
Error_t     block_AssembleReducedRHS(double *g, integer_t P, integer_t ku, integer_t kl, integer_t nrhs,
                                     const double *yt, const double *yb);

#ifdef __cplusplus
}
#endif
#endif
