/*
 * spike_ext.h -- extended entry points of the MI355X SPIKE library.
 *
 * The reference allows "an extended entry (handle + options) without changing" the legacy
 * ABI; it has a factor-once / solve-many handle only at the backend level
 * (OpenMP/spike_pardiso.h:87-115, OpenMP/superlu_test.c:129-147). This header exposes:
 *
 *   - a solver handle over a CSR matrix that already lives in device memory (HBM), with
 *     precision / partition / tolerance options and per-stage statistics;
 *   - access to the intermediate SPIKE quantities (spike tips, reduced solution) so callers
 *     can check them against the reference's reduced system (spike_matrix.c:657-740);
 *   - a batched solve of many independent systems of uneven size in one set of launches;
 *   - a continuous-batching request server with memory admission and resident factors;
 *   - the deterministic synthetic generators used by the benchmarks and tests.
 *
 * All functions return SPIKE_SUCCESS (0) or a negative SPIKE_ERR_* code (spike_matrix.h).
 * Device pointers are HIP device addresses; no torch or HIP types appear in the signatures
 * (streams are passed as void*).
 */
#ifndef HSPIKE_SPIKE_EXT_H
#define HSPIKE_SPIKE_EXT_H

#include <stdint.h>
#include <stddef.h>
#include "spike_matrix.h"

#ifdef __cplusplus
extern "C" {
#endif

/* value type of the caller's CSR values, RHS and solution arrays */
#define SPIKE_DTYPE_D 0   /* double                                   */
#define SPIKE_DTYPE_Z 1   /* complex16 (interleaved re, im doubles)   */

/* arithmetic of the factorization and spike sweeps */
#define SPIKE_PREC_FP64  0  /* plain FP64 SPIKE (FP64 refinement only if the residual misses tol) */
#define SPIKE_PREC_MIXED 1  /* FP32 factor + spikes, FP64 residual / reduced solve / correction   */

typedef struct {
    int    precision;      /* SPIKE_PREC_FP64 or SPIKE_PREC_MIXED                         */
    int    partitions;     /* 0 = autotune from (n, k, nrhs, dtype); otherwise fixed P     */
    int    max_refine;     /* cap on refinement iterations (default 8)                     */
    double tol;            /* target per-column ||b - A x|| / ||b|| (default 1e-10)        */
    int    reduced_maxit;  /* cap on block-Jacobi sweeps of the reduced system (default 64) */
    double reduced_tol;    /* relative update that ends the reduced iteration (1e-15)      */
    int    sweep_waves;    /* waves (16 columns each) per sweep workgroup: 0 auto, 1, 2, 4 */
    int    verbose;        /* 1 = stage timings on stderr                                   */
    int    reduced_solver; /* SPIKE_RED_AUTO (default), SPIKE_RED_JACOBI or SPIKE_RED_DIRECT  */
    int    graph;          /* 1: replay the solve phase from a captured HIP graph (see below)  */
    int    tune;           /* partition autotuning: SPIKE_TUNE_LATENCY (default) or _THROUGHPUT */
} spike_opts_t;

/* Partition autotuning (opts.partitions = 0).  LATENCY sizes partitions so that one system fills
 * the device (about 12 blocks each, capped at the factor workgroups the device runs at once);
 * THROUGHPUT uses partitions of about 96 blocks, for batches of many systems, where the device
 * is full anyway and every partition adds spike and reduced-system work.  Either way P depends
 * only on the system and the options, so a system gets the same bits alone and in a batch solved
 * with the same options. */
#define SPIKE_TUNE_LATENCY    0
#define SPIKE_TUNE_THROUGHPUT 1

/* Graph replay (opts.graph = 1).  The decisions of a solve -- which systems need another
 * refinement pass, which reduced iterations have converged -- are taken on the device, so a
 * solve is a fixed sequence of launches.  With graph = 1, a solve of at most 64 RHS columns is
 * captured once (after an eager solve of the same shapes has set the budget of passes and sweeps)
 * and replayed while the handle's or batch's arrays, shapes and options stay the same; the
 * results are the eager solve's bits.  A replay that would need more passes or sweeps than its
 * budget, a system on the direct reduced solve, or any change of the arrays falls back to the
 * eager solve (which sets a new budget).  The handle then works on a non-blocking stream of its
 * own unless spike_set_stream gave one; each call first waits for the default stream.  Per-stage
 * times (ms_ysweep .. ms_norms) are not measured in a replayed solve. */

/* Solver of the reduced system R (stage 3).  AUTO runs block-Jacobi sweeps and switches a
 * system to the direct solver when its sweeps stop contracting (update growing past 1e3 times
 * the iterate, non-finite, or not converged within reduced_maxit sweeps); DIRECT assembles R in
 * the reference's pattern and factors it with a banded LU with partial pivoting on the GPU (the
 * reference's PARDISO solve of R, spike_algebra.c:263-420); the factors of R stay resident for
 * the refinement passes and later solves until the matrix is factored again. */
#define SPIKE_RED_AUTO   0
#define SPIKE_RED_JACOBI 1
#define SPIKE_RED_DIRECT 2

typedef struct {
    int     n, ku, kl, b, P, nrhs;
    int64_t nnz;
    int     precision, dtype;
    int     refine_iters;      /* refinement corrections applied in the last solve          */
    int     reduced_iters;     /* block-Jacobi sweeps of the reduced system in the last solve */
    double  reduced_update;    /* last relative update of the reduced iteration             */
    double  max_rel_residual;  /* max over columns of ||b - A x|| / ||b|| (FP64, GPU)        */
    /* per-kernel times from HIP events on the handle's stream (milliseconds) */
    double  ms_convert;        /* CSR -> blocks                                               */
    double  ms_factor;         /* block LU (k_factor)                                         */
    double  ms_rtop;           /* top block rows of the U inverse when a separate launch forms
                                  them (k_rtop; otherwise the factor kernel does, inside ms_factor) */
    double  ms_spikes;         /* V / W spike tips (k_sweep)                                  */
    double  ms_red_setup;      /* reduced-system blocks (k_red_setup)                         */
    double  ms_ysweep;         /* RHS tips, summed over refinement passes (k_sweep)          */
    double  ms_reduced;        /* reduced-system iterations, summed                          */
    double  ms_recover;        /* recovery sweeps, summed (k_recover)                        */
    double  ms_residual;       /* FP64 residual launches, summed (k_residual)                */
    double  ms_factor_total;   /* spike_factor wall time                                     */
    double  ms_solve_total;    /* spike_solve_dev wall time                                  */
    size_t  device_bytes;      /* HBM held by the handle                                     */
    int     reduced_direct;    /* systems whose reduced system the direct band LU solved      */
    double  ms_norms;          /* per-column residual norm sums, summed (k_norm_reduce)      */
    int     solve_passes;      /* solve + refinement passes of the last column chunk         */
    int     launch_groups;     /* block-size groups: launches of each stage kernel per pass  */
    int     graph_replayed;    /* 1: the last solve was a replay of a captured graph          */
} spike_stats_t;

typedef struct spike_handle spike_handle_t;

void spike_default_opts(spike_opts_t *o);

int  spike_create(spike_handle_t **h, const spike_opts_t *o);
int  spike_destroy(spike_handle_t *h);
int  spike_set_opts(spike_handle_t *h, const spike_opts_t *o);

/* Bind a CSR matrix held in device memory; the pointers must stay valid until the next
 * spike_set_matrix_* or spike_destroy (zero-copy, like matrix_CreateFromComponents). */
int  spike_set_matrix_dev(spike_handle_t *h, int dtype, int n, int64_t nnz,
                          const int *d_rowptr, const int *d_colind, const void *d_aij);
/* Copy a host CSR to device memory owned by the handle. */
int  spike_set_matrix_host(spike_handle_t *h, int dtype, int n, int64_t nnz,
                           const int *rowptr, const int *colind, const void *aij);

/* Stages 0-2 for the matrix: bandwidth, schedule, CSR -> block conversion, block LU of every
 * partition, spike tips V/W, reduced-system setup.  Factors stay resident for later solves. */
int  spike_factor(spike_handle_t *h);

/* Solve for nrhs columns (device arrays, column-major, leading dims ldb/ldx >= n) with the
 * resident factors: RHS sweeps, FP64 reduced solve, recovery, FP64 residual and refinement.
 * Synchronous on return.  Calls spike_factor first if the matrix is not factored. */
int  spike_solve_dev(spike_handle_t *h, int nrhs, const void *d_b, int ldb, void *d_x, int ldx);

/* Host-array convenience wrapper: stages b to HBM, solves, copies x back. */
int  spike_solve_host(spike_handle_t *h, int nrhs, const void *b, int ldb, void *x, int ldx);

int  spike_get_stats(const spike_handle_t *h, spike_stats_t *s);

/* Work of the V / W spike chains in the last factorization, for throughput accounting:
 * w[0] = rows swept (per partition the union of its V and W chains), w[1] = the sum over those
 * rows of kl + ku + 1 (band entries the stage reads), w[2] = the sum over swept rows of
 * 2 (kl + ku) x the chain's spike columns.  A W chain stops once its spike has decayed below
 * 2^-70 of its first block and a V chain starts at the block its right-hand side reaches, so a
 * partition's chains sweep fewer rows than it holds.  Waits for the handle's stream. */
int  spike_get_spike_work(const spike_handle_t *h, double w[3]);

/* Schedule of the factored matrix: P and the first row of each partition (rows[P+1]). */
int  spike_get_schedule(const spike_handle_t *h, int *P, int *rows, int cap);

/* FP64 spike tips of the factored matrix and of the last solve's RHS, in the layout of
 * matrix_AssembleReducedSystem / block_AssembleReducedRHS (real dtype only). */
int  spike_get_tips(const spike_handle_t *h, double *Vt, double *Vb, double *Wt, double *Wb,
                    double *yt, double *yb);

/* Reduced-system solution of the last solve, N_r x nrhs column-major in the reference order
 * (per partition: x^t then x^b; spike_matrix.c:1254-1332). Real dtype, first chunk of RHS. */
int  spike_get_reduced_solution(const spike_handle_t *h, double *xr);

/* Set the HIP stream (hipStream_t passed as void*) the handle launches on; NULL = default. */
int  spike_set_stream(spike_handle_t *h, void *stream);

/* Static name of an error code. */
const char *spike_error_string(int code);
/* Detail message of the last error raised on the calling thread. */
const char *spike_last_error_message(void);

/* ---------------- batched solve of independent systems (one set of launches) ----------- */
typedef struct {
    int         n;
    int64_t     nnz;
    const int  *d_rowptr;
    const int  *d_colind;
    const void *d_aij;      /* dtype of the batch */
    int         nrhs;
    const void *d_b;        /* n x nrhs, ld = n */
    void       *d_x;        /* n x nrhs, ld = n */
} spike_system_t;

typedef struct spike_batch spike_batch_t;

int  spike_batch_create(spike_batch_t **bt, const spike_opts_t *o);
int  spike_batch_destroy(spike_batch_t *bt);
/* Factor and solve nsys systems; partitions of every system share the stage launches.
 * P per system is chosen from (n, k, nrhs, dtype) only, never from the batch mix. */
int  spike_batch_solve_dev(spike_batch_t *bt, int dtype, int nsys, const spike_system_t *sys);
int  spike_batch_get_stats(const spike_batch_t *bt, spike_stats_t *s,
                           double *max_rel_residual_per_system);
/* spike_get_spike_work for the batch's last factorization (its last window of systems when the
 * batch was solved in windows). */
int  spike_batch_get_spike_work(const spike_batch_t *bt, double w[3]);

/* Partition count the autotuner picks for one system (pure function of its arguments);
 * spike_autotune_partitions is the LATENCY mode. */
int  spike_autotune_partitions(int n, int ku, int kl, int nrhs, int dtype, int precision);
int  spike_autotune_partitions_tuned(int n, int ku, int kl, int nrhs, int dtype, int precision, int tune);

/* Relative residual ||A x - b|| / ||b|| of every column (host arrays, FP64 on the GPU). */
int  spike_residual_host(int dtype, int n, int64_t nnz, const int *rowptr, const int *colind,
                         const void *aij, int nrhs, const void *x, const void *b, double *rel);

/* Direct solve of a banded CSR matrix (host arrays) on the GPU: the band (ku, kl from the
 * stored entries) is factored by LU with partial pivoting (LAPACK gbtrf semantics: largest
 * modulus in the column, first on ties), the columns are solved, and FP64 refinement steps
 * follow while the largest relative residual halves (at most three).  rel (may be NULL)
 * receives every column's final ||b - A x|| / ||b||.  kl + (kl + ku) must stay below 4096
 * (FP64) or 2048 (complex).  Returns SPIKE_ERR_SINGULAR when a column has no nonzero pivot.
 * Backs directSolver_Host_Solve (the reduced-system solver slot, spike_algebra.h:157-164). */
int  spike_band_solve_host(int dtype, int n, int64_t nnz, const int *rowptr, const int *colind,
                           const void *aij, int nrhs, const void *b, int ldb, void *x, int ldx,
                           double *rel);

/* Allocation-free variant of spike_band_solve_host.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same computation and result bits as spike_band_solve_host,
 * on device arrays (CSR, b with leading dimension ldb, x with ldx) and on `stream` (a
 * hipStream_t, NULL = default); rel (host, nrhs entries) is required.  All workspace -- band,
 * pivots, residual panel, descriptors -- lives in a static device area of
 * SPIKE_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a solve that needs more
 * (about n * (2 kl + ku + 1 + nrhs) values plus small descriptors) returns SPIKE_ERR_OOM.  The
 * area keeps the last call's contents and is shared by every call on that GPU: calls that
 * overlap on one GPU race and must be serialized by the caller. */
#define SPIKE_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
int  spike_band_solve_dev_noalloc(int dtype, int n, int64_t nnz, const int *d_rowptr,
                                  const int *d_colind, const void *d_aij, int nrhs, const void *d_b,
                                  int ldb, void *d_x, int ldx, double *rel, void *stream);

/* ---------------- request server: continuous batching, memory admission, resident factors --
 * Requests (host arrays, kept alive and unmodified by the caller until spike_server_wait returns)
 * are admitted as worker slots and device memory free up; new matrices of one dtype are factored
 * and solved together (up to max_batch per job), jobs on different workers overlap on the GPU,
 * and every factored matrix stays resident in HBM (least-recently-used eviction) keyed by a
 * 64-bit fingerprint of its CSR or the caller's key, so a repeat request only runs the solve.
 * A request's result does not depend on what else is in flight.  No counterpart in the
 * reference (its driver solves one system per call). */
typedef struct {
    spike_opts_t opts;            /* precision, tolerance, refinement cap of every solve         */
    int          workers;         /* concurrent jobs (one HIP stream each), default 3            */
    int          max_batch;       /* new matrices per factor job, default 16                      */
    uint64_t     mem_budget_bytes;/* device bytes for resident factors + jobs; 0 = 85% of free   */
} spike_server_opts_t;

typedef struct {
    int          dtype;           /* SPIKE_DTYPE_D or SPIKE_DTYPE_Z                               */
    int          n;
    int64_t      nnz;
    const int   *rowptr;          /* host CSR, int32, 0-based                                     */
    const int   *colind;
    const void  *aij;
    int          nrhs;
    const void  *b;               /* host n x nrhs column-major                                   */
    void        *x;               /* host n x nrhs column-major, written on completion            */
    uint64_t     key;             /* 0: fingerprint the CSR; otherwise the caller's matrix key    */
} spike_request_t;

typedef struct {
    int      status;              /* SPIKE_SUCCESS, SPIKE_ERR_NOT_CONVERGED, or an error          */
    double   latency_ms;          /* submit -> solution copied back                               */
    double   max_rel_residual;
    int      resident_hit;        /* 1: solved on resident factors                                */
    uint64_t key;
} spike_request_result_t;

typedef struct {
    int64_t  completed, failed, resident_hits, rhs_solved;
    double   p50_ms, p99_ms, mean_ms, elapsed_ms;
    int64_t  resident_matrices;
    uint64_t resident_bytes, peak_bytes, budget_bytes;
    int64_t  oom_retries;         /* jobs that ran out of device memory and were run again        */
    double   est_ratio_max;       /* largest (device bytes a fresh job held) / (its admission
                                     estimate); <= 1 means every estimate covered its job       */
} spike_server_stats_t;

typedef struct spike_server spike_server_t;

void spike_server_default_opts(spike_server_opts_t *o);
int  spike_server_create(spike_server_t **s, const spike_server_opts_t *o);
/* Waits for every submitted request, then stops the workers. */
int  spike_server_destroy(spike_server_t *s);
int  spike_server_submit(spike_server_t *s, const spike_request_t *r, uint64_t *ticket);
/* Blocks until the request is done (timeout_ms < 0: no limit; returns 1 on timeout, else the
 * request's status); the ticket is released on return. */
int  spike_server_wait(spike_server_t *s, uint64_t ticket, double timeout_ms, spike_request_result_t *res);
int  spike_server_get_stats(spike_server_t *s, spike_server_stats_t *st);
int  spike_server_reset_stats(spike_server_t *s);

/* ---------------- synthetic generators (splitmix64 per row; deterministic) -------------- */
/* Banded matrix: each in-band off-diagonal entry present with probability `fill`
 * (fill >= 1 means all), values U[-1,1] (complex: re and im U[-1,1]); diagonal =
 * max(d_i * sum_j |a_ij| + diag_add, diag_floor) with d_i ~ U[dom_lo, dom_hi].
 * Pass NULL arrays to count: *nnz_out receives nnz. rowptr[n+1], colind/aij[nnz]. */
int  spike_gen_banded(int dtype, int n, int ku, int kl, double fill, double dom_lo, double dom_hi,
                      double diag_add, double diag_floor, uint64_t seed,
                      int *rowptr, int *colind, void *aij, int64_t *nnz_out);
/* Dense column-major n x nrhs values U[lo, hi] (complex: both parts). */
int  spike_gen_dense(int dtype, int64_t count, double lo, double hi, uint64_t seed, void *out);

/* ---------------- bandwidth-reducing reordering (reverse Cuthill-McKee) ------------------ */
/* perm[new] = old.  Works on the symmetrized pattern of A.  Returns the new bandwidths. */
int  spike_rcm_order(int n, const int *rowptr, const int *colind, int *perm, int *ku, int *kl);
/* B = P A P^T as CSR (host arrays), rowptr_out[n+1], colind_out/aij_out[nnz]. */
int  spike_permute_csr(int dtype, int n, const int *rowptr, const int *colind, const void *aij,
                       const int *perm, int *rowptr_out, int *colind_out, void *aij_out);

#ifdef __cplusplus
}
#endif
#endif
