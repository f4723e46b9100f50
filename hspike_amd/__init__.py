"""hspike_amd -- MI355X (gfx950) SPIKE solver for banded linear systems.

Python mirror of the reference's C interface (srodrb/HSPIKE OpenMP/spike_interfaces.h:28-72,
OpenMP/spike_superlu.h:107-135) over the C ABI of libhspike.so; every solve runs the HIP
kernels of hspike_amd/csrc.  Host arrays are numpy; device arrays may be torch tensors on a
ROCm device (only their data pointers are used).

  dspike_core_host(A, B) / zspike_core_host / sspike_core_host / cspike_core_host
  Solver            factor-once / solve-many handle (precision, partitions, tips, stats)
  Batch             many independent systems of uneven size in one set of launches
  Server            request loop: continuous batching, memory admission, resident factors
  gen_banded        deterministic synthetic banded matrices (splitmix64 per row)
  rcm / permute     reverse Cuthill-McKee reordering of a general sparse matrix
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import SPIKE_DTYPE_D, SPIKE_DTYPE_Z, SPIKE_PREC_FP64, SPIKE_PREC_MIXED, SpikeError, check, load

__all__ = [
    "CSRMatrix", "Solver", "Batch", "PreparedBatch", "Server", "gen_banded", "gen_dense", "dspike_core_host", "dspike_core_host_blocking",
    "zspike_core_host", "sspike_core_host", "cspike_core_host", "residual", "rcm", "permute", "autotune_partitions",
    "SpikeError", "load",
]


def _ptr(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


class CSRMatrix:
    """Square CSR matrix, int32 0-based indices (OpenMP/spike_matrix.h:39-53)."""

    def __init__(self, n, rowptr, colind, aij):
        self.n = int(n)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.colind = np.ascontiguousarray(colind, dtype=np.int32)
        aij = np.asarray(aij)
        self.aij = np.ascontiguousarray(aij, dtype=np.complex128 if np.iscomplexobj(aij) else np.float64)

    @property
    def nnz(self):
        return int(self.colind.size)

    @property
    def dtype_code(self):
        return SPIKE_DTYPE_Z if np.iscomplexobj(self.aij) else SPIKE_DTYPE_D

    def bandwidth(self):
        ku, kl = ctypes.c_int(), ctypes.c_int()
        check(load().matrix_ComputeBandwidth(self.n, _ptr(self.colind), _ptr(self.rowptr), None, ctypes.byref(ku),
                                             ctypes.byref(kl)), "matrix_ComputeBandwidth")
        return ku.value, kl.value

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.aij, self.colind, self.rowptr), shape=(self.n, self.n))

    @classmethod
    def from_scipy(cls, M):
        M = M.tocsr()
        return cls(M.shape[0], M.indptr, M.indices, M.data)


def gen_banded(n, ku, kl=None, fill=1.0, dom=(1.05, 1.05), diag_add=0.0, diag_floor=1.0, seed=0, complex=False):
    """Synthetic banded matrix: off-diagonals present with probability `fill`, U[-1,1];
    diagonal max(d_r * sum|a_rj| + diag_add, diag_floor) with d_r ~ U[dom[0], dom[1]]."""
    kl = ku if kl is None else kl
    dt = SPIKE_DTYPE_Z if complex else SPIKE_DTYPE_D
    L = load()
    nnz = ctypes.c_int64()
    check(L.spike_gen_banded(dt, n, ku, kl, fill, dom[0], dom[1], diag_add, diag_floor, seed, None, None, None,
                             ctypes.byref(nnz)), "spike_gen_banded")
    rowptr = np.zeros(n + 1, np.int32)
    colind = np.zeros(max(nnz.value, 1), np.int32)
    aij = np.zeros(max(nnz.value, 1), np.complex128 if complex else np.float64)
    check(L.spike_gen_banded(dt, n, ku, kl, fill, dom[0], dom[1], diag_add, diag_floor, seed, _ptr(rowptr),
                             _ptr(colind), _ptr(aij), ctypes.byref(nnz)), "spike_gen_banded")
    return CSRMatrix(n, rowptr, colind[:nnz.value], aij[:nnz.value])


def gen_dense(n, m, lo=-1.0, hi=1.0, seed=0, complex=False):
    """Dense column-major (Fortran-order) n x m panel with U[lo, hi] entries."""
    out = np.zeros((n, m), dtype=np.complex128 if complex else np.float64, order="F")
    check(load().spike_gen_dense(SPIKE_DTYPE_Z if complex else SPIKE_DTYPE_D, n * m, lo, hi, seed, _ptr(out)),
          "spike_gen_dense")
    return out


SPIKE_TUNE_LATENCY, SPIKE_TUNE_THROUGHPUT = 0, 1


def _tune_code(tune):
    codes = {"latency": SPIKE_TUNE_LATENCY, "throughput": SPIKE_TUNE_THROUGHPUT}
    if tune in codes.values():
        return int(tune)
    if tune not in codes:
        raise ValueError("tune must be 'latency' or 'throughput'")
    return codes[tune]


def autotune_partitions(n, ku, kl, nrhs=1, complex=False, precision=SPIKE_PREC_FP64, tune="latency"):
    return load().spike_autotune_partitions_tuned(n, ku, kl, nrhs, SPIKE_DTYPE_Z if complex else SPIKE_DTYPE_D,
                                                  precision, _tune_code(tune))


def _prec_code(precision):
    if precision in (SPIKE_PREC_FP64, "fp64", "double"):
        return SPIKE_PREC_FP64
    if precision in (SPIKE_PREC_MIXED, "mixed", "fp32+fp64"):
        return SPIKE_PREC_MIXED
    raise ValueError("precision must be 'fp64' or 'mixed'")


SPIKE_RED_AUTO, SPIKE_RED_JACOBI, SPIKE_RED_DIRECT = 0, 1, 2


def _red_code(solver):
    codes = {"auto": SPIKE_RED_AUTO, "jacobi": SPIKE_RED_JACOBI, "direct": SPIKE_RED_DIRECT}
    if solver in codes.values():
        return int(solver)
    if solver not in codes:
        raise ValueError("reduced_solver must be 'auto', 'jacobi' or 'direct'")
    return codes[solver]


class Solver:
    """Factor-once / solve-many GPU SPIKE handle (include/spike_ext.h)."""

    def __init__(self, precision="fp64", partitions=0, tol=1e-10, max_refine=8, reduced_maxit=64,
                 reduced_tol=1e-15, verbose=False, reduced_solver="auto", graph=False, tune="latency"):
        self._L = load()
        self._opts = _lib.default_opts(precision=_prec_code(precision), partitions=int(partitions), tol=float(tol),
                                       max_refine=int(max_refine), reduced_maxit=int(reduced_maxit),
                                       reduced_tol=float(reduced_tol), verbose=int(bool(verbose)),
                                       reduced_solver=_red_code(reduced_solver), graph=int(bool(graph)),
                                       tune=_tune_code(tune))
        h = ctypes.c_void_p()
        check(self._L.spike_create(ctypes.byref(h), ctypes.byref(self._opts)), "spike_create")
        self._h = h
        self._keep = ()
        self.n = 0
        self.complex = False

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.spike_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_matrix(self, A):
        """Copy a host CSRMatrix to HBM."""
        self.n, self.complex = A.n, A.dtype_code == SPIKE_DTYPE_Z
        check(self._L.spike_set_matrix_host(self._h, A.dtype_code, A.n, A.nnz, _ptr(A.rowptr), _ptr(A.colind),
                                            _ptr(A.aij)), "spike_set_matrix_host")

    def set_matrix_dev(self, n, rowptr, colind, aij):
        """Bind CSR arrays already in device memory (torch tensors; kept alive by the solver)."""
        import torch
        cplx = aij.dtype == torch.complex128
        self.n, self.complex = int(n), cplx
        self._keep = (rowptr, colind, aij)
        check(self._L.spike_set_matrix_dev(self._h, SPIKE_DTYPE_Z if cplx else SPIKE_DTYPE_D, int(n), colind.numel(),
                                           _ptr(rowptr), _ptr(colind), _ptr(aij)), "spike_set_matrix_dev")

    def factor(self):
        check(self._L.spike_factor(self._h), "spike_factor")

    def solve(self, B, allow_unconverged=False):
        """Solve for host panel B (n or n x m); returns X (numpy, Fortran order)."""
        dt = np.complex128 if self.complex else np.float64
        B = np.asfortranarray(np.asarray(B, dtype=dt).reshape(self.n, -1))
        X = np.zeros_like(B, order="F")
        rc = self._L.spike_solve_host(self._h, B.shape[1], _ptr(B), self.n, _ptr(X), self.n)
        if rc != 0 and not (allow_unconverged and rc == -5):
            check(rc, "spike_solve_host")
        return X

    def solve_dev(self, B, X, nrhs=None):
        """Solve with device panels B, X (torch tensors, column-major n x m, i.e. shape (m, n) rows)."""
        m = nrhs if nrhs is not None else (B.numel() // self.n)
        check(self._L.spike_solve_dev(self._h, m, _ptr(B), self.n, _ptr(X), self.n), "spike_solve_dev")

    def stats(self):
        s = _lib.Stats()
        check(self._L.spike_get_stats(self._h, ctypes.byref(s)), "spike_get_stats")
        return s.as_dict()

    def spike_work(self):
        """Work of the V / W spike chains in the last factorization: (rows swept, band entries read,
        flops) -- a W chain stops once its spike has decayed, a V chain starts where its
        right-hand side is nonzero (spike_get_spike_work)."""
        w = np.zeros(3)
        check(self._L.spike_get_spike_work(self._h, _ptr(w)), "spike_get_spike_work")
        return tuple(float(v) for v in w)

    def schedule(self):
        P = ctypes.c_int()
        check(self._L.spike_get_schedule(self._h, ctypes.byref(P), None, 0), "spike_get_schedule")
        rows = np.zeros(P.value + 1, np.int32)
        check(self._L.spike_get_schedule(self._h, ctypes.byref(P), _ptr(rows), P.value + 1), "spike_get_schedule")
        return rows

    def _tip_cols(self, nrhs):
        # the library writes the RHS tips of the last solve's first column chunk (<= 64 columns)
        m = min(self.stats()["nrhs"], 64)
        if nrhs is not None and nrhs > m:
            raise ValueError("the last solve had %d right-hand-side columns" % m)
        return m

    def tips(self, nrhs=None):
        """FP64 spike tips in the oracle's layout: dict name -> [P, rows, cols] (RHS tips: the first
        nrhs columns of the last solve, all of them by default)."""
        m = self._tip_cols(nrhs)
        st = self.stats()
        P, ku, kl = st["P"], st["ku"], st["kl"]
        shapes = {"Vt": (ku, ku), "Vb": (kl, ku), "Wt": (ku, kl), "Wb": (kl, kl), "yt": (ku, m), "yb": (kl, m)}
        bufs = {k: np.zeros(P * r * c + 1) for k, (r, c) in shapes.items()}
        check(self._L.spike_get_tips(self._h, *[_ptr(bufs[k]) for k in ("Vt", "Vb", "Wt", "Wb", "yt", "yb")]),
              "spike_get_tips")
        out = {k: bufs[k][:P * r * c].reshape(P, c, r).transpose(0, 2, 1) for k, (r, c) in shapes.items()}
        if nrhs is not None:
            out["yt"], out["yb"] = out["yt"][:, :, :nrhs], out["yb"][:, :, :nrhs]
        return out

    def reduced_solution(self, nrhs=None):
        """Reduced solution xr (reference order, N_r x columns) of the last solve's first columns."""
        m = self._tip_cols(nrhs)
        st = self.stats()
        Nr = st["P"] * (st["ku"] + st["kl"])
        xr = np.zeros(Nr * m)
        check(self._L.spike_get_reduced_solution(self._h, _ptr(xr)), "spike_get_reduced_solution")
        return xr.reshape(m, Nr).T[:, :(m if nrhs is None else nrhs)]


class Batch:
    """Factor and solve many independent device-resident systems together (include/spike_ext.h)."""

    def __init__(self, precision="mixed", tol=1e-10, max_refine=8, verbose=False, partitions=0, graph=False,
                 tune="latency"):
        self._L = load()
        self._opts = _lib.default_opts(precision=_prec_code(precision), tol=float(tol), max_refine=int(max_refine),
                                       verbose=int(bool(verbose)), partitions=int(partitions), graph=int(bool(graph)),
                                       tune=_tune_code(tune))
        h = ctypes.c_void_p()
        check(self._L.spike_batch_create(ctypes.byref(h), ctypes.byref(self._opts)), "spike_batch_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.spike_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prepare(self, systems):
        """The device pointers of `systems` (see solve_dev) as one descriptor array, built once for
        repeated solves of the same batch: building it costs about 4 us per system in Python."""
        return PreparedBatch(systems)

    def solve_dev(self, systems, complex=False, allow_unconverged=False):
        """systems: list of dicts with n, rowptr, colind, aij, b, x (torch device tensors), nrhs;
        or the PreparedBatch of such a list (Batch.prepare)."""
        prep = systems if isinstance(systems, PreparedBatch) else PreparedBatch(systems)
        rc = self._L.spike_batch_solve_dev(self._h, SPIKE_DTYPE_Z if complex else SPIKE_DTYPE_D, prep.count, prep.arr)
        if rc != 0 and not (allow_unconverged and rc == -5):
            check(rc, "spike_batch_solve_dev")
        rel = np.zeros(prep.count)
        st = _lib.Stats()
        check(self._L.spike_batch_get_stats(self._h, ctypes.byref(st), _ptr(rel)), "spike_batch_get_stats")
        return rel, st.as_dict()

    def spike_work(self):
        """Solver.spike_work for the batch's last factorization (spike_batch_get_spike_work)."""
        w = np.zeros(3)
        check(self._L.spike_batch_get_spike_work(self._h, _ptr(w)), "spike_batch_get_spike_work")
        return tuple(float(v) for v in w)


class PreparedBatch:
    """Descriptor array of a batch of device systems (spike_system_t per system).  It keeps the
    systems' tensors referenced, so the pointers it holds stay valid; a tensor replaced in a
    system's dict afterwards is not seen (prepare the batch again)."""

    def __init__(self, systems):
        self.systems = list(systems)
        self.count = len(self.systems)
        self.arr = (_lib.System * self.count)()
        for i, s in enumerate(self.systems):
            a = self.arr[i]
            a.n = s["n"]
            a.nnz = s["colind"].numel()
            a.d_rowptr = s["rowptr"].data_ptr()
            a.d_colind = s["colind"].data_ptr()
            a.d_aij = s["aij"].data_ptr()
            a.nrhs = s["nrhs"]
            a.d_b = s["b"].data_ptr()
            a.d_x = s["x"].data_ptr()


class Server:
    """Request server (include/spike_ext.h, spike_server_*): continuous batching of arriving solves,
    admission by device memory, factors kept resident for repeated right-hand sides.

        srv = Server("mixed", workers=3, max_batch=16)
        t = srv.submit(A, B)          # returns a ticket; A, B stay referenced until wait()
        X, info = srv.wait(t)         # info: status, latency_ms, max_rel_residual, resident_hit
    """

    def __init__(self, precision="mixed", tol=1e-10, max_refine=8, workers=3, max_batch=16, mem_budget_gb=0.0,
                 verbose=False):
        self._L = load()
        o = _lib.ServerOpts()
        self._L.spike_server_default_opts(ctypes.byref(o))
        o.opts = _lib.default_opts(precision=_prec_code(precision), tol=float(tol), max_refine=int(max_refine),
                                   verbose=int(bool(verbose)))
        o.workers, o.max_batch = int(workers), int(max_batch)
        o.mem_budget_bytes = int(mem_budget_gb * 2**30)
        h = ctypes.c_void_p()
        check(self._L.spike_server_create(ctypes.byref(h), ctypes.byref(o)), "spike_server_create")
        self._h = h
        self._live = {}

    def submit(self, A, B, key=0):
        cplx = A.dtype_code == SPIKE_DTYPE_Z
        dt = np.complex128 if cplx else np.float64
        B = np.asfortranarray(np.asarray(B, dtype=dt).reshape(A.n, -1))
        X = np.zeros_like(B, order="F")
        r = _lib.Request(dtype=A.dtype_code, n=A.n, nnz=A.nnz, rowptr=A.rowptr.ctypes.data, colind=A.colind.ctypes.data,
                         aij=A.aij.ctypes.data, nrhs=B.shape[1], b=B.ctypes.data, x=X.ctypes.data, key=int(key))
        t = ctypes.c_uint64()
        check(self._L.spike_server_submit(self._h, ctypes.byref(r), ctypes.byref(t)), "spike_server_submit")
        self._live[t.value] = (A, B, X)
        return t.value

    def wait(self, ticket, timeout_ms=-1.0):
        res = _lib.RequestResult()
        rc = self._L.spike_server_wait(self._h, ctypes.c_uint64(ticket), float(timeout_ms), ctypes.byref(res))
        if rc == 1:
            return None, None
        A, B, X = self._live.pop(ticket)
        info = {"status": res.status, "latency_ms": res.latency_ms, "max_rel_residual": res.max_rel_residual,
                "resident_hit": bool(res.resident_hit), "key": res.key}
        return X, info

    def stats(self):
        st = _lib.ServerStats()
        check(self._L.spike_server_get_stats(self._h, ctypes.byref(st)), "spike_server_get_stats")
        return st.as_dict()

    def reset_stats(self):
        check(self._L.spike_server_reset_stats(self._h), "spike_server_reset_stats")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.spike_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _legacy(fn_name, A, B, dt_np, cast_a=None):
    L = load()
    B = np.asfortranarray(np.asarray(B, dtype=dt_np).reshape(A.n, -1))
    X = np.zeros_like(B, order="F")
    aij = np.ascontiguousarray(A.aij.astype(dt_np)) if cast_a is None else cast_a
    rc = getattr(L, fn_name)(A.n, A.nnz, B.shape[1], _ptr(A.colind), _ptr(A.rowptr), _ptr(aij), _ptr(X), _ptr(B))
    check(rc, fn_name)
    return X


def dspike_core_host(A, B):
    """X = A^{-1} B, FP64 (OpenMP/spike_interfaces.h:37-44)."""
    return _legacy("dspike_core_host", A, B, np.float64)


def dspike_core_host_blocking(A, B):
    """Same with RHS columns processed in blocks of 200 (spike_interfaces.h:65-72)."""
    return _legacy("dspike_core_host_blocking", A, B, np.float64)


def zspike_core_host(A, B):
    """complex128 (spike_interfaces.h:55-62)."""
    return _legacy("zspike_core_host", A, B, np.complex128)


def sspike_core_host(A, B):
    """FP32 arrays; FP32 factorization with FP64 refinement (spike_interfaces.h:28-35)."""
    return _legacy("sspike_core_host", A, B, np.float32)


def cspike_core_host(A, B):
    """complex64 arrays (spike_interfaces.h:46-53)."""
    return _legacy("cspike_core_host", A, B, np.complex64)


def residual(A, X, B):
    """Per-column ||A x - b|| / ||b|| computed in FP64 on the GPU."""
    dt = np.complex128 if (A.dtype_code == SPIKE_DTYPE_Z or np.iscomplexobj(B)) else np.float64
    X = np.asfortranarray(np.asarray(X, dtype=dt).reshape(A.n, -1))
    B = np.asfortranarray(np.asarray(B, dtype=dt).reshape(A.n, -1))
    aij = np.ascontiguousarray(A.aij, dtype=dt)
    rel = np.zeros(X.shape[1])
    check(load().spike_residual_host(SPIKE_DTYPE_Z if dt == np.complex128 else SPIKE_DTYPE_D, A.n, A.nnz,
                                     _ptr(A.rowptr), _ptr(A.colind), _ptr(aij), X.shape[1], _ptr(X), _ptr(B),
                                     _ptr(rel)), "spike_residual_host")
    return rel


def band_solve(A, B):
    """X = A^{-1} B by the GPU band LU with partial pivoting and FP64 refinement
    (spike_band_solve_host); returns (X, per-column relative residuals)."""
    dt = np.complex128 if (A.dtype_code == SPIKE_DTYPE_Z or np.iscomplexobj(B)) else np.float64
    B = np.asfortranarray(np.asarray(B, dtype=dt).reshape(A.n, -1))
    X = np.zeros_like(B, order="F")
    aij = np.ascontiguousarray(A.aij, dtype=dt)
    rel = np.zeros(B.shape[1])
    check(load().spike_band_solve_host(SPIKE_DTYPE_Z if dt == np.complex128 else SPIKE_DTYPE_D, A.n, A.nnz,
                                       _ptr(A.rowptr), _ptr(A.colind), _ptr(aij), B.shape[1], _ptr(B), A.n, _ptr(X),
                                       A.n, _ptr(rel)), "spike_band_solve_host")
    return X, rel


def host_solve(A, B):
    """The reduced-system solver slot of the reference (directSolver_Host_Solve, spike_algebra.h:157-164)."""
    return _legacy("directSolver_Host_Solve", A, B, np.float64)


def rcm(A):
    """Reverse Cuthill-McKee order: returns (perm with perm[new] = old, ku, kl of P A P^T)."""
    perm = np.zeros(A.n, np.int32)
    ku, kl = ctypes.c_int(), ctypes.c_int()
    check(load().spike_rcm_order(A.n, _ptr(A.rowptr), _ptr(A.colind), _ptr(perm), ctypes.byref(ku), ctypes.byref(kl)),
          "spike_rcm_order")
    return perm, ku.value, kl.value


def permute(A, perm):
    """P A P^T as a new CSRMatrix (rows and columns renumbered by perm[new] = old)."""
    perm = np.ascontiguousarray(perm, dtype=np.int32)
    rp = np.zeros(A.n + 1, np.int32)
    ci = np.zeros(max(A.nnz, 1), np.int32)
    av = np.zeros(max(A.nnz, 1), A.aij.dtype)
    check(load().spike_permute_csr(A.dtype_code, A.n, _ptr(A.rowptr), _ptr(A.colind), _ptr(A.aij), _ptr(perm),
                                   _ptr(rp), _ptr(ci), _ptr(av)), "spike_permute_csr")
    return CSRMatrix(A.n, rp, ci[:A.nnz], av[:A.nnz])
