"""ctypes binding of libhspike.so (the C ABI declared in include/*.h).

The library is loaded from hspike_amd/lib/libhspike.so (built in-tree by hspike_amd.build).
There is no fallback: if the library is missing, loading raises, and every solve goes
through the HIP kernels.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libhspike.so")

SPIKE_DTYPE_D = 0
SPIKE_DTYPE_Z = 1
SPIKE_PREC_FP64 = 0
SPIKE_PREC_MIXED = 1

ERRORS = {
    0: "success", -1: "invalid argument", -2: "malformed CSR", -3: "out of memory",
    -4: "singular pivot", -5: "not converged", -6: "unsupported", -7: "device error",
    -8: "io error", -9: "cancelled",
}


class SpikeError(RuntimeError):
    def __init__(self, code, where, detail=""):
        self.code = code
        msg = "%s failed: %s (%d)" % (where, ERRORS.get(code, "unknown"), code)
        if detail:
            msg += " -- " + detail
        super().__init__(msg)


class Opts(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int), ("partitions", ctypes.c_int), ("max_refine", ctypes.c_int),
        ("tol", ctypes.c_double), ("reduced_maxit", ctypes.c_int), ("reduced_tol", ctypes.c_double),
        ("sweep_waves", ctypes.c_int), ("verbose", ctypes.c_int), ("reduced_solver", ctypes.c_int),
        ("graph", ctypes.c_int), ("tune", ctypes.c_int),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("ku", ctypes.c_int), ("kl", ctypes.c_int), ("b", ctypes.c_int),
        ("P", ctypes.c_int), ("nrhs", ctypes.c_int), ("nnz", ctypes.c_int64),
        ("precision", ctypes.c_int), ("dtype", ctypes.c_int), ("refine_iters", ctypes.c_int),
        ("reduced_iters", ctypes.c_int), ("reduced_update", ctypes.c_double),
        ("max_rel_residual", ctypes.c_double), ("ms_convert", ctypes.c_double),
        ("ms_factor", ctypes.c_double), ("ms_rtop", ctypes.c_double), ("ms_spikes", ctypes.c_double),
        ("ms_red_setup", ctypes.c_double), ("ms_ysweep", ctypes.c_double), ("ms_reduced", ctypes.c_double),
        ("ms_recover", ctypes.c_double), ("ms_residual", ctypes.c_double),
        ("ms_factor_total", ctypes.c_double), ("ms_solve_total", ctypes.c_double),
        ("device_bytes", ctypes.c_size_t), ("reduced_direct", ctypes.c_int), ("ms_norms", ctypes.c_double),
        ("solve_passes", ctypes.c_int), ("launch_groups", ctypes.c_int), ("graph_replayed", ctypes.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class System(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("nnz", ctypes.c_int64), ("d_rowptr", ctypes.c_void_p),
        ("d_colind", ctypes.c_void_p), ("d_aij", ctypes.c_void_p), ("nrhs", ctypes.c_int),
        ("d_b", ctypes.c_void_p), ("d_x", ctypes.c_void_p),
    ]


class ServerOpts(ctypes.Structure):
    _fields_ = [("opts", Opts), ("workers", ctypes.c_int), ("max_batch", ctypes.c_int),
                ("mem_budget_bytes", ctypes.c_uint64)]


class Request(ctypes.Structure):
    _fields_ = [
        ("dtype", ctypes.c_int), ("n", ctypes.c_int), ("nnz", ctypes.c_int64), ("rowptr", ctypes.c_void_p),
        ("colind", ctypes.c_void_p), ("aij", ctypes.c_void_p), ("nrhs", ctypes.c_int), ("b", ctypes.c_void_p),
        ("x", ctypes.c_void_p), ("key", ctypes.c_uint64),
    ]


class RequestResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("latency_ms", ctypes.c_double), ("max_rel_residual", ctypes.c_double),
                ("resident_hit", ctypes.c_int), ("key", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("completed", ctypes.c_int64), ("failed", ctypes.c_int64), ("resident_hits", ctypes.c_int64),
        ("rhs_solved", ctypes.c_int64), ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
        ("mean_ms", ctypes.c_double), ("elapsed_ms", ctypes.c_double), ("resident_matrices", ctypes.c_int64),
        ("resident_bytes", ctypes.c_uint64), ("peak_bytes", ctypes.c_uint64), ("budget_bytes", ctypes.c_uint64),
        ("oom_retries", ctypes.c_int64), ("est_ratio_max", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# every symbol declared in include/*.h (checked by the CPU test suite)
EXPORTED = [
    # spike_interfaces.h
    "sspike_core_host", "dspike_core_host", "cspike_core_host", "zspike_core_host", "dspike_core_host_blocking",
    # spike_matrix.h
    "matrix_LoadCSR", "matrix_SaveCSR", "matrix_LoadCSR_z", "matrix_LoadCSR_zl", "matrix_Deallocate", "matrix_z_Deallocate",
    "matrix_CreateEmptyMatrix", "matrix_CreateFromComponents", "matrix_AreEqual", "matrix_CheckCSR",
    "matrix_ComputeBandwidth", "matrix_ExtractMatrix", "matrix_ExtractBlock", "block_CreateEmptyBlock",
    "block_Deallocate", "spike_solve_analysis", "spike_solve_analysis_p", "schedule_Destroy",
    "matrix_CreateEmptyReducedSystem",
    "matrix_AssembleReducedSystem", "block_AssembleReducedRHS",
    # spike_memory.h
    "_spike_malloc", "spike_free", "spike_nullify", "spike_memory_is_pinned",
    # spike_hip.h
    "directSolver_CreateHandler", "directSolver_Configure", "directSolver_Factorize", "directSolver_SolveForRHS",
    "directSolver_ShowStatistics", "directSolver_Finalize", "directSolver_Solve", "directSolver_Host_Solve",
    "ComputeResidualOfLinearSystem",
    # spike_ext.h
    "spike_default_opts", "spike_create", "spike_destroy", "spike_set_opts", "spike_set_matrix_dev",
    "spike_set_matrix_host", "spike_factor", "spike_solve_dev", "spike_solve_host", "spike_get_stats",
    "spike_get_spike_work", "spike_batch_get_spike_work",
    "spike_get_schedule", "spike_get_tips", "spike_get_reduced_solution", "spike_set_stream",
    "spike_error_string", "spike_last_error_message", "spike_batch_create", "spike_batch_destroy", "spike_batch_solve_dev",
    "spike_batch_get_stats", "spike_autotune_partitions", "spike_autotune_partitions_tuned", "spike_residual_host", "spike_band_solve_host",
    "spike_band_solve_dev_noalloc",
    "spike_gen_banded",
    "spike_gen_dense", "spike_rcm_order", "spike_permute_csr",
    "spike_server_default_opts", "spike_server_create", "spike_server_destroy", "spike_server_submit",
    "spike_server_wait", "spike_server_get_stats", "spike_server_reset_stats",
]

_LIB = None


def load():
    """Load libhspike.so and declare the C signatures (raises if it was not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise OSError("libhspike.so not built: run `python -m hspike_amd.build` (expected %s)" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
    sig = {
        "spike_default_opts": (None, [ctypes.POINTER(Opts)]),
        "spike_create": (i32, [ctypes.POINTER(vp), ctypes.POINTER(Opts)]),
        "spike_destroy": (i32, [vp]),
        "spike_set_opts": (i32, [vp, ctypes.POINTER(Opts)]),
        "spike_set_matrix_dev": (i32, [vp, i32, i32, i64, vp, vp, vp]),
        "spike_set_matrix_host": (i32, [vp, i32, i32, i64, vp, vp, vp]),
        "spike_factor": (i32, [vp]),
        "spike_solve_dev": (i32, [vp, i32, vp, i32, vp, i32]),
        "spike_solve_host": (i32, [vp, i32, vp, i32, vp, i32]),
        "spike_get_stats": (i32, [vp, ctypes.POINTER(Stats)]),
        "spike_get_spike_work": (i32, [vp, vp]),
        "spike_batch_get_spike_work": (i32, [vp, vp]),
        "spike_get_schedule": (i32, [vp, ctypes.POINTER(i32), vp, i32]),
        "spike_get_tips": (i32, [vp, vp, vp, vp, vp, vp, vp]),
        "spike_get_reduced_solution": (i32, [vp, vp]),
        "spike_set_stream": (i32, [vp, vp]),
        "spike_error_string": (ctypes.c_char_p, [i32]),
        "spike_last_error_message": (ctypes.c_char_p, []),
        "spike_batch_create": (i32, [ctypes.POINTER(vp), ctypes.POINTER(Opts)]),
        "spike_batch_destroy": (i32, [vp]),
        "spike_batch_solve_dev": (i32, [vp, i32, i32, ctypes.POINTER(System)]),
        "spike_batch_get_stats": (i32, [vp, ctypes.POINTER(Stats), vp]),
        "spike_autotune_partitions": (i32, [i32, i32, i32, i32, i32, i32]),
        "spike_autotune_partitions_tuned": (i32, [i32, i32, i32, i32, i32, i32, i32]),
        "spike_residual_host": (i32, [i32, i32, i64, vp, vp, vp, i32, vp, vp, vp]),
        "spike_band_solve_host": (i32, [i32, i32, i64, vp, vp, vp, i32, vp, i32, vp, i32, vp]),
        "spike_band_solve_dev_noalloc": (i32, [i32, i32, i64, vp, vp, vp, i32, vp, i32, vp, i32, vp, vp]),
        "spike_gen_banded": (i32, [i32, i32, i32, i32, dbl, dbl, dbl, dbl, dbl, ctypes.c_uint64, vp, vp, vp,
                                   ctypes.POINTER(i64)]),
        "spike_gen_dense": (i32, [i32, i64, dbl, dbl, ctypes.c_uint64, vp]),
        "spike_rcm_order": (i32, [i32, vp, vp, vp, ctypes.POINTER(i32), ctypes.POINTER(i32)]),
        "spike_permute_csr": (i32, [i32, i32, vp, vp, vp, vp, vp, vp, vp]),
        "dspike_core_host": (i32, [i32, i32, i32, vp, vp, vp, vp, vp]),
        "dspike_core_host_blocking": (i32, [i32, i32, i32, vp, vp, vp, vp, vp]),
        "sspike_core_host": (i32, [i32, i32, i32, vp, vp, vp, vp, vp]),
        "cspike_core_host": (i32, [i32, i32, i32, vp, vp, vp, vp, vp]),
        "zspike_core_host": (i32, [i32, i32, i32, vp, vp, vp, vp, vp]),
        "directSolver_CreateHandler": (vp, []),
        "directSolver_Configure": (i32, [vp, i32]),
        "directSolver_Factorize": (i32, [vp, i32, i32, vp, vp, vp]),
        "directSolver_SolveForRHS": (i32, [vp, i32, vp, vp]),
        "directSolver_ShowStatistics": (i32, [vp]),
        "directSolver_Finalize": (i32, [vp]),
        "directSolver_Solve": (i32, [i32, i32, i32, vp, vp, vp, vp, vp]),
        "directSolver_Host_Solve": (i32, [i32, i32, i32, vp, vp, vp, vp, vp]),
        "ComputeResidualOfLinearSystem": (i32, [vp, vp, vp, vp, vp, i32, i32]),
        "matrix_CheckCSR": (i32, [i32, ctypes.c_uint64, vp, vp]),
        "matrix_ComputeBandwidth": (i32, [i32, vp, vp, vp, ctypes.POINTER(i32), ctypes.POINTER(i32)]),
        "matrix_CreateEmptyReducedSystem": (vp, [i32, vp, vp, vp]),
        "spike_solve_analysis": (vp, [vp, i32]),
        "spike_solve_analysis_p": (vp, [vp, i32, i32]),
        "schedule_Destroy": (None, [vp]),
        "_spike_malloc": (vp, [i32, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_char_p, i32]),
        "spike_free": (None, [vp]),
        "spike_nullify": (None, [vp]),
        "spike_memory_is_pinned": (i32, [vp]),
        "matrix_AssembleReducedSystem": (i32, [vp, i32, i32, i32, vp, vp, vp, vp]),
        "block_AssembleReducedRHS": (i32, [vp, i32, i32, i32, i32, vp, vp]),
        "matrix_Deallocate": (i32, [vp]),
        "matrix_LoadCSR": (vp, [ctypes.c_char_p]),
        "matrix_LoadCSR_z": (vp, [ctypes.c_char_p]),
        "matrix_LoadCSR_zl": (vp, [ctypes.c_char_p, i32]),
        "matrix_z_Deallocate": (i32, [vp]),
        "matrix_SaveCSR": (i32, [ctypes.c_char_p, vp]),
        "matrix_ExtractMatrix": (vp, [vp, i32, i32, i32, i32, i32]),
        "matrix_AreEqual": (i32, [vp, vp]),
        "spike_server_default_opts": (None, [ctypes.POINTER(ServerOpts)]),
        "spike_server_create": (i32, [ctypes.POINTER(vp), ctypes.POINTER(ServerOpts)]),
        "spike_server_destroy": (i32, [vp]),
        "spike_server_submit": (i32, [vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]),
        "spike_server_wait": (i32, [vp, ctypes.c_uint64, dbl, ctypes.POINTER(RequestResult)]),
        "spike_server_get_stats": (i32, [vp, ctypes.POINTER(ServerStats)]),
        "spike_server_reset_stats": (i32, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def check(code, where):
    if code != 0:
        detail = ""
        try:
            detail = load().spike_last_error_message().decode(errors="replace")
        except Exception:
            pass
        raise SpikeError(code, where, detail)
    return code


def default_opts(**kw):
    o = Opts()
    load().spike_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TypeError("unknown option %r" % k)
        setattr(o, k, v)
    return o
