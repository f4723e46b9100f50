"""CPU oracle for the SPIKE banded-solve path -- ctypes wrapper of liboracle.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by
bench.py's ``cpu_baseline`` leg, always as the checker or the timed CPU baseline, never as
the product path.  The product (``hspike_amd``) never imports this module.

What it restates (srodrb/HSPIKE):
  * ``load_csr``        -- the binary CSR format of OpenMP/spike_matrix.c:11-51
                           (int32 n, nnz, dtype; aij[nnz]; colind[nnz]; rowptr[n+1]); complex
                           files written by generator_banded.py:134-141 are planar (all real
                           parts, then all imaginary parts) and are read as such.
  * ``bandwidth``       -- OpenMP/spike_algebra.c:56-95
  * ``schedule``        -- OpenMP/spike_analysis.c:38-76 (partition rows, reduced offsets)
  * ``extract_csr``     -- OpenMP/spike_matrix.c:563-645
  * ``spike_solve``     -- OpenMP/spike_interfaces.c:43-419 (all four stages, with tips,
                           reduced system R / g / xr); see spike_oracle_body.h
  * ``direct_solve``    -- banded partial-pivot LU of the whole matrix
  * ``residual``        -- OpenMP/spike_algebra.c:97-151 (every column)
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build(force=False):
    """Compile liboracle.so with the Makefile beside this file."""
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "all"])
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        _LIB = ctypes.CDLL(so)
        vp = ctypes.c_void_p
        for suf in ("d", "z"):
            f = getattr(_LIB, "or_spike_solve_" + suf)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_int, ctypes.c_int, vp, vp] + [vp] * 11 + [ctypes.c_int]
            f = getattr(_LIB, "or_band_direct_solve_" + suf)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_int, vp, vp]
            f = getattr(_LIB, "or_residual_" + suf)
            f.restype = None
            f.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_int, vp, vp, vp]
        _LIB.or_bandwidth.argtypes = [ctypes.c_int, vp, vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        _LIB.or_schedule.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp]
        _LIB.or_reduced_nnz.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _LIB.or_reduced_nnz.restype = ctypes.c_int
        _LIB.or_extract_csr.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp]
        _LIB.or_extract_csr.restype = ctypes.c_int
        _LIB.or_num_threads.restype = ctypes.c_int
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class CSR:
    """Square CSR matrix with int32 0-based indices (OpenMP/spike_matrix.h:39-53)."""

    def __init__(self, n, rowptr, colind, aij, planar=False):
        self.n = int(n)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.colind = np.ascontiguousarray(colind, dtype=np.int32)
        self.aij = np.ascontiguousarray(aij)
        self.planar = planar

    @property
    def nnz(self):
        return int(self.colind.size)

    @property
    def is_complex(self):
        return np.iscomplexobj(self.aij)

    def to_dense(self):
        A = np.zeros((self.n, self.n), dtype=self.aij.dtype)
        for r in range(self.n):
            for idx in range(self.rowptr[r], self.rowptr[r + 1]):
                A[r, self.colind[idx]] += self.aij[idx]
        return A

    def matvec(self, X):
        X = np.asarray(X)
        rows = np.repeat(np.arange(self.n), np.diff(self.rowptr))
        prod = self.aij[:, None] * X.reshape(self.n, -1)[self.colind]
        out = np.zeros((self.n, prod.shape[1]), dtype=np.result_type(self.aij, X))
        np.add.at(out, rows, prod)
        return out.reshape(X.shape)


def load_csr(path):
    """Read the reference's binary CSR file (OpenMP/spike_matrix.c:11-51).

    dtype 1 (complex) files as written by generator_banded.py:134-141 hold all real parts and
    then all imaginary parts (planar).  The reference's loader reads them as interleaved
    complex_t (spike_matrix.c:34-35) and so misreads them; this reader follows the writer.
    """
    raw = np.fromfile(path, dtype=np.uint8)
    n, nnz, dtype = np.frombuffer(raw[:12].tobytes(), dtype=np.int32)
    off = 12
    if dtype == 1:
        re = np.frombuffer(raw[off:off + 8 * nnz].tobytes(), dtype=np.float64); off += 8 * nnz
        im = np.frombuffer(raw[off:off + 8 * nnz].tobytes(), dtype=np.float64); off += 8 * nnz
        aij = re + 1j * im
    else:
        aij = np.frombuffer(raw[off:off + 8 * nnz].tobytes(), dtype=np.float64).copy(); off += 8 * nnz
    colind = np.frombuffer(raw[off:off + 4 * nnz].tobytes(), dtype=np.int32).copy(); off += 4 * nnz
    rowptr = np.frombuffer(raw[off:off + 4 * (n + 1)].tobytes(), dtype=np.int32).copy(); off += 4 * (n + 1)
    if off != raw.size:
        raise ValueError("%s: %d trailing bytes" % (path, raw.size - off))
    return CSR(n, rowptr, colind, aij, planar=(dtype == 1))


def bandwidth(A):
    ku, kl = ctypes.c_int(), ctypes.c_int()
    lib().or_bandwidth(A.n, _p(A.rowptr), _p(A.colind), ctypes.byref(ku), ctypes.byref(kl))
    return ku.value, kl.value


def schedule(n, P, ku, kl):
    rows = np.zeros(P + 1, dtype=np.int32)
    red = np.zeros(P + 1, dtype=np.int32)
    lib().or_schedule(n, P, ku, kl, _p(rows), _p(red))
    return rows, red


def extract_csr(A, r0, rf, c0, cf):
    nnz = lib().or_extract_csr(_p(A.rowptr), _p(A.colind), _p(A.aij), r0, rf, c0, cf, None, None, None)
    rp = np.zeros(rf - r0 + 1, dtype=np.int32)
    ci = np.zeros(max(nnz, 1), dtype=np.int32)
    av = np.zeros(max(nnz, 1), dtype=np.float64)
    lib().or_extract_csr(_p(A.rowptr), _p(A.colind), _p(A.aij), r0, rf, c0, cf, _p(rp), _p(ci), _p(av))
    return CSR(rf - r0, rp, ci[:nnz], av[:nnz])


def spike_solve(A, B, P, nthreads=0, want_reduced=True):
    """Run the oracle's explicit SPIKE solve; returns (X, info dict) with tips and R/g/xr."""
    cplx = A.is_complex or np.iscomplexobj(B)
    dt = np.complex128 if cplx else np.float64
    B = np.asfortranarray(np.asarray(B, dtype=dt).reshape(A.n, -1))
    nrhs = B.shape[1]
    aij = np.ascontiguousarray(A.aij, dtype=dt)
    ku, kl = bandwidth(A)
    X = np.zeros((A.n, nrhs), dtype=dt, order="F")
    k2 = ku + kl
    Nr = P * k2
    tips = {
        "Vt": np.zeros(P * ku * ku + 1, dt), "Vb": np.zeros(P * kl * ku + 1, dt),
        "Wt": np.zeros(P * ku * kl + 1, dt), "Wb": np.zeros(P * kl * kl + 1, dt),
        "yt": np.zeros(P * ku * nrhs + 1, dt), "yb": np.zeros(P * kl * nrhs + 1, dt),
    }
    nnzR = lib().or_reduced_nnz(P, ku, kl)
    Rp = np.zeros(Nr + 1, np.int32)
    Rc = np.zeros(max(nnzR, 1), np.int32)
    Ra = np.zeros(max(nnzR, 1), dt)
    g = np.zeros(Nr * nrhs + 1, dt)
    xr = np.zeros(Nr * nrhs + 1, dt)
    fn = lib().or_spike_solve_z if cplx else lib().or_spike_solve_d
    st = fn(A.n, _p(A.rowptr), _p(A.colind), _p(aij), P, nrhs, _p(B), _p(X),
            _p(tips["Vt"]), _p(tips["Vb"]), _p(tips["Wt"]), _p(tips["Wb"]), _p(tips["yt"]), _p(tips["yb"]),
            _p(Rp) if want_reduced else None, _p(Rc) if want_reduced else None,
            _p(Ra) if want_reduced else None, _p(g) if want_reduced else None,
            _p(xr) if want_reduced else None, int(nthreads))
    if st != 0:
        raise RuntimeError("oracle spike solve failed with status %d" % st)
    info = {"ku": ku, "kl": kl, "P": P, "Nr": Nr}
    shapes = {"Vt": (ku, ku), "Vb": (kl, ku), "Wt": (ku, kl), "Wb": (kl, kl), "yt": (ku, nrhs), "yb": (kl, nrhs)}
    for k, (r, c) in shapes.items():
        info[k] = tips[k][:P * r * c].reshape(P, c, r).transpose(0, 2, 1)   # [p] -> r x c
    if want_reduced:
        info["R"] = CSR(Nr, Rp, Rc[:nnzR], Ra[:nnzR])
        info["g"] = g[:Nr * nrhs].reshape(nrhs, Nr).T
        info["xr"] = xr[:Nr * nrhs].reshape(nrhs, Nr).T
    return X, info


def direct_solve(A, B):
    cplx = A.is_complex or np.iscomplexobj(B)
    dt = np.complex128 if cplx else np.float64
    B = np.asfortranarray(np.asarray(B, dtype=dt).reshape(A.n, -1))
    X = np.zeros_like(B, order="F")
    aij = np.ascontiguousarray(A.aij, dtype=dt)
    fn = lib().or_band_direct_solve_z if cplx else lib().or_band_direct_solve_d
    st = fn(A.n, _p(A.rowptr), _p(A.colind), _p(aij), B.shape[1], _p(B), _p(X))
    if st != 0:
        raise RuntimeError("oracle direct solve failed with status %d" % st)
    return X


def residual(A, X, B):
    cplx = A.is_complex or np.iscomplexobj(B) or np.iscomplexobj(X)
    dt = np.complex128 if cplx else np.float64
    X = np.asfortranarray(np.asarray(X, dtype=dt).reshape(A.n, -1))
    B = np.asfortranarray(np.asarray(B, dtype=dt).reshape(A.n, -1))
    rel = np.zeros(X.shape[1], np.float64)
    aij = np.ascontiguousarray(A.aij, dtype=dt)
    fn = lib().or_residual_z if cplx else lib().or_residual_d
    fn(A.n, _p(A.rowptr), _p(A.colind), _p(aij), X.shape[1], _p(X), _p(B), _p(rel))
    return rel


def num_threads():
    return lib().or_num_threads()
