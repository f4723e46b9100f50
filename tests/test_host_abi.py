"""CPU-only checks of the library: it loads, exports every symbol include/*.h declares, and its
host-side logic (CSR IO, validation, schedule, reduced-system pattern, generators, RCM) is right.
No GPU compute is called here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, ref_path


@pytest.fixture(scope="module")
def L():
    import hspike_amd
    return hspike_amd.load()


def declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*[\s\*]+\**([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt, re.M):
            if m.group(1) not in ("if", "while", "return", "sizeof"):
                names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported(L):
    from hspike_amd import _lib
    names = declared_symbols()
    assert len(names) > 50
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(_lib.EXPORTED) <= set(names) | {"spike_last_error_message"}


def test_reference_signatures_kept():
    txt = open(os.path.join(ROOT, "include", "spike_interfaces.h")).read()
    for p in "sdcz":
        assert re.search(r"Error_t %sspike_core_host\(const integer_t n, const integer_t nnz, const integer_t nrhs," % p, txt)
    assert "dspike_core_host_blocking" in txt


def test_load_csr_and_compare(L, tmp_path):
    import hspike_amd as hs
    import oracle
    A = oracle.load_csr(ref_path("split/matrix.bin"))
    M = L.matrix_LoadCSR(ref_path("split/matrix.bin").encode())
    assert M
    out = str(tmp_path / "m.bin")
    assert L.matrix_SaveCSR(out.encode(), M) == 0
    assert open(out, "rb").read() == open(ref_path("split/matrix.bin"), "rb").read()
    # ExtractMatrix == the split fixture blocks (exact, OpenMP/split.c:43-75)
    for (r0, rf), name in (((0, 5), "block1"), ((5, 10), "block2")):
        E = L.matrix_ExtractMatrix(M, r0, rf, r0, rf, 1)
        G = L.matrix_LoadCSR(ref_path("split/%s.bin" % name).encode())
        assert L.matrix_AreEqual(E, G) == 1
        L.matrix_Deallocate(E)
        L.matrix_Deallocate(G)
    L.matrix_Deallocate(M)
    assert A.n == 10
    assert hs.CSRMatrix(A.n, A.rowptr, A.colind, A.aij).bandwidth() == (1, 1)


def test_bandwidth_fixture(L):
    import hspike_amd as hs
    import oracle
    A = oracle.load_csr(ref_path("bandwidth/banded.bin"))
    assert hs.CSRMatrix(A.n, A.rowptr, A.colind, A.aij).bandwidth() == (2, 4)


def test_complex_loader_detects_planar(L):
    class MZ(ctypes.Structure):
        _fields_ = [("n", ctypes.c_int), ("nnz", ctypes.c_uint64), ("ku", ctypes.c_int), ("kl", ctypes.c_int),
                    ("type", ctypes.c_int), ("colind", ctypes.c_void_p), ("rowptr", ctypes.c_void_p),
                    ("aij", ctypes.POINTER(ctypes.c_double)), ("planar_file", ctypes.c_int)]
    import oracle
    for f in ("complex16/penta_15.z", "complex16/penta_1k.z"):
        p = L.matrix_LoadCSR_z(ref_path(f).encode())
        assert p
        M = ctypes.cast(p, ctypes.POINTER(MZ)).contents
        assert M.planar_file == 1
        ref = oracle.load_csr(ref_path(f))
        vals = np.ctypeslib.as_array(M.aij, (2 * M.nnz,)).copy()
        np.testing.assert_array_equal(vals[0::2] + 1j * vals[1::2], ref.aij)
        L.matrix_z_Deallocate(p)


def test_malformed_csr_rejected(L):
    rp = np.array([0, 2, 1, 3], np.int32)         # not monotone
    ci = np.array([0, 1, 2], np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert L.matrix_CheckCSR(3, 3, p(ci), p(rp)) == -2
    rp = np.array([0, 1, 2, 3], np.int32)
    ci = np.array([0, 5, 2], np.int32)            # column out of range
    assert L.matrix_CheckCSR(3, 3, p(ci), p(rp)) == -2
    ci = np.array([0, 1, 2], np.int32)
    assert L.matrix_CheckCSR(3, 3, p(ci), p(rp)) == 0
    # the reference aborts on nnz < 0 (spike_interfaces.c:53-56); here: an error code, no abort
    x = np.zeros(3)
    assert L.dspike_core_host(3, -1, 1, p(ci), p(rp), p(x), p(x), p(x)) == -1
    assert L.dspike_core_host(3, 3, 1, p(np.array([0, 1, 7], np.int32)), p(rp), p(x), p(x), p(x)) == -2


def test_reduced_pattern_matches_oracle(L):
    """matrix_CreateEmptyReducedSystem + AssembleReducedSystem vs the oracle's R (spike_matrix.c:657-740)."""
    import oracle
    rng = np.random.RandomState(3)
    import scipy.sparse as sp
    n, ku, kl, P = 300, 4, 3, 5
    D = np.zeros((n, n))
    for d in range(-kl, ku + 1):
        idx = np.arange(max(0, -d), min(n, n - d))
        D[idx, idx + d] = rng.uniform(-1, 1, idx.size)
    D[np.arange(n), np.arange(n)] += np.abs(D).sum(1) * 1.1
    S = sp.csr_matrix(D)
    A = oracle.CSR(n, S.indptr, S.indices, S.data)
    _, info = oracle.spike_solve(A, np.ones((n, 2)), P)
    kua = np.full(P, ku, np.int32)
    kla = np.full(P, kl, np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    R = L.matrix_CreateEmptyReducedSystem(P, None, p(kua), p(kla))
    cat = {k: np.ascontiguousarray(info[k].transpose(0, 2, 1)).ravel() for k in ("Vt", "Vb", "Wt", "Wb", "yt", "yb")}
    assert L.matrix_AssembleReducedSystem(R, P, ku, kl, p(cat["Vt"]), p(cat["Vb"]), p(cat["Wt"]), p(cat["Wb"])) == 0

    class M(ctypes.Structure):
        _fields_ = [("n", ctypes.c_int), ("nnz", ctypes.c_uint64), ("ku", ctypes.c_int), ("kl", ctypes.c_int),
                    ("type", ctypes.c_int), ("colind", ctypes.POINTER(ctypes.c_int)),
                    ("rowptr", ctypes.POINTER(ctypes.c_int)), ("aij", ctypes.POINTER(ctypes.c_double))]
    m = ctypes.cast(R, ctypes.POINTER(M)).contents
    rp = np.ctypeslib.as_array(m.rowptr, (m.n + 1,))
    ci = np.ctypeslib.as_array(m.colind, (m.nnz,))
    av = np.ctypeslib.as_array(m.aij, (m.nnz,))
    Ro = info["R"]
    np.testing.assert_array_equal(rp, Ro.rowptr)
    np.testing.assert_array_equal(ci, Ro.colind)
    np.testing.assert_allclose(av, Ro.aij, rtol=0, atol=0)
    g = np.zeros(P * (ku + kl) * 2)
    assert L.block_AssembleReducedRHS(p(g), P, ku, kl, 2, p(cat["yt"]), p(cat["yb"])) == 0
    np.testing.assert_array_equal(g.reshape(2, -1).T, info["g"])
    L.matrix_Deallocate(R)


def test_generator_properties():
    import hspike_amd as hs
    A = hs.gen_banded(5000, 7, 5, fill=0.3, dom=(1.01, 2.0), diag_floor=1.0, seed=9)
    assert A.bandwidth() == (7, 5)
    S = A.to_scipy()
    d = np.abs(S.diagonal())
    off = np.asarray(abs(S).sum(1)).ravel() - d
    scaled = d > 1.0                                     # rows not lifted to the diagonal floor
    ratio = d[scaled] / off[scaled]
    assert ratio.min() >= 1.01 - 1e-12 and ratio.max() <= 2.0 + 1e-12
    assert (d[~scaled] == 1.0).all() and (off[~scaled] * 1.01 <= 1.0 + 1e-12).all()  # floor only when needed
    assert (d[off == 0] == 1.0).all()                 # diagonal floor for rows without off-diagonals
    frac = (A.nnz - 5000) / (5000 * 12.0)
    assert 0.27 < frac < 0.33                           # in-band fill
    B = hs.gen_banded(5000, 7, 5, fill=0.3, dom=(1.01, 2.0), diag_floor=1.0, seed=9)
    np.testing.assert_array_equal(A.aij, B.aij)         # deterministic
    Z = hs.gen_banded(300, 3, 3, fill=1.0, dom=(1.2, 1.2), seed=4, complex=True)
    Sz = Z.to_scipy()
    dz = np.abs(Sz.diagonal())
    np.testing.assert_allclose(dz, (np.asarray(abs(Sz).sum(1)).ravel() - dz) * 1.2)
    X = hs.gen_dense(100, 3, lo=-1, hi=1, seed=1)
    assert X.shape == (100, 3) and X.min() >= -1 and X.max() <= 1


def test_rcm_recovers_band():
    import hspike_amd as hs
    A = hs.gen_banded(3000, 4, 4, fill=0.6, dom=(1.1, 1.1), seed=2)
    perm = np.random.RandomState(1).permutation(A.n).astype(np.int32)
    Ap = hs.permute(A, perm)
    assert max(Ap.bandwidth()) > 1000
    order, ku, kl = hs.rcm(Ap)
    Ar = hs.permute(Ap, order)
    assert Ar.bandwidth() == (ku, kl)
    assert max(ku, kl) <= 16
    assert sorted(order.tolist()) == list(range(A.n))
    # P A P^T preserves the matrix: same entries, renumbered
    Ds = Ap.to_scipy().toarray()
    Dr = Ar.to_scipy().toarray()
    np.testing.assert_array_equal(Dr, Ds[np.ix_(order, order)])


def test_autotuner_is_pure():
    import hspike_amd as hs
    a = hs.autotune_partitions(1_000_000, 64, 64, 16)
    assert a == hs.autotune_partitions(1_000_000, 64, 64, 16)
    assert 32 <= a <= 8192
    assert hs.autotune_partitions(1000, 2, 2, 1) >= 1
    assert hs.autotune_partitions(100, 60, 60, 1) == 1
    # throughput tuning: longer partitions, never more of them than the latency tuning
    for n, k in ((200_000, 64), (37_000, 4), (1_000, 16)):
        t = hs.autotune_partitions(n, k, k, 4, tune="throughput")
        assert 1 <= t <= hs.autotune_partitions(n, k, k, 4)


# ---------------------------------------------------------------- drop-in boundary (headers, memory ABI)
REF_TU = r"""
/* the reference's own typedefs and containers, as spike_datatypes.h:29-91 and
 * spike_matrix.h:35-72 declare them under _DATATYPE_D_, ahead of the backend header
 * (spike_algebra.h:30-36 includes the backend after them) */
#define _SPIKE_DATATYPES_H_
#include <stdint.h>
typedef double real_t;
typedef double complex_t;
typedef int integer_t;
typedef int Error_t;
typedef int Bool_t;
typedef double spike_timer_t;
typedef uint64_t uLong_t;
typedef enum { _C_BLOCK_, _DIAG_BLOCK_, _B_BLOCK_ } matrixtype_t;
typedef struct { integer_t n; uLong_t nnz; integer_t ku; integer_t kl; matrixtype_t type;
                 integer_t *colind; integer_t *rowptr; complex_t *aij; } matrix_t;
typedef enum { _V_BLOCK_, _W_BLOCK_, _RHS_BLOCK_ } blocktype_t;
typedef enum { _TOP_SECTION_, _CENTRAL_SECTION_, _BOTTOM_SECTION_, _WHOLE_SECTION_ } blocksection_t;
typedef struct { blocktype_t type; blocksection_t section; integer_t n, m, ku, kl; complex_t *aij; } block_t;
#define _HIP_BACKEND_
#if defined (_HIP_BACKEND_)
    #include "spike_hip.h"
#endif
/* the reference's prototypes of the functions the plugin also provides (spike_algebra.h:122-164) */
Error_t ComputeResidualOfLinearSystem(integer_t *restrict colind, integer_t *restrict rowptr,
                                      complex_t *restrict aij, complex_t *restrict x, complex_t *restrict b,
                                      integer_t n, integer_t nrhs);
Error_t directSolver_Host_Solve(integer_t n, integer_t nnz, integer_t nrhs, integer_t *restrict colind,
                                integer_t *restrict rowptr, complex_t *restrict aij, complex_t *restrict x,
                                complex_t *restrict b);
/* a driver's use of the slot (spike_interfaces.c:113-143) */
int drive(matrix_t *A, block_t *x, block_t *b) {
    DirectSolverHander_t *h = directSolver_CreateHandler();
    directSolver_Configure(h, b->m);
    directSolver_Factorize(h, A->n, (integer_t)A->nnz, A->colind, A->rowptr, A->aij);
    directSolver_SolveForRHS(h, b->m, x->aij, b->aij);
    directSolver_ShowStatistics(h);
    return directSolver_Finalize(h);
}
"""


@pytest.mark.parametrize("std", ["c99", "gnu11"])
def test_plugin_header_in_reference_translation_unit(tmp_path, std):
    """spike_hip.h takes a backend's place after the reference's typedefs (spike_algebra.h:30-36)
    without redeclaring any of them, and its prototypes agree with the reference's."""
    import subprocess
    src = tmp_path / "tu.c"
    src.write_text(REF_TU)
    r = subprocess.run(["gcc", "-std=" + std, "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_public_headers_compile_together(tmp_path):
    import subprocess
    src = tmp_path / "all.c"
    src.write_text('#include "spike_hip.h"\n#include "spike_interfaces.h"\n#include "spike_memory.h"\n'
                   '#include "spike_ext.h"\n#include "spike_matrix.h"\nint main(void){return 0;}\n')
    for std in ("c99", "c11"):
        r = subprocess.run(["gcc", "-std=" + std, "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                            str(src)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-x", "c++", "-I",
                        os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_reference_l1_signatures():
    """matrix_ComputeBandwidth(n, colind, rowptr, aij, ku, kl) (spike_algebra.h:114-119),
    matrix_CreateEmptyReducedSystem(P, n, ku, kl) (spike_matrix.h:242-245),
    spike_solve_analysis(A, nrhs) (spike_analysis.h:65), the memory ABI (spike_memory.h:33-37)."""
    flat = lambda t: re.sub(r"\s+", " ", re.sub(r"/\*.*?\*/", "", t, flags=re.S))
    m = flat(open(os.path.join(ROOT, "include", "spike_matrix.h")).read())
    assert re.search(r"Error_t matrix_ComputeBandwidth\(integer_t n, const integer_t \*colind, const integer_t \*rowptr, "
                     r"const double \*aij, integer_t \*ku, integer_t \*kl\);", m)
    assert re.search(r"matrix_CreateEmptyReducedSystem\(integer_t P, const integer_t \*n, const integer_t \*ku, "
                     r"const integer_t \*kl\);", m)
    assert re.search(r"sm_schedule_t \*spike_solve_analysis\(matrix_t \*A, const integer_t nrhs\);", m)
    mem = flat(open(os.path.join(ROOT, "include", "spike_memory.h")).read())
    assert "void *_spike_malloc(const int alignment, const size_t nmemb, const size_t size, const char *function, const int line);" in mem
    assert "void spike_free(void *ptr);" in mem and "void spike_nullify(void *ptr);" in mem
    assert "extern unsigned int cnt_alloc;" in mem and "extern unsigned int cnt_free;" in mem


def test_bandwidth_and_analysis(L):
    """Bandwidth over the stored entries (the reference fixture Tests/bandwidth: ku = 2, kl = 4) and
    the schedule of spike_solve_analysis: p = 4, n[i] = i * floor(n / 4), the remainder to the last
    partition, r[i] = i (ku + kl) (spike_analysis.c:38-76)."""
    M = L.matrix_LoadCSR(ref_path("bandwidth/banded.bin").encode())
    assert M

    class Mat(ctypes.Structure):
        _fields_ = [("n", ctypes.c_int), ("nnz", ctypes.c_uint64), ("ku", ctypes.c_int), ("kl", ctypes.c_int),
                    ("type", ctypes.c_int), ("colind", ctypes.c_void_p), ("rowptr", ctypes.c_void_p),
                    ("aij", ctypes.c_void_p)]

    class Sched(ctypes.Structure):
        _fields_ = [("p", ctypes.c_int), ("max_n", ctypes.c_int), ("max_m", ctypes.c_int), ("max_nrhs", ctypes.c_int),
                    ("blockingDistance", ctypes.c_int), ("n", ctypes.POINTER(ctypes.c_int)),
                    ("r", ctypes.POINTER(ctypes.c_int)), ("ku", ctypes.POINTER(ctypes.c_int)),
                    ("kl", ctypes.POINTER(ctypes.c_int))]
    m = ctypes.cast(M, ctypes.POINTER(Mat)).contents
    ku, kl = ctypes.c_int(-1), ctypes.c_int(-1)
    assert L.matrix_ComputeBandwidth(m.n, m.colind, m.rowptr, m.aij, ctypes.byref(ku), ctypes.byref(kl)) == 0
    assert (ku.value, kl.value) == (2, 4)
    m.ku = m.kl = 0                                     # spike_solve_analysis recomputes them into A
    old = os.environ.pop("HSPIKE_PARTITIONS", None)
    try:
        S = L.spike_solve_analysis(M, 3)
    finally:
        if old is not None:
            os.environ["HSPIKE_PARTITIONS"] = old
    assert S
    m = ctypes.cast(M, ctypes.POINTER(Mat)).contents
    assert (m.ku, m.kl) == (2, 4)
    s = ctypes.cast(S, ctypes.POINTER(Sched)).contents
    assert s.p == 4 and s.max_nrhs == 4 and s.blockingDistance == 200
    assert [s.n[i] for i in range(5)] == [0, 2, 4, 6, 10]
    assert [s.r[i] for i in range(5)] == [0, 6, 12, 18, 24]
    L.schedule_Destroy(S)
    L.matrix_Deallocate(M)


def test_memory_abi(L):
    """_spike_malloc / spike_free / spike_nullify (spike_memory.h:33-37): aligned, counted, and
    NULL instead of abort() on a request that cannot be met."""
    cnt_a = ctypes.c_uint.in_dll(L, "cnt_alloc")
    cnt_f = ctypes.c_uint.in_dll(L, "cnt_free")
    a0, f0 = cnt_a.value, cnt_f.value
    ptrs = []
    for align, nmemb, size in ((16, 10, 8), (32, 1000, 16), (256, 3, 1), (4096, 1, 100)):
        ptr = L._spike_malloc(align, nmemb, size, b"test", 1)
        assert ptr and ptr % max(align, 64) == 0
        ctypes.memset(ptr, 0x5a, nmemb * size)           # the whole buffer is writable
        ptrs.append(ptr)
    assert cnt_a.value == a0 + 4
    L.spike_free(ptrs[0])
    L.spike_nullify(ptrs[1])
    L.spike_free(ptrs[2])
    L.spike_free(ptrs[3])
    L.spike_free(None)                                    # ignored, still counted (spike_memory_host.c:42-50)
    assert cnt_f.value == f0 + 5
    assert not L._spike_malloc(16, 2 ** 62, 16, b"test", 2)   # overflow: NULL, no abort
    assert cnt_a.value == a0 + 4


def test_prepared_batch_descriptors():
    """Batch.prepare / PreparedBatch: one spike_system_t per system carrying its sizes and the
    tensors' data pointers, the tensors kept referenced (CPU tensors stand in for device ones; no
    library call is made)."""
    import torch
    import hspike_amd as hs
    systems = []
    for i in range(3):
        n, nnz, m = 100 + i, 300 + i, 1 + i
        systems.append(dict(n=n, nrhs=m, rowptr=torch.zeros(n + 1, dtype=torch.int32),
                            colind=torch.zeros(nnz, dtype=torch.int32), aij=torch.zeros(nnz, dtype=torch.float64),
                            b=torch.zeros((m, n), dtype=torch.float64), x=torch.zeros((m, n), dtype=torch.float64)))
    pb = hs.PreparedBatch(systems)
    assert pb.count == 3 and len(pb.systems) == 3
    for a, s in zip(pb.arr, systems):
        assert (a.n, a.nnz, a.nrhs) == (s["n"], s["colind"].numel(), s["nrhs"])
        assert (a.d_rowptr, a.d_colind, a.d_aij, a.d_b, a.d_x) == tuple(
            s[k].data_ptr() for k in ("rowptr", "colind", "aij", "b", "x"))
