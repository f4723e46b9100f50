"""GPU parity of the HIP SPIKE path against the CPU oracle (oracle/), the reference's fixtures
and their golden values.  Every test runs the HIP kernels through the C ABI of libhspike.so.

Tolerances: FP64 path -- spike tips and reduced solution within 1e-10 (absolute, entries are
O(1)), solution within 1e-9 relative to the oracle and every column's relative residual
||A x - b|| / ||b|| <= 1e-10 (the metric's bar).  Mixed path -- the same residual bar after FP64
refinement and ||x - x64|| / ||x64|| <= 1e-8 against the oracle's FP64 solution.
"""
import numpy as np
import pytest

from conftest import ref_path

pytestmark = pytest.mark.gpu

PENTA10_R = np.array([
    [1.0, 0.0, 0.0, 0.0, 0.00458, -0.00108, 0.0, 0.0],
    [0.0, 1.0, 0.0, 0.0, 0.02556, 0.01347, 0.0, 0.0],
    [0.0, 0.0, 1.0, 0.0, 0.44477, -0.006, 0.0, 0.0],
    [0.0, 0.0, 0.0, 1.0, 0.44704, 0.15165, 0.0, 0.0],
    [0.0, 0.0, 0.30871, 0.23544, 1.0, 0.0, 0.0, 0.0],
    [0.0, 0.0, -0.10528, 0.02555, 0.0, 1.0, 0.0, 0.0],
    [0.0, 0.0, 0.16531, 0.10707, 0.0, 0.0, 1.0, 0.0],
    [0.0, 0.0, -0.09041, -0.05356, 0.0, 0.0, 0.0, 1.0]])
PENTA10_G = np.array([0.32497, 0.4458, 0.25393, 0.32481, 0.4902, 0.40042, 0.21562, 0.65147])
SOLVE_SMALL_X = np.array([0.358, 0.409, 0.543, 0.295, 0.261, 0.439, 0.408, 0.644, -0.012, 0.801])


@pytest.fixture(scope="module")
def hs():
    import hspike_amd
    hspike_amd.load()
    return hspike_amd


def to_h(hs, A):
    return hs.CSRMatrix(A.n, A.rowptr, A.colind, A.aij)


def to_o(oracle_mod, A):
    return oracle_mod.CSR(A.n, A.rowptr, A.colind, A.aij)


def rel_err(x, ref):
    return np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300)


def test_penta10_reduced_system_matches_golden(hs, oracle_mod):
    A = oracle_mod.load_csr(ref_path("spike/penta_10.bin"))
    b = np.ones((A.n, 1))
    s = hs.Solver("fp64", partitions=2)
    s.set_matrix(to_h(hs, A))
    s.factor()
    X = s.solve(b)
    st = s.stats()
    assert st["P"] == 2 and st["ku"] == 2 and st["kl"] == 2
    t = s.tips(1)
    Xo, info = oracle_mod.spike_solve(A, b, 2)
    for k in ("Vt", "Vb", "Wt", "Wb", "yt", "yb"):
        np.testing.assert_allclose(t[k], info[k], atol=1e-12, err_msg=k)
    # reference-pattern R and g assembled from the GPU tips == Tests/spike/penta_10.readme
    import ctypes
    L = hs.load()
    ku_a = np.array([2, 2], np.int32)
    Rp = L.matrix_CreateEmptyReducedSystem(2, None, ku_a.ctypes.data_as(ctypes.c_void_p), ku_a.ctypes.data_as(ctypes.c_void_p))
    assert Rp

    class M(ctypes.Structure):
        _fields_ = [("n", ctypes.c_int), ("nnz", ctypes.c_uint64), ("ku", ctypes.c_int), ("kl", ctypes.c_int),
                    ("type", ctypes.c_int), ("colind", ctypes.POINTER(ctypes.c_int)),
                    ("rowptr", ctypes.POINTER(ctypes.c_int)), ("aij", ctypes.POINTER(ctypes.c_double))]
    cat = {k: np.ascontiguousarray(t[k].transpose(0, 2, 1)).ravel() for k in t}
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert L.matrix_AssembleReducedSystem(Rp, 2, 2, 2, p(cat["Vt"]), p(cat["Vb"]), p(cat["Wt"]), p(cat["Wb"])) == 0
    R = ctypes.cast(Rp, ctypes.POINTER(M)).contents
    n, nnz = R.n, R.nnz
    rp = np.ctypeslib.as_array(R.rowptr, (n + 1,)).copy()
    ci = np.ctypeslib.as_array(R.colind, (nnz,)).copy()
    av = np.ctypeslib.as_array(R.aij, (nnz,)).copy()
    L.matrix_Deallocate(Rp)
    dense = np.zeros((n, n))
    for r in range(n):
        dense[r, ci[rp[r]:rp[r + 1]]] = av[rp[r]:rp[r + 1]]
    np.testing.assert_allclose(dense, PENTA10_R, atol=6e-6)
    g = np.zeros(8)
    assert L.block_AssembleReducedRHS(p(g), 2, 2, 2, 1, p(cat["yt"]), p(cat["yb"])) == 0
    np.testing.assert_allclose(g, PENTA10_G, atol=6e-6)
    np.testing.assert_allclose(s.reduced_solution(1), info["xr"], atol=1e-12)
    np.testing.assert_allclose(X, Xo, rtol=1e-12, atol=1e-13)
    assert st["max_rel_residual"] <= 1e-10


def test_solve_small_golden(hs, oracle_mod):
    A = oracle_mod.load_csr(ref_path("solve_small/tridiagonal.bin"))
    for P in (1, 2, 4):
        s = hs.Solver("fp64", partitions=P)
        s.set_matrix(to_h(hs, A))
        X = s.solve(np.ones((A.n, 1)))
        np.testing.assert_allclose(X[:, 0], SOLVE_SMALL_X, atol=5.1e-4)
        assert hs.residual(to_h(hs, A), X, np.ones((A.n, 1)))[0] <= 1e-14


@pytest.mark.parametrize("n,ku,kl,fill,P,nrhs", [
    (2000, 2, 2, 1.0, 4, 1),      # C1 shape (pentadiagonal, reference P=4), smaller n
    (3001, 5, 3, 0.5, 7, 3),      # ku != kl, uneven last partition
    (5000, 16, 16, 0.3, 6, 16),   # b = 16 exactly
    (6000, 30, 21, 0.3, 5, 20),   # b = 32
    (20000, 64, 64, 0.15, 8, 16), # C2 shape at a reduced size
    (12000, 100, 90, 0.2, 4, 5),  # b = 128
])
def test_fp64_parity_random(hs, oracle_mod, n, ku, kl, fill, P, nrhs):
    A = hs.gen_banded(n, ku, kl, fill=fill, dom=(1.05, 1.05), diag_floor=1.0, seed=n + ku)
    B = hs.gen_dense(n, nrhs, seed=3)
    s = hs.Solver("fp64", partitions=P)
    s.set_matrix(A)
    X = s.solve(B)
    Xo, info = oracle_mod.spike_solve(to_o(oracle_mod, A), B, P)
    t = s.tips(nrhs)
    for k in ("Vt", "Vb", "Wt", "Wb", "yt", "yb"):
        np.testing.assert_allclose(t[k], info[k], atol=1e-10, err_msg=k)
    np.testing.assert_allclose(s.reduced_solution(nrhs), info["xr"], atol=1e-10)
    assert rel_err(X, Xo) <= 1e-9
    rel = hs.residual(A, X, B)
    assert rel.max() <= 1e-10, rel.max()
    # residual kernel vs the oracle's on an O(1) residual (a perturbed solution)
    Xp = X + hs.gen_dense(n, nrhs, seed=9) * 1e-3
    np.testing.assert_allclose(hs.residual(A, Xp, B), oracle_mod.residual(to_o(oracle_mod, A), Xp, B), rtol=1e-10)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_single_pass_tail_partition_and_determinism(hs, oracle_mod, prec):
    """64 partitions of 769 rows and a last one of 829 (the C2 layout at 1/20 size): the last
    partition's workgroup runs alone at the end of every sweep.  FP64 must reach the bar in one
    pass (no refinement) with every partition's tips equal to the oracle's, and two solves must
    agree bit for bit (fixed-order arithmetic, no races)."""
    n, k, P, nrhs = 64 * 769 + 60, 64, 64, 16
    A = hs.gen_banded(n, k, k, fill=0.15, dom=(1.05, 1.05), diag_floor=1.0, seed=1000 + n + k)
    B = hs.gen_dense(n, nrhs, seed=7)
    runs = []
    for _ in range(2):
        s = hs.Solver(prec, partitions=P, max_refine=0 if prec == "fp64" else 8, tol=1e-12)
        s.set_matrix(A)
        X = s.solve(B)
        runs.append((X, s.tips(nrhs), s.stats()))
        s.close()
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    for key in ("Vt", "Vb", "Wt", "Wb", "yt", "yb"):
        np.testing.assert_array_equal(runs[0][1][key], runs[1][1][key], err_msg=key)
    if prec == "fp64":
        assert runs[0][2]["refine_iters"] == 0 and runs[0][2]["max_rel_residual"] <= 1e-14
        Xo, info = oracle_mod.spike_solve(to_o(oracle_mod, A), B, P, want_reduced=False)
        for key in ("Vt", "Vb", "Wt", "Wb", "yt", "yb"):
            np.testing.assert_allclose(runs[0][1][key], info[key], atol=1e-12, err_msg=key)
        assert rel_err(runs[0][0], Xo) <= 1e-12


@pytest.mark.parametrize("n,k,fill,dom,nrhs", [
    (20000, 64, 0.15, (1.05, 1.05), 16),
    (30000, 128, 0.2, (1.01, 2.0), 8),    # C3 shape at a reduced size
    (8000, 8, 0.3, (1.1, 1.1), 3),
])
def test_mixed_parity_autotuned(hs, oracle_mod, n, k, fill, dom, nrhs):
    A = hs.gen_banded(n, k, k, fill=fill, dom=dom, diag_floor=1.0, seed=11 + k)
    B = hs.gen_dense(n, nrhs, seed=5)
    s = hs.Solver("mixed", tol=1e-12)
    s.set_matrix(A)
    X = s.solve(B)
    st = s.stats()
    X64 = oracle_mod.direct_solve(to_o(oracle_mod, A), B)
    assert rel_err(X, X64) <= 1e-8
    assert st["max_rel_residual"] <= 1e-12
    assert hs.residual(A, X, B).max() <= 1e-12


def test_legacy_abi_dspike(hs, oracle_mod):
    A = hs.gen_banded(4000, 3, 3, fill=1.0, dom=(1.0, 1.0), diag_add=1.0, seed=1)
    B = hs.gen_dense(4000, 4, seed=2)
    X = hs.dspike_core_host(A, B)
    Xo = oracle_mod.direct_solve(to_o(oracle_mod, A), B)
    assert rel_err(X, Xo) <= 1e-12
    # every column written (the reference copies column 0 into every column)
    assert not np.allclose(X[:, 0], X[:, 1])
    Xb = hs.dspike_core_host_blocking(A, B)
    np.testing.assert_allclose(Xb, X, rtol=1e-13, atol=1e-15)


def test_complex128_planar_fixture(hs, oracle_mod, golden):
    A = oracle_mod.load_csr(ref_path("complex16/penta_1k.z"))
    b = golden["b_zpenta1k"]
    X = hs.zspike_core_host(to_h(hs, A), b)
    np.testing.assert_allclose(X, golden["x_zpenta1k"], rtol=1e-9, atol=1e-12)
