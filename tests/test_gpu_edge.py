"""Edge cases of the HIP SPIKE path on the GPU: degenerate shapes, more RHS than one 64-column
chunk, the reference's own fixture matrices, error paths, the single-precision legacy entry
points, bandwidth reordering ahead of the solve, and the backend plugin ABI.

Every test runs the HIP kernels through the C ABI of libhspike.so.  Bars: every column's relative
residual ||A x - b|| / ||b|| <= 1e-10 (the metric's bar; 1e-12 where the test asks the solver for
it) and agreement with the oracle's FP64 direct solve within 1e-9 relative (FP32 entry points:
1e-5, their inputs and outputs being FP32).
"""
import ctypes

import numpy as np
import pytest

from conftest import ref_path

pytestmark = pytest.mark.gpu


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


def solve_and_check(hs, oracle_mod, A, B, prec, P=0, tol=1e-12):
    s = hs.Solver(prec, partitions=P, tol=tol)
    s.set_matrix(A)
    X = s.solve(B)
    st = s.stats()
    s.close()
    X64 = oracle_mod.direct_solve(to_o(oracle_mod, A), B)
    assert rel_err(X, X64) <= 1e-9
    assert st["max_rel_residual"] <= tol
    assert hs.residual(A, X, B).max() <= tol
    return X, st


@pytest.mark.parametrize("prec,nrhs", [("mixed", 70), ("fp64", 70), ("mixed", 66), ("mixed", 129)])
def test_rhs_beyond_one_chunk(hs, oracle_mod, prec, nrhs):
    """RHS are solved in chunks of 64 columns; each chunk's convergence is its own (a chunk after
    the first once reported its unrefined first-pass residual)."""
    A = hs.gen_banded(6000, 8, 8, fill=0.5, dom=(1.1, 1.1), diag_floor=1.0, seed=17)
    B = hs.gen_dense(6000, nrhs, seed=4)
    solve_and_check(hs, oracle_mod, A, B, prec, P=5)


@pytest.mark.parametrize("n,ku,kl,P,prec,nrhs", [
    (1, 0, 0, 1, "fp64", 1),         # one row
    (1, 0, 0, 1, "mixed", 2),
    (15, 2, 2, 1, "fp64", 3),        # fewer rows than one block
    (500, 1, 0, 4, "fp64", 2),       # upper bidiagonal: no left couplings
    (500, 0, 3, 3, "mixed", 2),      # lower banded: no right couplings
    (1000, 0, 0, 2, "mixed", 3),     # diagonal
    (3000, 4, 4, 90, "fp64", 4),     # many short partitions (33 rows each)
    (6000, 128, 128, 3, "mixed", 3), # the widest supported band, b = 128
    (6000, 128, 128, 3, "fp64", 2),
])
def test_degenerate_and_extreme_shapes(hs, oracle_mod, n, ku, kl, P, prec, nrhs):
    A = hs.gen_banded(n, ku, kl, fill=1.0 if ku + kl < 10 else 0.1, dom=(1.1, 1.1), diag_floor=1.0, seed=n + ku)
    B = hs.gen_dense(n, nrhs, seed=6)
    solve_and_check(hs, oracle_mod, A, B, prec, P=P)


@pytest.mark.parametrize("rel", ["heptadiagonal/small.bin", "heptadiagonal/medium.bin", "pentadiagonal/small.bin",
                                 "pentadiagonal/large.bin", "dummy/tridiagonal.bin", "bandwidth/banded.bin"])
def test_reference_fixture_matrices(hs, oracle_mod, rel):
    """The reference's test matrices through dspike_core_host with B = ones and a random column
    (main.c:103-111); the partition count is autotuned here (one partition for these small n),
    the reference's own P = 4 and P = 2-3 are covered in test_gpu_direct.py."""
    Ao = oracle_mod.load_csr(ref_path(rel))
    A = to_h(hs, Ao)
    B = np.ones((A.n, 2))
    B[:, 1] = hs.gen_dense(A.n, 1, seed=2)[:, 0]
    X = hs.dspike_core_host(A, B)
    X64 = oracle_mod.direct_solve(Ao, B)
    assert rel_err(X, X64) <= 1e-9
    assert hs.residual(A, X, B).max() <= 1e-10


def pair_swap_matrix(hs, n, k, complex_=False, seed=0):
    """Zero diagonal, 2 at (i, i xor 1), band noise with row sums <= 0.8: nonsingular (singular
    values >= 1.2) but every 4x4 pivot block of the block inversion needs row interchanges."""
    rng = np.random.default_rng(seed)
    i = np.repeat(np.arange(n), 2 * k + 1)
    j = i + np.tile(np.arange(-k, k + 1), n)
    keep = (j >= 0) & (j < n)
    i, j = i[keep], j[keep]
    v = 0.8 / (2 * k) * rng.uniform(-1, 1, i.size)
    if complex_:
        v = v * np.exp(2j * np.pi * rng.uniform(0, 1, i.size))
    v[j == i] = 0.0
    v[j == (i ^ 1)] = 2.0
    rp = np.zeros(n + 1, dtype=np.int32)
    np.add.at(rp, i + 1, 1)
    return hs.CSRMatrix(n, np.cumsum(rp).astype(np.int32), j.astype(np.int32), v)


@pytest.mark.parametrize("k,prec,cplx", [(8, "fp64", False), (8, "mixed", False), (40, "mixed", False),
                                         (40, "fp64", False), (8, "mixed", True), (24, "fp64", True)])
def test_interchanges_inside_pivot_blocks(hs, oracle_mod, k, prec, cplx):
    """A zero diagonal: the 4x4 pivot blocks are inverted with row interchanges (without them the
    factorization stops at a zero pivot)."""
    A = pair_swap_matrix(hs, 3000, k, cplx, seed=k)
    B = hs.gen_dense(A.n, 3, seed=9, complex=cplx)
    solve_and_check(hs, oracle_mod, A, B, prec, P=3)


def test_singular_pivot_reported(hs):
    n = 200
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    av = np.ones(n)
    av[77] = 0.0
    s = hs.Solver("fp64", partitions=2)
    s.set_matrix(hs.CSRMatrix(n, rp, ci, av))
    with pytest.raises(hs.SpikeError) as ex:
        s.factor()
    assert ex.value.code == -4
    s.close()


def test_bandwidth_above_256_unsupported(hs):
    A = hs.gen_banded(3000, 300, 300, fill=0.05, dom=(1.1, 1.1), diag_floor=1.0, seed=3)
    s = hs.Solver("mixed")
    s.set_matrix(A)
    with pytest.raises(hs.SpikeError) as ex:
        s.factor()
    assert ex.value.code == -6
    s.close()


def test_malformed_device_csr_rejected(hs):
    """A column index out of range in device-resident CSR is caught by the validation kernel."""
    import torch
    A = hs.gen_banded(4000, 4, 4, fill=1.0, dom=(1.1, 1.1), diag_floor=1.0, seed=8)
    ci = A.colind.copy()
    ci[1234] = A.n                  # one past the last column
    dev = torch.device("cuda:0")
    d_rp = torch.from_numpy(A.rowptr).to(dev)
    d_ci = torch.from_numpy(ci).to(dev)
    d_av = torch.from_numpy(A.aij).to(dev)
    s = hs.Solver("fp64")
    s.set_matrix_dev(A.n, d_rp, d_ci, d_av)
    with pytest.raises(hs.SpikeError) as ex:
        s.factor()
    assert ex.value.code == -2
    s.close()


def test_single_precision_legacy_abi(hs, oracle_mod):
    A = hs.gen_banded(5000, 6, 6, fill=1.0, dom=(1.1, 1.1), diag_floor=1.0, seed=12)
    B = hs.gen_dense(5000, 3, seed=13)
    A32 = hs.CSRMatrix(A.n, A.rowptr, A.colind, A.aij.astype(np.float32).astype(np.float64))
    X = hs.sspike_core_host(A, B)
    assert X.dtype == np.float32
    X64 = oracle_mod.direct_solve(to_o(oracle_mod, A32), B.astype(np.float32).astype(np.float64))
    assert rel_err(X.astype(np.float64), X64) <= 1e-5
    Bz = hs.gen_dense(5000, 2, seed=14, complex=True)
    Az = hs.gen_banded(5000, 6, 6, fill=1.0, dom=(1.2, 1.2), diag_floor=1.0, seed=15, complex=True)
    Xz = hs.cspike_core_host(Az, Bz)
    assert Xz.dtype == np.complex64
    Az32 = hs.CSRMatrix(Az.n, Az.rowptr, Az.colind, Az.aij.astype(np.complex64).astype(np.complex128))
    Xz64 = oracle_mod.direct_solve(to_o(oracle_mod, Az32), Bz.astype(np.complex64).astype(np.complex128))
    assert rel_err(Xz.astype(np.complex128), Xz64) <= 1e-5


def test_rcm_reordering_then_solve(hs, oracle_mod):
    """A banded matrix with its rows and columns shuffled: RCM restores a narrow band, the GPU
    solves P A P^T y = P b, and x = P^T y solves the shuffled system."""
    A0 = hs.gen_banded(3000, 5, 5, fill=1.0, dom=(1.1, 1.1), diag_floor=1.0, seed=21)
    shuffle = np.random.default_rng(5).permutation(A0.n).astype(np.int32)
    A = hs.permute(A0, shuffle)
    assert A.bandwidth()[0] > 1000
    perm, ku, kl = hs.rcm(A)
    assert max(ku, kl) <= 32
    Ar = hs.permute(A, perm)
    B = hs.gen_dense(A.n, 3, seed=22)
    s = hs.Solver("mixed", tol=1e-12)
    s.set_matrix(Ar)
    Y = s.solve(B[perm])
    s.close()
    X = np.empty_like(Y)
    X[perm] = Y
    assert hs.residual(A, X, B).max() <= 1e-12
    assert rel_err(X, oracle_mod.direct_solve(to_o(oracle_mod, A), B)) <= 1e-9


def test_backend_plugin_abi(hs, oracle_mod):
    """directSolver_* (spike_hip.h): factor once, solve two RHS panels, one-shot Solve."""
    L = hs.load()
    A = hs.gen_banded(8000, 20, 12, fill=0.4, dom=(1.1, 1.1), diag_floor=1.0, seed=31)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    h = L.directSolver_CreateHandler()
    assert h
    assert L.directSolver_Configure(h, 8) == 0
    assert L.directSolver_Factorize(h, A.n, A.nnz, p(A.colind), p(A.rowptr), p(A.aij)) == 0
    for seed, m in ((41, 3), (42, 8)):
        B = np.asfortranarray(hs.gen_dense(A.n, m, seed=seed))
        X = np.zeros_like(B, order="F")
        assert L.directSolver_SolveForRHS(h, m, p(X), p(B)) == 0
        assert rel_err(X, oracle_mod.direct_solve(to_o(oracle_mod, A), B)) <= 1e-9
    assert L.directSolver_Finalize(h) == 0
    B = np.asfortranarray(hs.gen_dense(A.n, 2, seed=43))
    X = np.zeros_like(B, order="F")
    assert L.directSolver_Solve(A.n, A.nnz, 2, p(A.colind), p(A.rowptr), p(A.aij), p(X), p(B)) == 0
    assert hs.residual(A, X, B).max() <= 1e-10


def cross_panel_matrix(hs, n, k, shift, seed=0):
    """Zero diagonal, 2 at (i, i xor shift), band noise with row sums <= 0.8: nonsingular (singular
    values >= 1.2), but with shift = 4 every 4x4 pivot block of the block inversion holds only the
    noise (its pivots lie in the neighbouring panel), and with shift = 16 every diagonal 16 x 16
    block does (its pivots lie in the neighbouring block)."""
    rng = np.random.default_rng(seed)
    i = np.repeat(np.arange(n), 2 * k + 1)
    j = i + np.tile(np.arange(-k, k + 1), n)
    keep = (j >= 0) & (j < n)
    i, j = i[keep], j[keep]
    v = 0.8 / (2 * k) * rng.uniform(-1, 1, i.size)
    v[j == i] = 0.0
    v[j == (i ^ shift)] = 2.0
    rp = np.zeros(n + 1, dtype=np.int32)
    np.add.at(rp, i + 1, 1)
    return hs.CSRMatrix(n, np.cumsum(rp).astype(np.int32), j.astype(np.int32), v)


@pytest.mark.parametrize("shift,k,prec", [(4, 8, "fp64"), (4, 8, "mixed"), (4, 20, "fp64"), (16, 24, "fp64"),
                                          (16, 24, "mixed")])
def test_pivots_outside_the_pivot_panels(hs, oracle_mod, shift, k, prec):
    """The block LU interchanges rows only inside each 4x4 pivot block (no pivoting across panels or
    blocks, unlike the reference's SuperLU with threshold 1.0, spike_superlu.c:76).  A nonsingular
    matrix whose pivots lie outside their panel must either come back right (the FP64 residual and
    refinement vouch for it) or fail cleanly with SPIKE_ERR_SINGULAR / SPIKE_ERR_NOT_CONVERGED --
    never rc = 0 with a wrong x."""
    A = cross_panel_matrix(hs, 4000, k, shift, seed=shift + k)
    B = hs.gen_dense(A.n, 2, seed=3)
    s = hs.Solver(prec, partitions=4, max_refine=4)
    s.set_matrix(A)
    try:
        s.factor()
        X = s.solve(B)
    except hs.SpikeError as ex:
        assert ex.code in (-4, -5), ex
        return
    assert hs.residual(A, X, B).max() <= 1e-10
    Xo = oracle_mod.direct_solve(oracle_mod.CSR(A.n, A.rowptr, A.colind, A.aij), B)
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) <= 1e-9


def test_host_panels_with_leading_dimensions(hs):
    """spike_solve_host and spike_band_solve_host with leading dimensions above n (ldb = n + 7,
    ldx = n + 3) on panels of 43 MB, wider than one 32 MiB pinned staging chunk, so chunk bounds
    fall inside columns: the host panels are packed into and out of the staging buffers column by
    column.  The solutions equal those of the ld = n calls bit for bit, and X's padding rows are
    left as they were."""
    L = hs.load()
    vp = ctypes.c_void_p
    p = lambda a: a.ctypes.data_as(vp)
    A = hs.gen_banded(600000, 8, 8, fill=0.5, dom=(1.1, 1.1), diag_floor=1.0, seed=61)
    n, m = A.n, 9
    B = np.asfortranarray(hs.gen_dense(n, m, seed=62))
    ldb, ldx = n + 7, n + 3
    Bp = np.zeros((ldb, m), order="F")
    Bp[:n] = B
    s = hs.Solver("mixed", tol=1e-10)
    s.set_matrix(A)
    X0 = s.solve(B)
    Xp = np.full((ldx, m), 7.0, order="F")
    assert L.spike_solve_host(s._h, m, p(Bp), ldb, p(Xp), ldx) == 0
    s.close()
    np.testing.assert_array_equal(Xp[:n], X0)
    assert np.all(Xp[n:] == 7.0)
    assert hs.residual(A, X0, B).max() <= 1e-10
    X1, _ = hs.band_solve(A, B)
    rel = np.zeros(m)
    Xq = np.full((ldx, m), 7.0, order="F")
    assert L.spike_band_solve_host(hs.SPIKE_DTYPE_D, n, A.nnz, p(A.rowptr), p(A.colind), p(A.aij), m, p(Bp), ldb,
                                   p(Xq), ldx, p(rel)) == 0
    np.testing.assert_array_equal(Xq[:n], X1)
    assert np.all(Xq[n:] == 7.0)
