"""GPU parity of the direct reduced-system solver: the device band LU with partial pivoting
(spike_band_solve_host, directSolver_Host_Solve) and SPIKE's stage 3 on it (reduced_solver
"direct", and the automatic switch when block-Jacobi does not contract), against the oracle's
band LU of R, the reference's fixtures and their scipy (SuperLU) golden solutions.

Tolerances: the band solve and every SPIKE solve meet ||A x - b|| / ||b|| <= 1e-10 per column
(the metric's bar) and agree with the oracle's FP64 solution within 1e-9 relative; reduced
solutions xr agree with the oracle's within 1e-9 relative (R's entries are O(1)).
"""
import numpy as np
import pytest

from conftest import ref_path

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hs():
    import hspike_amd
    hspike_amd.load()
    return hspike_amd


def rel_err(x, ref):
    return np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300)


def pivoting_band(hs, n, ku, kl, seed, complex_=False):
    """Banded, well conditioned and not diagonally dominant: 2 at (i, s(i)) with s reversing
    blocks of min(ku, kl) + 1 rows (a permutation inside the band), zero diagonal elsewhere, and
    band noise with row sums <= 0.8, so singular values are >= 1.2 and the LU needs row
    interchanges across up to min(ku, kl) rows."""
    rng = np.random.default_rng(seed)
    i = np.repeat(np.arange(n), ku + kl + 1)
    j = i + np.tile(np.arange(-kl, ku + 1), n)
    keep = (j >= 0) & (j < n)
    i, j = i[keep], j[keep]
    v = rng.uniform(-1, 1, i.size) * (0.8 / max(ku + kl, 1))
    if complex_:
        v = v * np.exp(2j * np.pi * rng.uniform(0, 1, i.size))
    s = min(ku, kl) + 1
    blk = (i // s) * s
    rev = np.minimum(blk + s - 1, n - 1) - (i - blk)          # reversed position inside the block
    v[i == j] = 0.0
    v[j == rev] = 2.0
    rp = np.zeros(n + 1, dtype=np.int32)
    np.add.at(rp, i + 1, 1)
    return hs.CSRMatrix(n, np.cumsum(rp).astype(np.int32), j.astype(np.int32), v)


@pytest.mark.parametrize("n,ku,kl,nrhs,cplx", [
    (1, 0, 0, 1, False), (7, 3, 3, 2, False), (500, 4, 9, 3, False), (2000, 31, 17, 5, False),
    (1500, 64, 64, 2, False), (3000, 191, 191, 3, False), (800, 12, 5, 2, True), (1200, 40, 40, 3, True),
])
def test_band_solve_matches_oracle(hs, oracle_mod, n, ku, kl, nrhs, cplx):
    A = pivoting_band(hs, n, ku, kl, seed=n + ku, complex_=cplx)
    if n == 1:
        A = hs.CSRMatrix(1, np.array([0, 1], np.int32), np.array([0], np.int32),
                         np.array([2.5 + (0.5j if cplx else 0)]))
    B = hs.gen_dense(n, nrhs, seed=3, complex=cplx)
    X, rel = hs.band_solve(A, B)
    Xo = oracle_mod.direct_solve(oracle_mod.CSR(A.n, A.rowptr, A.colind, A.aij), B)
    assert rel.max() <= 1e-12, rel
    assert rel_err(X, Xo) <= 1e-9
    np.testing.assert_allclose(hs.residual(A, X, B), rel, rtol=1e-6, atol=1e-16)


def test_band_solve_singular_and_errors(hs):
    n = 50
    A = pivoting_band(hs, n, 2, 2, seed=1)
    aij = A.aij.copy()
    col7 = A.colind == 7
    aij[col7] = 0.0                                     # column 7 of the band is zero: singular
    S = hs.CSRMatrix(n, A.rowptr, A.colind, aij)
    with pytest.raises(hs.SpikeError) as ex:
        hs.band_solve(S, np.ones((n, 1)))
    assert ex.value.code == -4
    bad = hs.CSRMatrix(n, A.rowptr, np.where(col7, n + 3, A.colind).astype(np.int32), A.aij)
    with pytest.raises(hs.SpikeError) as ex:
        hs.band_solve(bad, np.ones((n, 1)))
    assert ex.value.code == -2


@pytest.mark.parametrize("k,P,n", [(64, 4, 20000), (16, 8, 6000), (3, 4, 40)])
def test_host_solve_on_reduced_system(hs, oracle_mod, k, P, n):
    """directSolver_Host_Solve (spike_algebra.h:157-164) on the reference's R, assembled by the
    oracle for a C2-shaped matrix (fill 0.15, diagonal 1.05 sum|offdiag|): bandwidths ku + 2 kl - 1
    and 2 ku + kl - 1 (191 at k = 64), solved against the oracle's xr."""
    A = hs.gen_banded(n, k, k, fill=0.15 if k > 8 else 1.0, dom=(1.05, 1.05), diag_floor=1.0, seed=k)
    B = hs.gen_dense(n, 3, seed=4)
    Ao = oracle_mod.CSR(A.n, A.rowptr, A.colind, A.aij)
    _, info = oracle_mod.spike_solve(Ao, B, P)
    R, g = info["R"], info["g"]
    Rh = hs.CSRMatrix(R.n, R.rowptr, R.colind, R.aij)
    assert Rh.bandwidth() == (2 * k + k - 1, k + 2 * k - 1)
    xr = hs.host_solve(Rh, g)
    assert rel_err(xr, info["xr"]) <= 1e-9
    assert hs.residual(Rh, xr, g).max() <= 1e-12


@pytest.mark.parametrize("n,k,P,prec,dom", [
    (20000, 64, 4, "fp64", 1.05), (20000, 64, 4, "mixed", 1.05), (6000, 16, 16, "fp64", 1.1),
    (9000, 32, 3, "mixed", 1.1), (4000, 5, 7, "fp64", 1.0),
])
def test_direct_reduced_solve_matches_oracle(hs, oracle_mod, n, k, P, prec, dom):
    A = hs.gen_banded(n, k, k, fill=0.3, dom=(dom, dom), diag_floor=1.0, seed=n + k)
    B = hs.gen_dense(n, 4, seed=5)
    s = hs.Solver(prec, partitions=P, reduced_solver="direct")
    s.set_matrix(A)
    s.factor()
    X = s.solve(B)
    st = s.stats()
    assert st["P"] == P and st["reduced_direct"] == 1 and st["reduced_iters"] == 0
    Ao = oracle_mod.CSR(A.n, A.rowptr, A.colind, A.aij)
    Xo, info = oracle_mod.spike_solve(Ao, B, P)
    assert rel_err(X, Xo) <= 1e-9
    assert hs.residual(A, X, B).max() <= 1e-10
    if prec == "fp64" and st["refine_iters"] == 0:
        assert rel_err(s.reduced_solution(4), info["xr"]) <= 1e-9
    # the factored R stays resident: a second solve with new right-hand sides
    B2 = hs.gen_dense(n, 2, seed=6)
    X2 = s.solve(B2)
    assert rel_err(X2, oracle_mod.direct_solve(Ao, B2)) <= 1e-9


def test_reference_default_run(hs, oracle_mod, golden, monkeypatch):
    """The reference's own driver run (OpenMP/main.c:86-111): Tests/heptadiagonal/medium.bin, 2 RHS
    of ones, 4 partitions, dspike_core_host_blocking.  The matrix is not diagonally dominant: its
    reduced system is solved by the direct band LU (block-Jacobi does not contract there)."""
    Ao = oracle_mod.load_csr(ref_path("heptadiagonal/medium.bin"))
    A = hs.CSRMatrix(Ao.n, Ao.rowptr, Ao.colind, Ao.aij)
    B = np.ones((A.n, 2))
    monkeypatch.setenv("HSPIKE_PARTITIONS", "4")
    X = hs.dspike_core_host_blocking(A, B)
    for c in range(2):
        assert rel_err(X[:, c], golden["x_hepta40_ones"].ravel()) <= 1e-9
    assert hs.residual(A, X, B).max() <= 1e-10
    Xo, _ = oracle_mod.spike_solve(Ao, B, 4)
    assert rel_err(X, Xo) <= 1e-9
    # the same run through the handle: 4 partitions, and the automatic switch to the direct solver
    s = hs.Solver("fp64", partitions=4)
    s.set_matrix(A)
    s.factor()
    X1 = s.solve(B)
    st = s.stats()
    assert st["P"] == 4 and st["reduced_direct"] == 1
    assert rel_err(X1, Xo) <= 1e-9


@pytest.mark.parametrize("rel,gname,P", [
    ("pentadiagonal/small.bin", "penta10", 2), ("pentadiagonal/small.bin", "penta10", 3),
    ("spike/penta_15.bin", "penta15", 2), ("spike/penta_15.bin", "penta15", 3), ("spike/penta_15.bin", "penta15", 4),
    ("heptadiagonal/small.bin", "hepta20", 2), ("heptadiagonal/small.bin", "hepta20", 3),
    ("heptadiagonal/medium.bin", "hepta40", 2), ("heptadiagonal/medium.bin", "hepta40", 4),
    ("solve_small/tridiagonal.bin", "tridiag10", 4), ("bandwidth/banded.bin", "banded10", 2),
])
@pytest.mark.parametrize("solver", ["auto", "direct"])
def test_reference_fixtures_at_reference_partitions(hs, oracle_mod, golden, rel, gname, P, solver):
    """Every fixture of the reference at 2-4 partitions (the reference's P = 4 and smaller), with a
    random and an all-ones right-hand side, against the scipy SuperLU golden solutions; the
    dense-inverse golden tips where they were recorded at this P."""
    Ao = oracle_mod.load_csr(ref_path(rel))
    A = hs.CSRMatrix(Ao.n, Ao.rowptr, Ao.colind, Ao.aij)
    B = np.column_stack([np.ones(A.n), golden["b_%s_rand" % gname]])
    s = hs.Solver("fp64", partitions=P, reduced_solver=solver)
    s.set_matrix(A)
    s.factor()
    X = s.solve(B)
    assert s.stats()["P"] == P
    assert rel_err(X[:, :1], golden["x_%s_ones" % gname]) <= 1e-9
    assert rel_err(X[:, 1:], golden["x_%s_rand" % gname]) <= 1e-9
    assert hs.residual(A, X, B).max() <= 1e-10
    key = "tips_%s_P%d_Vt" % (gname, P)
    if key in golden.files:
        t = s.tips(2)
        for name in ("Vt", "Vb", "Wt", "Wb"):
            np.testing.assert_allclose(t[name], golden["tips_%s_P%d_%s" % (gname, P, name)], atol=1e-10, err_msg=name)


def test_jacobi_only_reports_not_converged_on_nondominant(hs, oracle_mod):
    """With the direct path switched off, a reduced system whose block-Jacobi sweeps diverge ends in
    SPIKE_ERR_NOT_CONVERGED, never in rc = 0 with a wrong x."""
    Ao = oracle_mod.load_csr(ref_path("heptadiagonal/medium.bin"))
    A = hs.CSRMatrix(Ao.n, Ao.rowptr, Ao.colind, Ao.aij)
    s = hs.Solver("fp64", partitions=4, reduced_solver="jacobi", max_refine=2)
    s.set_matrix(A)
    s.factor()
    B = np.ones((A.n, 1))
    with pytest.raises(hs.SpikeError) as ex:
        s.solve(B)
    assert ex.value.code == -5
