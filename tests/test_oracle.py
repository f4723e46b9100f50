"""Pins of the CPU oracle (CPU only).

Each oracle function is checked against something it does not compute itself: the
reference's fixtures and golden readmes, scipy/LAPACK golden vectors, dense brute force on
the fixture sizes, or a closed form.  A mistake such as a transposed tip, an off-by-one
partition bound or a dropped coupling term fails at least one of these.
"""
import hashlib

import numpy as np
import pytest

from conftest import ref_path

# Tests/spike/penta_10.readme: reduced system and RHS for P = 2, b = ones (5 decimals)
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
# Tests/solve_small/readme.txt: x for tridiagonal.bin, b = ones (3 decimals)
SOLVE_SMALL_X = np.array([0.358, 0.409, 0.543, 0.295, 0.261, 0.439, 0.408, 0.644, -0.012, 0.801])


def test_load_headers(oracle_mod):
    # sizes stated in the reference readmes
    A = oracle_mod.load_csr(ref_path("pentadiagonal/large.bin"))
    assert (A.n, A.nnz) == (20000, 99994)       # Tests/pentadiagonal/large.readme
    A = oracle_mod.load_csr(ref_path("solve_small/tridiagonal.bin"))
    assert (A.n, A.nnz) == (10, 28)              # Tests/solve_small/readme.txt
    A = oracle_mod.load_csr(ref_path("bandwidth/banded.bin"))
    assert (A.n, A.nnz) == (10, 31)              # Tests/bandwidth/readme.txt


def test_bandwidth_fixture(oracle_mod):
    # Tests/bandwidth/readme.txt: upper 2, lower 4 (also OpenMP/bandwidth.c:32-42)
    A = oracle_mod.load_csr(ref_path("bandwidth/banded.bin"))
    assert oracle_mod.bandwidth(A) == (2, 4)


def test_extract_matches_split_fixture(oracle_mod):
    # OpenMP/split.c:43-75: ExtractMatrix(A,0,5,0,5) == block1, (A,5,10,5,10) == block2, exact
    A = oracle_mod.load_csr(ref_path("split/matrix.bin"))
    for (r0, rf), name in (((0, 5), "block1"), ((5, 10), "block2")):
        E = oracle_mod.extract_csr(A, r0, rf, r0, rf)
        G = oracle_mod.load_csr(ref_path("split/%s.bin" % name))
        assert E.n == G.n
        np.testing.assert_array_equal(E.rowptr, G.rowptr)
        np.testing.assert_array_equal(E.colind, G.colind)
        np.testing.assert_array_equal(E.aij, G.aij)


def test_schedule_bounds(oracle_mod):
    # spike_analysis.c:38-76: rows[i] = i*floor(n/P), remainder in the last partition
    rows, red = oracle_mod.schedule(10, 4, 1, 1)
    assert rows.tolist() == [0, 2, 4, 6, 10]
    assert red.tolist() == [0, 2, 4, 6, 8]
    rows, red = oracle_mod.schedule(1000003, 32, 64, 64)
    assert rows[1] == 31250 and rows[-1] == 1000003 and red[-1] == 32 * 128


def test_penta10_reduced_system_golden(oracle_mod):
    A = oracle_mod.load_csr(ref_path("spike/penta_10.bin"))
    b = np.ones((A.n, 1))
    X, info = oracle_mod.spike_solve(A, b, 2)
    np.testing.assert_allclose(info["R"].to_dense(), PENTA10_R, atol=6e-6)
    np.testing.assert_allclose(info["g"][:, 0], PENTA10_G, atol=6e-6)
    # "SPIKE error: 3.51E-16" in the readme: machine precision
    assert oracle_mod.residual(A, X, b)[0] < 1e-14


def test_penta10_reduced_pattern(oracle_mod):
    # CSR pattern of spike_matrix.c:657-740: nnz = k2*(ku+1) + k2*(kl+1) for P = 2
    A = oracle_mod.load_csr(ref_path("spike/penta_10.bin"))
    _, info = oracle_mod.spike_solve(A, np.ones((A.n, 1)), 2)
    R = info["R"]
    assert R.nnz == 4 * 3 + 4 * 3
    assert R.rowptr.tolist() == [0, 3, 6, 9, 12, 15, 18, 21, 24]
    assert R.colind[:3].tolist() == [0, 4, 5] and R.colind[12:15].tolist() == [2, 3, 4]


def test_solve_small_golden(oracle_mod):
    A = oracle_mod.load_csr(ref_path("solve_small/tridiagonal.bin"))
    b = np.ones((A.n, 1))
    for P in (1, 2, 4):
        X, _ = oracle_mod.spike_solve(A, b, P)
        np.testing.assert_allclose(X[:, 0], SOLVE_SMALL_X, atol=5.1e-4)
        assert oracle_mod.residual(A, X, b)[0] < 1e-15 * 10


@pytest.mark.parametrize("name,rel,P", [
    ("hepta40", "heptadiagonal/medium.bin", 4), ("banded10", "bandwidth/banded.bin", 2),
    ("penta15", "spike/penta_15.bin", 3), ("tridiag10", "solve_small/tridiagonal.bin", 4)])
def test_tips_vs_dense_inverse(oracle_mod, golden, name, rel, P):
    A = oracle_mod.load_csr(ref_path(rel))
    b = golden["b_%s_rand" % name]
    X, info = oracle_mod.spike_solve(A, b, P)
    for k in ("Vt", "Vb", "Wt", "Wb", "yt", "yb"):
        np.testing.assert_allclose(info[k], golden["tips_%s_P%d_%s" % (name, P, k)], rtol=1e-9, atol=1e-12,
                                   err_msg=k)
    np.testing.assert_allclose(X, golden["x_%s_rand" % name], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("name,rel", [
    ("tridiag10", "solve_small/tridiagonal.bin"), ("penta10", "spike/penta_10.bin"),
    ("penta15", "spike/penta_15.bin"), ("hepta20", "heptadiagonal/small.bin"),
    ("hepta40", "heptadiagonal/medium.bin"), ("banded10", "bandwidth/banded.bin"),
    ("penta20k", "pentadiagonal/large.bin")])
def test_fixture_solutions_vs_superlu(oracle_mod, golden, name, rel):
    A = oracle_mod.load_csr(ref_path(rel))
    n = A.n
    P = 4 if n >= 40 else 2
    X1, _ = oracle_mod.spike_solve(A, np.ones((n, 1)), P)
    np.testing.assert_allclose(X1, golden["x_%s_ones" % name], rtol=1e-8, atol=1e-11)
    X3, _ = oracle_mod.spike_solve(A, golden["b_%s_rand" % name], P)
    np.testing.assert_allclose(X3, golden["x_%s_rand" % name], rtol=1e-8, atol=1e-11)
    Xd = oracle_mod.direct_solve(A, golden["b_%s_rand" % name])
    np.testing.assert_allclose(Xd, golden["x_%s_rand" % name], rtol=1e-8, atol=1e-11)


@pytest.mark.parametrize("key", ["band_300_5_3", "band_257_16_16", "band_1000_2_7"])
def test_band_lu_vs_lapack_gbsv(oracle_mod, golden, key):
    n = int(key.split("_")[1])
    A = oracle_mod.CSR(n, golden[key + "_rowptr"], golden[key + "_colind"], golden[key + "_aij"])
    b = golden[key + "_b"]
    np.testing.assert_allclose(oracle_mod.direct_solve(A, b), golden[key + "_x"], rtol=1e-10, atol=1e-13)
    for P in (2, 5):
        X, _ = oracle_mod.spike_solve(A, b, P)
        np.testing.assert_allclose(X, golden[key + "_x"], rtol=1e-10, atol=1e-13)


@pytest.mark.parametrize("name,rel", [("zpenta15", "complex16/penta_15.z"), ("zpenta1k", "complex16/penta_1k.z")])
def test_complex_planar_fixtures(oracle_mod, golden, name, rel):
    A = oracle_mod.load_csr(ref_path(rel))
    assert A.planar and A.is_complex
    b = golden["b_%s" % name]
    X, _ = oracle_mod.spike_solve(A, b, 3)
    np.testing.assert_allclose(X, golden["x_%s" % name], rtol=1e-9, atol=1e-12)
    assert oracle_mod.residual(A, X, b).max() < 1e-13


def test_residual_closed_form(oracle_mod):
    # r = A x - b with x = e_j picks column j of A: ||A e_j - 0|| / ||0|| -> absolute norm
    A = oracle_mod.load_csr(ref_path("heptadiagonal/small.bin"))
    D = A.to_dense()
    x = np.zeros((A.n, 2)); x[3, 0] = 1.0; x[7, 1] = 2.0
    b = np.ones((A.n, 2))
    rel = oracle_mod.residual(A, x, b)
    want = np.linalg.norm(D @ x - b, axis=0) / np.linalg.norm(b, axis=0)
    np.testing.assert_allclose(rel, want, rtol=1e-13)


def test_generator_seed314_regenerates_reference_file(golden):
    # Tests/large/tridiagonal.bin (4 MB, not copied) regenerates bit for bit from the seed-314
    # tridiagonal generator (generator_banded.py:10-26); sha recorded by make_golden.py.
    import scipy.sparse as sp
    np.random.seed(314)
    n = 100000
    A = (sp.diags(np.random.rand(n) + 1.0, 0) + sp.diags(np.random.rand(n - 1), 1)
         + sp.diags(np.random.rand(n - 1), -1)).tocsr()
    blob = np.array([n, A.nnz, 0], np.int32).tobytes() + A.data.tobytes() + \
        A.indices.astype(np.int32).tobytes() + A.indptr.astype(np.int32).tobytes()
    assert hashlib.sha256(blob).hexdigest() == bytes(golden["sha_tridiag_100000"]).decode() == \
        "e50ba765e51659c9234b16ee46947479bdaff6d9cac19aa8fb463af1f84ec989"


def test_spike_invariants_random(oracle_mod):
    # A_i V_i = B_i (checked through the reduced identity): full solution tips == xr
    rng = np.random.RandomState(5)
    n, k = 400, 6
    D = np.zeros((n, n))
    for d in range(-k, k + 1):
        idx = np.arange(max(0, -d), min(n, n - d))
        D[idx, idx + d] = rng.uniform(-1, 1, idx.size)
    D[np.arange(n), np.arange(n)] += np.abs(D).sum(1) * 1.05
    import scipy.sparse as sp
    S = sp.csr_matrix(D)
    A = oracle_mod.CSR(n, S.indptr, S.indices, S.data)
    b = rng.uniform(-1, 1, (n, 3))
    P = 5
    X, info = oracle_mod.spike_solve(A, b, P)
    rows, _ = oracle_mod.schedule(n, P, k, k)
    xr = info["xr"]
    for p in range(P):
        r0, rf = rows[p], rows[p + 1]
        np.testing.assert_allclose(xr[p * 2 * k:p * 2 * k + k], X[r0:r0 + k], rtol=1e-10, atol=1e-13)
        np.testing.assert_allclose(xr[p * 2 * k + k:(p + 1) * 2 * k], X[rf - k:rf], rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(X, np.linalg.solve(D, b), rtol=1e-10, atol=1e-13)
