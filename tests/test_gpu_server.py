"""Request server (spike_server_*): continuous batching with overlapping jobs, memory admission
and resident factors.  Each result must equal -- bit for bit, the stated tolerance is zero --
the solution of the same request solved alone through a fresh handle, whatever else was in
flight; every column meets the 1e-10 relative-residual bar, and every request agrees with the
oracle's FP64 band LU within 1e-9 (relative)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hs():
    import hspike_amd
    hspike_amd.load()
    return hspike_amd


def make_mats(hs, seed, count):
    rng = np.random.RandomState(seed)
    mats = []
    for i in range(count):
        cplx = rng.rand() < 0.25
        n = int(np.exp(rng.uniform(np.log(2000), np.log(60000))))
        k = int(rng.randint(2, 257))
        mats.append(hs.gen_banded(n, k, k, fill=0.3, dom=(1.2, 1.2), diag_floor=1.0, seed=seed * 100 + i, complex=cplx))
    return mats


def alone(hs, A, B, prec):
    s = hs.Solver(prec, tol=1e-10)
    s.set_matrix(A)
    X = s.solve(B)
    s.close()
    return X


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_mix_equals_alone(hs, oracle_mod, prec):
    mats = make_mats(hs, 3 if prec == "mixed" else 4, 8)
    rng = np.random.RandomState(11)
    srv = hs.Server(prec, tol=1e-10, workers=3, max_batch=4)
    reqs = []
    for r in range(20):                       # every matrix once, then repeats with new RHS
        A = mats[r] if r < len(mats) else mats[rng.randint(len(mats))]
        B = hs.gen_dense(A.n, int(rng.randint(1, 17)), seed=700 + r, complex=A.dtype_code == 1)
        reqs.append((A, B, srv.submit(A, B)))
    hits = 0
    out = []
    for A, B, t in reqs:
        X, info = srv.wait(t)
        assert info["status"] == 0, info
        assert info["max_rel_residual"] <= 1e-10
        hits += info["resident_hit"]
        out.append((A, B, X))
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert hits >= 1 and st["p99_ms"] >= st["p50_ms"] > 0
    for A, B, X in out:
        assert hs.residual(A, X, B).max() <= 1e-10
        Xo = oracle_mod.direct_solve(oracle_mod.CSR(A.n, A.rowptr, A.colind, A.aij), B)
        assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) <= 1e-9
        np.testing.assert_array_equal(X, alone(hs, A, B, prec))


def test_memory_admission_and_eviction(hs):
    mats = [hs.gen_banded(40000, 64, 64, fill=0.2, dom=(1.2, 1.2), seed=50 + i) for i in range(6)]
    small = 0.35                                   # GiB: room for about one such job at a time
    srv = hs.Server("mixed", workers=3, max_batch=4, mem_budget_gb=small)
    tick = [(A, B, srv.submit(A, B)) for A in mats for B in [hs.gen_dense(A.n, 8, seed=A.nnz % 1000)]]
    for A, B, t in tick:
        X, info = srv.wait(t)
        assert info["status"] == 0
        assert hs.residual(A, X, B).max() <= 1e-10
    st = srv.stats()
    assert st["completed"] == len(mats)
    assert st["resident_bytes"] <= st["budget_bytes"]
    # a request that can never fit fails cleanly
    big = hs.gen_banded(2_000_000, 64, 64, fill=0.2, dom=(1.2, 1.2), seed=99)
    X, info = srv.wait(srv.submit(big, np.ones(big.n)))
    assert info["status"] == -3
    srv.close()


def test_repeat_rhs_uses_resident_factors(hs):
    A = hs.gen_banded(30000, 32, 32, fill=0.3, dom=(1.2, 1.2), seed=7)
    B = hs.gen_dense(A.n, 4, seed=8)
    srv = hs.Server("mixed", workers=2)
    X1, i1 = srv.wait(srv.submit(A, B))
    X2, i2 = srv.wait(srv.submit(A, B))
    srv.close()
    assert not i1["resident_hit"] and i2["resident_hit"] and i1["key"] == i2["key"]
    np.testing.assert_array_equal(X1, X2)


def test_large_requests_cross_staging_chunks(hs):
    """Requests whose arrays span several 32 MiB pinned staging chunks of a worker, up and down
    (values 105 MB, column indices 53 MB, RHS and solution 51 MB, real; a complex one of 140 MB of
    values): the solution equals the same request solved alone through a fresh handle bit for bit,
    and meets the residual bar."""
    srv = hs.Server("mixed", tol=1e-10, workers=2, max_batch=2)
    reqs = []
    for i, (n, k, m, cplx) in enumerate(((400000, 16, 16, False), (350000, 12, 5, True))):
        A = hs.gen_banded(n, k, k, fill=1.0, dom=(1.2, 1.2), diag_floor=1.0, seed=9300 + i, complex=cplx)
        B = hs.gen_dense(n, m, seed=9400 + i, complex=cplx)
        reqs.append((A, B, srv.submit(A, B)))
    out = []
    for A, B, t in reqs:
        X, info = srv.wait(t)
        assert info["status"] == 0, info
        out.append((A, B, X))
    srv.close()
    for A, B, X in out:
        assert hs.residual(A, X, B).max() <= 1e-10
        np.testing.assert_array_equal(X, alone(hs, A, B, "mixed"))
