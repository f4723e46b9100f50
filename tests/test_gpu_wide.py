"""Block sizes whose factor state lives in global memory (spike_kernels_wide.inc): b = 256 for
half-bandwidths 129..256 in every precision, and complex FP64 at b = 128.  The same SPIKE
quantities as the register kernels, checked against the CPU oracle.

Tolerances: FP64 -- spike tips and reduced solution within 1e-10 (absolute, entries are O(1)),
solution within 1e-9 relative to the oracle, every column's relative residual <= 1e-10.  Mixed
-- the residual bar after FP64 refinement and ||x - x64|| / ||x64|| <= 1e-8 against the oracle's
FP64 direct solve.  Alone vs batch / server: exact equality."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hs():
    import hspike_amd
    hspike_amd.load()
    return hspike_amd


def to_o(oracle_mod, A):
    return oracle_mod.CSR(A.n, A.rowptr, A.colind, A.aij)


def rel_err(x, ref):
    return np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300)


@pytest.mark.parametrize("n,ku,kl,fill,P,nrhs,cplx", [
    (6000, 150, 150, 0.2, 4, 3, False),     # b = 256
    (5000, 256, 256, 0.1, 3, 2, False),     # the widest band
    (7001, 200, 140, 0.15, 5, 17, False),   # ku != kl, uneven last partition, two RHS tiles
    (4000, 100, 90, 0.2, 4, 2, True),       # complex FP64 at b = 128
    (5000, 180, 180, 0.1, 3, 2, True),      # complex FP64 at b = 256
])
def test_wide_fp64_parity(hs, oracle_mod, n, ku, kl, fill, P, nrhs, cplx):
    """Spike tips, reduced solution and solution against the oracle's SPIKE at the same P."""
    A = hs.gen_banded(n, ku, kl, fill=fill, dom=(1.05, 1.05), diag_floor=1.0, seed=n + ku, complex=cplx)
    B = hs.gen_dense(n, nrhs, seed=3, complex=cplx)
    s = hs.Solver("fp64", partitions=P)
    s.set_matrix(A)
    X = s.solve(B)
    st = s.stats()
    assert st["b"] == (256 if max(ku, kl) > 128 else 128)
    Xo, info = oracle_mod.spike_solve(to_o(oracle_mod, A), B, P)
    if not cplx:
        t = s.tips(nrhs)
        for k in ("Vt", "Vb", "Wt", "Wb", "yt", "yb"):
            np.testing.assert_allclose(t[k], info[k], atol=1e-10, err_msg=k)
        np.testing.assert_allclose(s.reduced_solution(nrhs), info["xr"], atol=1e-10)
    assert rel_err(X, Xo) <= 1e-9
    rel = hs.residual(A, X, B)
    assert rel.max() <= 1e-10, rel.max()
    s.close()


@pytest.mark.parametrize("n,k,cplx,tol", [
    (40000, 200, False, 1e-12),
    (30000, 256, False, 1e-10),
    (20000, 150, True, 1e-10),
])
def test_wide_mixed_autotuned(hs, oracle_mod, n, k, cplx, tol):
    A = hs.gen_banded(n, k, k, fill=0.1, dom=(1.05, 1.2), diag_floor=1.0, seed=17 + k, complex=cplx)
    B = hs.gen_dense(n, 4, seed=5, complex=cplx)
    s = hs.Solver("mixed", tol=tol)
    s.set_matrix(A)
    X = s.solve(B)
    st = s.stats()
    assert st["b"] == 256 and st["P"] > 1
    assert st["max_rel_residual"] <= tol
    X64 = oracle_mod.direct_solve(to_o(oracle_mod, A), B)
    assert rel_err(X, X64) <= 1e-8
    assert hs.residual(A, X, B).max() <= tol
    # a second solve on the resident factors gives the same bits
    np.testing.assert_array_equal(s.solve(B), X)
    s.close()


def test_wide_single_pass_and_determinism(hs, oracle_mod):
    """FP64 at b = 256 reaches the bar without refinement, and two solves agree bit for bit."""
    n, k, P, nrhs = 24 * 1100 + 37, 140, 24, 5
    A = hs.gen_banded(n, k, k, fill=0.15, dom=(1.05, 1.05), diag_floor=1.0, seed=77)
    B = hs.gen_dense(n, nrhs, seed=8)
    runs = []
    for _ in range(2):
        s = hs.Solver("fp64", partitions=P, max_refine=0, tol=1e-12)
        s.set_matrix(A)
        runs.append((s.solve(B), s.stats()))
        s.close()
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    assert runs[0][1]["refine_iters"] == 0 and runs[0][1]["max_rel_residual"] <= 1e-13
    Xo, _ = oracle_mod.spike_solve(to_o(oracle_mod, A), B, P, want_reduced=False)
    assert rel_err(runs[0][0], Xo) <= 1e-12


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_wide_in_batch_matches_alone(hs, oracle_mod, prec):
    """Systems of b = 16, 64 and 256 (and complex b = 128 for FP64) in one batch: each against
    the oracle and bit for bit equal to its solve alone."""
    import torch
    dev = torch.device("cuda:0")
    spec = [(3000, 8), (9000, 60), (12000, 180), (5000, 250), (7000, 30)]
    systems = []
    for i, (n, k) in enumerate(spec):
        A = hs.gen_banded(n, k, k, fill=0.2, dom=(1.1, 1.1), diag_floor=1.0, seed=300 + i)
        B = hs.gen_dense(n, 1 + i % 3, seed=400 + i)
        systems.append(dict(n=n, nrhs=B.shape[1], A=A, B=B, rowptr=torch.from_numpy(A.rowptr).to(dev),
                            colind=torch.from_numpy(A.colind).to(dev), aij=torch.from_numpy(A.aij).to(dev),
                            b=torch.from_numpy(np.ascontiguousarray(B.T)).to(dev),
                            x=torch.zeros((B.shape[1], n), dtype=torch.float64, device=dev)))
    bt = hs.Batch(prec, tol=1e-10)
    rel, _ = bt.solve_dev(systems)
    torch.cuda.synchronize()
    assert rel.max() <= 1e-10, rel
    for s in systems:
        X = s["x"].cpu().numpy().T
        X64 = oracle_mod.direct_solve(to_o(oracle_mod, s["A"]), s["B"])
        assert rel_err(X, X64) <= (1e-8 if prec == "mixed" else 1e-11), (s["n"],)
        solo = hs.Solver(prec, tol=1e-10)
        solo.set_matrix_dev(s["n"], s["rowptr"], s["colind"], s["aij"])
        xs = torch.zeros_like(s["x"])
        solo.solve_dev(s["b"], xs, s["nrhs"])
        torch.cuda.synchronize()
        assert torch.equal(xs, s["x"]), (s["n"],)
        solo.close()
    bt.close()


def test_wide_direct_reduced_solve(hs, oracle_mod):
    """The reduced system of a b = 256 matrix through the direct band LU (SPIKE_RED_DIRECT)."""
    n, k, P = 6000, 160, 5
    A = hs.gen_banded(n, k, k, fill=0.15, dom=(1.02, 1.02), diag_floor=1.0, seed=91)
    B = hs.gen_dense(n, 2, seed=92)
    s = hs.Solver("fp64", partitions=P, reduced_solver="direct")
    s.set_matrix(A)
    X = s.solve(B)
    assert s.stats()["reduced_direct"] == 1
    Xo, info = oracle_mod.spike_solve(to_o(oracle_mod, A), B, P)
    np.testing.assert_allclose(s.reduced_solution(2), info["xr"], atol=1e-10)
    assert rel_err(X, Xo) <= 1e-9
    s.close()
