"""Every benchmark workload at its full size on the GPU, through the C ABI, with the inputs
bench.py generates (same seeds, same generators).

  C1  n = 10^4 pentadiagonal, a_ii = sum|a_ij| + 1, 4 partitions, 1 RHS, FP64: tips, reduced
      solution and x against the oracle's SPIKE at 4 partitions (the reference's P).
  C2  n = 10^6, half-bandwidth 64, 15% fill, diagonal 1.05 sum|offdiag|, 16 RHS, FP64 and mixed,
      at the configured 32 partitions and autotuned: x against the oracle's SPIKE at 32
      partitions (||x - x64|| / ||x64|| <= 1e-9), every column's residual <= 1e-10; at 32
      partitions in FP64 also the tips and the reduced solution.
  C3  n = 2*10^6, half-bandwidth 128, 20% fill, dominance U[1.01, 2], 64 RHS, mixed refined to
      1e-12: every column's GPU residual <= 1e-12, re-checked on the CPU (scipy) for 4 sampled
      columns, and the mixed solution against the plain FP64 GPU solution (<= 1e-9); the oracle
      at a tenth of the rows with the same shape.
  C4  the full 4096-system batch: every system's residual <= 1e-10, 48 sampled systems against
      the oracle's band LU (<= 1e-9), 6 of them solved alone equal bit for bit.
  C5  the request stream (Poisson arrivals, 80% real / 20% complex128, 30% repeats, nrhs 1-64)
      with rows capped at 1.5*10^5 so the oracle solves every request: each against the oracle's
      band LU (<= 1e-9) and equal bit for bit to the same request solved alone.
The residual bar is the metric's: ||A x - b|| / ||b|| per column, FP64."""
import os

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hs():
    import hspike_amd
    hspike_amd.load()
    return hspike_amd


def rel_err(x, ref):
    return np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300)


def cfg_matrix(hs, name, n=None):
    c = bench.CONFIGS[name]
    A = hs.gen_banded(n or c["n"], c["k"], c["k"], fill=c["fill"], dom=c["dom"], diag_add=c["diag_add"], diag_floor=1.0,
                      seed=bench.SEED_A)
    return A, c


def threads():
    return max(1, min(32, os.cpu_count() or 1))


def test_c1_full(hs, oracle_mod, monkeypatch):
    A, c = cfg_matrix(hs, "c1")
    assert c["P_cfg"] == 4 and A.bandwidth() == (2, 2) and A.nnz == 5 * A.n - 6
    B = hs.gen_dense(A.n, c["nrhs"], seed=bench.SEED_B)
    Ao = oracle_mod.CSR(A.n, A.rowptr, A.colind, A.aij)
    Xo, info = oracle_mod.spike_solve(Ao, B, 4)
    s = hs.Solver("fp64", partitions=4)
    s.set_matrix(A)
    s.factor()
    X = s.solve(B)
    t = s.tips()
    for k in ("Vt", "Vb", "Wt", "Wb", "yt", "yb"):
        np.testing.assert_allclose(t[k], info[k], atol=1e-12, err_msg=k)
    np.testing.assert_allclose(s.reduced_solution(), info["xr"], atol=1e-12)
    assert rel_err(X, Xo) <= 1e-12
    assert hs.residual(A, X, B).max() <= 1e-10
    monkeypatch.setenv("HSPIKE_PARTITIONS", "4")            # the legacy entry point at the same P
    assert rel_err(hs.dspike_core_host(A, B), Xo) <= 1e-12


@pytest.fixture(scope="module")
def c2(hs, oracle_mod):
    A, c = cfg_matrix(hs, "c2")
    B = hs.gen_dense(A.n, c["nrhs"], seed=bench.SEED_B)
    Xo, info = oracle_mod.spike_solve(oracle_mod.CSR(A.n, A.rowptr, A.colind, A.aij), B, c["P_cfg"], nthreads=threads())
    return A, B, Xo, info, c


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("parts", [32, 0])
def test_c2_full(hs, c2, prec, parts):
    A, B, Xo, info, c = c2
    s = hs.Solver(prec, partitions=parts, tol=c["tol"])
    s.set_matrix(A)
    s.factor()
    X = s.solve(B)
    st = s.stats()
    assert st["b"] == 64 and (parts == 0 or st["P"] == 32)
    assert rel_err(X, Xo) <= 1e-9
    rel = hs.residual(A, X, B)
    assert rel.max() <= 1e-10 and abs(rel.max() - st["max_rel_residual"]) <= 1e-3 * st["max_rel_residual"] + 1e-17
    if prec == "fp64" and parts == 32:
        t = s.tips()
        for k in ("Vt", "Vb", "Wt", "Wb", "yt", "yb"):
            np.testing.assert_allclose(t[k], info[k], atol=1e-10, err_msg=k)
        if st["refine_iters"] == 0:
            np.testing.assert_allclose(s.reduced_solution(), info["xr"], atol=1e-10)


def test_c3_full(hs):
    A, c = cfg_matrix(hs, "c3")
    B = hs.gen_dense(A.n, c["nrhs"], seed=bench.SEED_B)
    s = hs.Solver("mixed", tol=c["tol"])
    s.set_matrix(A)
    s.factor()
    X = s.solve(B)
    st = s.stats()
    assert st["b"] == 128 and st["max_rel_residual"] <= 1e-12
    assert hs.residual(A, X, B).max() <= 1e-12
    S = A.to_scipy()
    for col in (0, 21, 42, 63):                                 # CPU re-check, FP64 sparse product
        r = S @ X[:, col] - B[:, col]
        assert np.linalg.norm(r) / np.linalg.norm(B[:, col]) <= 1.5e-12
    f = hs.Solver("fp64", tol=c["tol"])
    f.set_matrix(A)
    f.factor()
    assert rel_err(X, f.solve(B)) <= 1e-9


def test_c3_shape_vs_oracle(hs, oracle_mod):
    A, c = cfg_matrix(hs, "c3", n=200_000)
    B = hs.gen_dense(A.n, 4, seed=bench.SEED_B)
    Xo, _ = oracle_mod.spike_solve(oracle_mod.CSR(A.n, A.rowptr, A.colind, A.aij), B, 16, nthreads=threads(),
                                   want_reduced=False)
    s = hs.Solver("mixed", tol=c["tol"])
    s.set_matrix(A)
    s.factor()
    X = s.solve(B)
    assert rel_err(X, Xo) <= 1e-10
    assert hs.residual(A, X, B).max() <= 1e-12


def test_c4_full_batch(hs, oracle_mod):
    import torch
    dev = torch.device("cuda:0")
    spec = bench.c4_systems(bench.C4["nsys"], bench.C4["seed"])
    sample = set(range(0, len(spec), 85))                     # 49 systems checked on the CPU
    systems, host = [], {}
    for i, (n, k, nrhs, sa, sb) in enumerate(spec):
        A = hs.gen_banded(n, k, k, fill=bench.C4["fill"], dom=(bench.C4["dom"], bench.C4["dom"]), diag_floor=1.0, seed=sa)
        B = hs.gen_dense(n, nrhs, seed=sb)
        systems.append(dict(n=n, nrhs=nrhs, rowptr=torch.from_numpy(A.rowptr).to(dev),
                            colind=torch.from_numpy(A.colind).to(dev), aij=torch.from_numpy(A.aij).to(dev),
                            b=torch.from_numpy(np.ascontiguousarray(B.T)).to(dev),
                            x=torch.zeros((nrhs, n), dtype=torch.float64, device=dev)))
        if i in sample:
            host[i] = (A, B)
    bt = hs.Batch("mixed", tol=bench.C4["tol"], tune="throughput")        # as bench.py solves it
    rel, st = bt.solve_dev(systems)
    torch.cuda.synchronize()
    assert len(rel) == 4096 and rel.max() <= 1e-10
    for n_done, (i, (A, B)) in enumerate(sorted(host.items())):
        X = systems[i]["x"].cpu().numpy().T
        Xo = oracle_mod.direct_solve(oracle_mod.CSR(A.n, A.rowptr, A.colind, A.aij), B)
        assert rel_err(X, Xo) <= 1e-9, (i, spec[i])
        if n_done < 6:                                         # the same system alone: same bits
            one = [dict(systems[i])]
            one[0]["x"] = torch.zeros_like(systems[i]["x"])
            hs.Batch("mixed", tol=bench.C4["tol"], tune="throughput").solve_dev(one)
            torch.cuda.synchronize()
            assert torch.equal(one[0]["x"], systems[i]["x"]), i


def test_c5_stream_vs_oracle_and_alone(hs, oracle_mod):
    reqs, mats = bench.c5_stream(30, bench.C5["seed"], n_hi=8e4)
    A = [hs.gen_banded(n, k, k, fill=bench.C5["fill"], dom=(bench.C5["dom"], bench.C5["dom"]), diag_floor=1.0, seed=sd,
                       complex=cplx) for (cplx, n, k, sd) in mats]
    Bs = [hs.gen_dense(A[mid].n, nrhs, seed=sb, complex=mats[mid][0]) for (_, mid, nrhs, sb) in reqs]
    assert any(m[0] for m in mats) and len(mats) < len(reqs)    # complex requests and repeats are in the mix
    srv = hs.Server("mixed", tol=bench.C5["tol"], workers=3, max_batch=8)
    tickets = [srv.submit(A[mid], Bs[i], key=mid + 1) for i, (_, mid, _, _) in enumerate(reqs)]
    outs = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["failed"] == 0 and st["completed"] == len(reqs)
    assert sum(info["resident_hit"] for _, info in outs) >= 1
    for i, ((X, info), (_, mid, _, _)) in enumerate(zip(outs, reqs)):
        assert info["status"] == 0 and info["max_rel_residual"] <= 1e-10
        Xo = oracle_mod.direct_solve(oracle_mod.CSR(A[mid].n, A[mid].rowptr, A[mid].colind, A[mid].aij), Bs[i])
        assert rel_err(X, Xo) <= 1e-9, (i, mats[mid])
        s = hs.Solver("mixed", tol=bench.C5["tol"])
        s.set_matrix(A[mid])
        np.testing.assert_array_equal(X, s.solve(Bs[i]))      # alone == in the mix
        s.close()
