"""Batched solves of many independent systems (spike_batch_solve_dev): every system meets the
residual bar, agrees with the oracle's FP64 direct solve, and its solution is bit-for-bit the one
it gets when solved alone (the schedule depends only on the system, reductions run in a fixed
order per system, and converged systems are masked out of later passes).

Tolerances: per-column relative residual <= 1e-10; ||x - x64|| / ||x64|| <= 1e-8 (mixed) or
1e-11 (FP64) against the oracle; alone vs batch: exact equality."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import hspike_amd
    hspike_amd.load()
    return torch, hspike_amd


def make_batch(torch, hs, nsys, seed, cplx=False, n_hi=20000):
    rng = np.random.RandomState(seed)
    dev = torch.device("cuda:0")
    ks = (4, 8, 16, 32, 64) if not cplx else (2, 4, 16, 40, 64)
    out = []
    for i in range(nsys):
        n = int(np.exp(rng.uniform(np.log(600), np.log(n_hi))))
        k = int(rng.choice(ks))
        nrhs = int(rng.randint(1, 9))
        A = hs.gen_banded(n, k, k, fill=0.3, dom=(1.1, 1.1), diag_floor=1.0, seed=seed * 1000 + i, complex=cplx)
        B = hs.gen_dense(n, nrhs, seed=seed * 1000 + 500 + i, complex=cplx)
        dt = torch.complex128 if cplx else torch.float64
        out.append(dict(n=n, k=k, nrhs=nrhs, A=A, B=B, rowptr=torch.from_numpy(A.rowptr).to(dev),
                        colind=torch.from_numpy(A.colind).to(dev), aij=torch.from_numpy(A.aij).to(dev),
                        b=torch.from_numpy(np.ascontiguousarray(B.T)).to(dev), x=torch.zeros((nrhs, n), dtype=dt, device=dev)))
    return out


@pytest.mark.parametrize("prec,cplx", [("mixed", False), ("fp64", False), ("mixed", True)])
def test_batch_matches_oracle_and_alone(env, oracle_mod, prec, cplx):
    torch, hs = env
    systems = make_batch(torch, hs, 24 if not cplx else 10, seed=5 if not cplx else 6, cplx=cplx)
    bt = hs.Batch(prec, tol=1e-10)
    rel, st = bt.solve_dev(systems, complex=cplx)
    torch.cuda.synchronize()
    assert rel.max() <= 1e-10, rel.max()
    bound = 1e-8 if prec == "mixed" else 1e-11
    for s in systems:
        X = s["x"].cpu().numpy().T
        r = hs.residual(s["A"], X, s["B"])
        assert r.max() <= 1e-10, (s["n"], s["k"], r.max())
        X64 = oracle_mod.direct_solve(oracle_mod.CSR(s["A"].n, s["A"].rowptr, s["A"].colind, s["A"].aij), s["B"])
        assert np.linalg.norm(X - X64) / np.linalg.norm(X64) <= bound
    # each system alone: same partitions (autotuned from the system only), same bits
    for s in systems:
        solo = hs.Solver(prec, tol=1e-10)
        solo.set_matrix_dev(s["n"], s["rowptr"], s["colind"], s["aij"])
        xs = torch.zeros_like(s["x"])
        solo.solve_dev(s["b"], xs, s["nrhs"])
        torch.cuda.synchronize()
        assert torch.equal(xs, s["x"]), (s["n"], s["k"], s["nrhs"])
        solo.close()
    bt.close()


def test_batch_repeat_is_deterministic(env):
    torch, hs = env
    systems = make_batch(torch, hs, 12, seed=9)
    bt = hs.Batch("mixed", tol=1e-10)
    bt.solve_dev(systems)
    first = [s["x"].clone() for s in systems]
    for s in systems:
        s["x"].zero_()
    bt.solve_dev(bt.prepare(systems))          # the same batch through its prepared descriptors
    torch.cuda.synchronize()
    for a, s in zip(first, systems):
        assert torch.equal(a, s["x"])
    bt.close()


def test_windowed_batch_matches_whole(env, monkeypatch):
    """A batch too large for HBM is solved in windows of systems (the window halves on every
    failed allocation); HSPIKE_BATCH_WINDOW starts it small.  Windows change no bits."""
    torch, hs = env
    systems = make_batch(torch, hs, 11, seed=21)
    whole = hs.Batch("mixed", tol=1e-10)
    rel_whole, st_whole = whole.solve_dev(systems)
    ref = [s["x"].clone() for s in systems]
    whole.close()
    for s in systems:
        s["x"].zero_()
    monkeypatch.setenv("HSPIKE_BATCH_WINDOW", "3")
    win = hs.Batch("mixed", tol=1e-10)
    rel_win, st_win = win.solve_dev(systems)
    torch.cuda.synchronize()
    for a, s in zip(ref, systems):
        assert torch.equal(a, s["x"])
    np.testing.assert_array_equal(rel_whole, rel_win)
    assert st_win["max_rel_residual"] <= 1e-10
    win.close()


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_throughput_tuning_matches_alone_and_oracle(env, oracle_mod, prec):
    """Partitions autotuned for throughput (long partitions, for batches of many systems): each
    system's bits equal its solve alone with the same tuning, and the oracle's solution."""
    torch, hs = env
    systems = make_batch(torch, hs, 16, seed=31, n_hi=60000)
    bt = hs.Batch(prec, tol=1e-10, tune="throughput")
    rel, st = bt.solve_dev(systems)
    torch.cuda.synchronize()
    assert rel.max() <= 1e-10
    for s in systems:
        X = s["x"].cpu().numpy().T
        X64 = oracle_mod.direct_solve(oracle_mod.CSR(s["A"].n, s["A"].rowptr, s["A"].colind, s["A"].aij), s["B"])
        assert np.linalg.norm(X - X64) / np.linalg.norm(X64) <= (1e-8 if prec == "mixed" else 1e-11)
        solo = hs.Solver(prec, tol=1e-10, tune="throughput")
        solo.set_matrix_dev(s["n"], s["rowptr"], s["colind"], s["aij"])
        xs = torch.zeros_like(s["x"])
        solo.solve_dev(s["b"], xs, s["nrhs"])
        torch.cuda.synchronize()
        assert torch.equal(xs, s["x"])
        assert solo.stats()["P"] == hs.autotune_partitions(s["n"], s["k"], s["k"], s["nrhs"],
                                                           precision=hs.SPIKE_PREC_MIXED if prec == "mixed" else 0,
                                                           tune="throughput")
        solo.close()
    bt.close()


def test_oversubscribed_b64_factor_matches_alone(env, oracle_mod):
    """A mixed-precision batch of half-bandwidth-64 systems (b = 64) whose partitions outnumber the
    factor workgroups the device runs at once takes the factor variant that keeps one LDS copy of
    R's top rows; each system alone (fewer partitions than that) takes the two-copy variant.  Both
    form the same products, so the solutions are bit-for-bit equal; the first is also checked
    against the oracle's FP64 direct solve (||x - x64|| / ||x64|| <= 1e-8)."""
    torch, hs = env
    dev = torch.device("cuda:0")
    n, k = 300000, 64
    cap = hs.autotune_partitions(10 ** 8, k, k, precision=hs.SPIKE_PREC_MIXED)    # one round of the factor
    p1 = hs.autotune_partitions(n, k, k, precision=hs.SPIKE_PREC_MIXED)
    nsys = cap // p1 + 1
    assert nsys * p1 > cap and p1 < cap
    systems = []
    for i in range(nsys):
        A = hs.gen_banded(n, k, k, fill=0.2, dom=(1.1, 1.1), diag_floor=1.0, seed=7100 + i)
        B = hs.gen_dense(n, 2, seed=7200 + i)
        systems.append(dict(n=n, nrhs=2, A=A, B=B, rowptr=torch.from_numpy(A.rowptr).to(dev),
                            colind=torch.from_numpy(A.colind).to(dev), aij=torch.from_numpy(A.aij).to(dev),
                            b=torch.from_numpy(np.ascontiguousarray(B.T)).to(dev),
                            x=torch.zeros((2, n), dtype=torch.float64, device=dev)))
    bt = hs.Batch("mixed", tol=1e-10)
    rel, st = bt.solve_dev(systems)
    torch.cuda.synchronize()
    assert rel.max() <= 1e-10, rel.max()
    bt.close()
    s = systems[0]
    X64 = oracle_mod.direct_solve(oracle_mod.CSR(s["A"].n, s["A"].rowptr, s["A"].colind, s["A"].aij), s["B"])
    X = s["x"].cpu().numpy().T
    assert np.linalg.norm(X - X64) / np.linalg.norm(X64) <= 1e-8
    for s in systems:
        solo = hs.Solver("mixed", tol=1e-10)
        solo.set_matrix_dev(s["n"], s["rowptr"], s["colind"], s["aij"])
        xs = torch.zeros_like(s["x"])
        solo.solve_dev(s["b"], xs, s["nrhs"])
        torch.cuda.synchronize()
        assert torch.equal(xs, s["x"])
        solo.close()


def test_narrow_coupling_blocks_after_wide_ones(env, oracle_mod):
    """With FP32 factors the empty column groups of the coupling blocks B_j / C_j are neither
    written nor read.  One Batch object first solves full-band systems (every group of its 16- and
    64-row blocks holds values), then asymmetric narrow ones of the same block sizes in the same
    arrays: (ku, kl) = (3, 12), (12, 3), (1, 7) at b = 16 and (40, 20), (20, 40) at b = 64.  Their
    solutions must be bit-for-bit those of a fresh solver (stale values never reach them, and B
    follows ku, C follows kl), and match the oracle's FP64 direct solve (<= 1e-8)."""
    torch, hs = env
    dev = torch.device("cuda:0")

    def systems_for(shapes, seed):
        out = []
        for i, (n, ku, kl) in enumerate(shapes):
            A = hs.gen_banded(n, ku, kl, fill=0.6, dom=(1.1, 1.1), diag_floor=1.0, seed=seed + i)
            B = hs.gen_dense(n, 3, seed=seed + 50 + i)
            out.append(dict(n=n, nrhs=3, A=A, B=B, rowptr=torch.from_numpy(A.rowptr).to(dev),
                            colind=torch.from_numpy(A.colind).to(dev), aij=torch.from_numpy(A.aij).to(dev),
                            b=torch.from_numpy(np.ascontiguousarray(B.T)).to(dev),
                            x=torch.zeros((3, n), dtype=torch.float64, device=dev)))
        return out

    bt = hs.Batch("mixed", tol=1e-10)
    wide = systems_for([(40000, 16, 16), (30000, 15, 16), (40000, 64, 64), (30000, 60, 64)], 8100)
    rel, _ = bt.solve_dev(wide)
    torch.cuda.synchronize()
    assert rel.max() <= 1e-10
    narrow = systems_for([(40000, 3, 12), (30000, 12, 3), (40000, 1, 7), (40000, 40, 20), (30000, 20, 40)], 8200)
    rel, _ = bt.solve_dev(narrow)
    torch.cuda.synchronize()
    assert rel.max() <= 1e-10, rel.max()
    bt.close()
    for s in narrow:
        X64 = oracle_mod.direct_solve(oracle_mod.CSR(s["A"].n, s["A"].rowptr, s["A"].colind, s["A"].aij), s["B"])
        X = s["x"].cpu().numpy().T
        assert np.linalg.norm(X - X64) / np.linalg.norm(X64) <= 1e-8
        fresh = hs.Batch("mixed", tol=1e-10)
        xs = dict(s)
        xs["x"] = torch.zeros_like(s["x"])
        fresh.solve_dev([xs])
        torch.cuda.synchronize()
        assert torch.equal(xs["x"], s["x"])
        fresh.close()
