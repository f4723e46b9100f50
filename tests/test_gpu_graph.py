"""Graph replay of the solve (opts.graph): the decisions of a solve are taken on the device, so a
solve is one fixed sequence of launches, captured once and replayed.  Every replay -- with eager
work of the program and synchronizations in between -- must give the eager solve's bits (the
stated tolerance is zero), for a single handle and for a batch, in both precisions."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import hspike_amd
    hspike_amd.load()
    return torch, hspike_amd


@pytest.mark.parametrize("prec,n,k,nrhs", [("mixed", 200_000, 64, 16), ("fp64", 120_000, 32, 5), ("mixed", 50_000, 8, 1)])
def test_handle_replay_equals_eager(env, oracle_mod, prec, n, k, nrhs):
    torch, hs = env
    dev = torch.device("cuda:0")
    A = hs.gen_banded(n, k, k, fill=0.3, dom=(1.05, 1.05), diag_floor=1.0, seed=n + k)
    B = hs.gen_dense(n, nrhs, seed=3)
    rp, ci, av = (torch.from_numpy(a).to(dev) for a in (A.rowptr, A.colind, A.aij))
    b = torch.from_numpy(np.ascontiguousarray(B.T)).to(dev)
    ref = hs.Solver(prec)
    ref.set_matrix_dev(n, rp, ci, av)
    ref.factor()
    xe = torch.zeros_like(b)
    ref.solve_dev(b, xe, nrhs)
    torch.cuda.synchronize()
    s = hs.Solver(prec, graph=True)
    s.set_matrix_dev(n, rp, ci, av)
    s.factor()
    x = torch.zeros_like(b)
    replays = 0
    for step in range(5):
        x.fill_(float("nan"))                               # eager work of the program between solves
        s.solve_dev(b, x, nrhs)
        torch.cuda.synchronize()
        replays += s.stats()["graph_replayed"]
        assert torch.equal(x, xe), step
        if step == 2:
            s.factor()                                      # a new factorization keeps the graph
    assert replays >= 3
    st = s.stats()
    assert st["max_rel_residual"] <= 1e-10
    Xo = oracle_mod.direct_solve(oracle_mod.CSR(A.n, A.rowptr, A.colind, A.aij), B)
    X = x.cpu().numpy().T
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) <= 1e-9
    # new right-hand-side arrays: captured again, same bits as eager
    b2 = torch.from_numpy(np.ascontiguousarray(hs.gen_dense(n, nrhs, seed=4).T)).to(dev)
    x2, x2e = torch.zeros_like(b2), torch.zeros_like(b2)
    ref.solve_dev(b2, x2e, nrhs)
    for _ in range(3):
        s.solve_dev(b2, x2, nrhs)
    torch.cuda.synchronize()
    assert s.stats()["graph_replayed"] == 1 and torch.equal(x2, x2e)


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_batch_replay_equals_eager(env, prec):
    torch, hs = env
    dev = torch.device("cuda:0")
    rng = np.random.RandomState(9)
    systems = []
    for i in range(48):
        n = int(np.exp(rng.uniform(np.log(1000), np.log(40000))))
        k = int(rng.choice([4, 8, 16, 32, 64]))
        nrhs = int(rng.randint(1, 9))
        A = hs.gen_banded(n, k, k, fill=0.3, dom=(1.1, 1.1), diag_floor=1.0, seed=500 + i)
        B = hs.gen_dense(n, nrhs, seed=900 + i)
        systems.append(dict(n=n, nrhs=nrhs, rowptr=torch.from_numpy(A.rowptr).to(dev),
                            colind=torch.from_numpy(A.colind).to(dev), aij=torch.from_numpy(A.aij).to(dev),
                            b=torch.from_numpy(np.ascontiguousarray(B.T)).to(dev),
                            x=torch.zeros((nrhs, n), dtype=torch.float64, device=dev)))
    hs.Batch(prec).solve_dev(systems)
    torch.cuda.synchronize()
    eager = [s["x"].clone() for s in systems]
    bt = hs.Batch(prec, graph=True)
    replays = 0
    for step in range(4):
        for s in systems:
            s["x"].zero_()
        rel, st = bt.solve_dev(systems)
        torch.cuda.synchronize()
        replays += st["graph_replayed"]
        assert rel.max() <= 1e-10
        for i, s in enumerate(systems):
            assert torch.equal(s["x"], eager[i]), (step, i)
    assert replays >= 2
