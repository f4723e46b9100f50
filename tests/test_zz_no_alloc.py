"""Allocation-free band solve (spike_band_solve_dev_noalloc) against the allocating entry point
(spike_band_solve_host) through the C ABI, on the same inputs: identical solution and residual
bits on two calls in a row (the second smaller, over the scratch area the first one left
behind), agreement with the oracle's FP64 band LU within 1e-9, and a solve too large for the
scratch area rejected with SPIKE_ERR_OOM."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import hspike_amd
    hspike_amd.load()
    assert torch.cuda.is_available()
    return hspike_amd, torch


def well_conditioned_band(hs, n, ku, kl, seed, complex_=False):
    """2 at (i, s(i)) with s reversing blocks of min(ku, kl) + 1 rows, band noise with row sums
    <= 0.8: singular values >= 1.2, row interchanges needed."""
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
    v[i == j] = 0.0
    v[j == np.minimum(blk + s - 1, n - 1) - (i - blk)] = 2.0
    rp = np.zeros(n + 1, dtype=np.int32)
    np.add.at(rp, i + 1, 1)
    return hs.CSRMatrix(n, np.cumsum(rp).astype(np.int32), j.astype(np.int32), v)


def noalloc_solve(hs, torch, A, B):
    dev = torch.device("cuda:0")
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    rp, ci, av = (torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (A.rowptr, A.colind, A.aij))
    b = torch.from_numpy(np.ascontiguousarray(B.T)).to(dev)          # column-major n x m
    x = torch.zeros_like(b)
    rel = np.zeros(B.shape[1])
    torch.cuda.synchronize()
    dt = hs.SPIKE_DTYPE_Z if np.iscomplexobj(B) else hs.SPIKE_DTYPE_D
    rc = hs.load().spike_band_solve_dev_noalloc(dt, A.n, A.nnz, p(rp), p(ci), p(av), B.shape[1], p(b), A.n, p(x), A.n,
                                                rel.ctypes.data_as(ctypes.c_void_p), None)
    torch.cuda.synchronize()
    return rc, x.cpu().numpy().T, rel


@pytest.mark.parametrize("cplx", [False, True])
def test_noalloc_matches_allocating_entry(env, oracle_mod, cplx):
    hs, torch = env
    for n, ku, kl, m in ((3000, 40, 40, 3), (700, 9, 4, 2)):          # second call smaller
        A = well_conditioned_band(hs, n, ku, kl, seed=n, complex_=cplx)
        B = hs.gen_dense(n, m, seed=n + 1, complex=cplx)
        rc, X, rel = noalloc_solve(hs, torch, A, B)
        assert rc == 0, hs.load().spike_last_error_message()
        Xh, relh = hs.band_solve(A, B)
        assert np.array_equal(X, Xh)                                   # same kernels, same bits
        assert np.array_equal(rel, relh)
        assert rel.max() <= 1e-12
        Xo = oracle_mod.direct_solve(oracle_mod.CSR(A.n, A.rowptr, A.colind, A.aij), B)
        assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) <= 1e-9


def test_noalloc_rejects_beyond_scratch(env):
    hs, torch = env
    A = hs.gen_banded(400_000, 8, 8, fill=0.1, dom=(1.1, 1.1), diag_floor=1.0, seed=3)   # band: 400k x 25 x 8 B = 80 MB
    B = hs.gen_dense(A.n, 1, seed=4)
    rc, _, _ = noalloc_solve(hs, torch, A, B)
    assert rc == -3
    assert b"scratch" in hs.load().spike_last_error_message() or b"workspace" in hs.load().spike_last_error_message()
    # the entry point still works after the rejection
    A2 = well_conditioned_band(hs, 500, 5, 5, seed=5)
    B2 = hs.gen_dense(500, 2, seed=6)
    rc, X2, _ = noalloc_solve(hs, torch, A2, B2)
    assert rc == 0
    assert np.array_equal(X2, hs.band_solve(A2, B2)[0])
