"""Generate the golden vectors in tests/golden/golden.npz (run once; output committed).

Every expected value here comes from numpy/scipy (LAPACK and SuperLU), computed from the
reference's own fixture matrices (tests/golden/ref, copied from srodrb/HSPIKE Tests/) or
from small seeded banded matrices -- never from this repository's oracle or GPU code:

  * x_<name>_ones / x_<name>_rand : scipy.sparse.linalg.splu(A).solve(b) (SuperLU, the
    library family of the reference's SuperLU_MT backend) for b = ones and a seeded U[-1,1]
    panel with 3 columns;
  * tips_<name>_P<p>_* : explicit spike tips from a dense inverse of each diagonal block
    (V_i = A_i^{-1} A[r0:rf, rf:rf+ku], W_i = A_i^{-1} A[r0:rf, r0-kl:r0], top = first ku
    rows, bottom = last kl rows) with the partition rows of spike_analysis.c:38-76;
  * band_<k> : scipy.linalg.solve_banded (LAPACK gbsv) on a seeded banded matrix;
  * sha256 of the seed-314 tridiagonal generator's n=100000 output (generator_banded.py:10-26,
    export_csr2bin :70-172), checked against Tests/large/tridiagonal.bin of the reference.

Usage:  python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))


def load(path):
    raw = np.fromfile(path, dtype=np.uint8)
    n, nnz, dtype = np.frombuffer(raw[:12].tobytes(), dtype=np.int32)
    off = 12
    if dtype == 1:
        re = np.frombuffer(raw[off:off + 8 * nnz].tobytes(), np.float64); off += 8 * nnz
        im = np.frombuffer(raw[off:off + 8 * nnz].tobytes(), np.float64); off += 8 * nnz
        a = re + 1j * im
    else:
        a = np.frombuffer(raw[off:off + 8 * nnz].tobytes(), np.float64); off += 8 * nnz
    ci = np.frombuffer(raw[off:off + 4 * nnz].tobytes(), np.int32); off += 4 * nnz
    rp = np.frombuffer(raw[off:off + 4 * (n + 1)].tobytes(), np.int32)
    return sp.csr_matrix((a, ci, rp), shape=(n, n))


def bw(A):
    C = A.tocoo()
    return int(max(0, (C.col - C.row).max())), int(max(0, (C.row - C.col).max()))


def tips(A, P, nrhs_b):
    n = A.shape[0]
    ku, kl = bw(A)
    D = A.toarray()
    rows = [i * (n // P) for i in range(P)] + [n]
    out = {k: [] for k in ("Vt", "Vb", "Wt", "Wb", "yt", "yb")}
    for p in range(P):
        r0, rf = rows[p], rows[p + 1]
        Ai = D[r0:rf, r0:rf]
        inv = np.linalg.inv(Ai)
        y = inv @ nrhs_b[r0:rf]
        out["yt"].append(y[:ku]); out["yb"].append(y[rf - r0 - kl:])
        if p < P - 1:
            V = inv @ D[r0:rf, rf:rf + ku]
        else:
            V = np.zeros((rf - r0, ku))
        if p > 0:
            W = inv @ D[r0:rf, r0 - kl:r0]
        else:
            W = np.zeros((rf - r0, kl))
        out["Vt"].append(V[:ku]); out["Vb"].append(V[rf - r0 - kl:])
        out["Wt"].append(W[:ku]); out["Wb"].append(W[rf - r0 - kl:])
    return {k: np.stack(v) for k, v in out.items()}


def seeded_banded(n, ku, kl, seed):
    rng = np.random.RandomState(seed)
    D = np.zeros((n, n))
    for d in range(-kl, ku + 1):
        if d == 0:
            continue
        idx = np.arange(max(0, -d), min(n, n - d))
        D[idx, idx + d] = rng.uniform(-1, 1, idx.size)
    D[np.arange(n), np.arange(n)] = np.abs(D).sum(1) * 1.1 + 0.5
    return sp.csr_matrix(D)


def regen_tridiagonal_sha(n):
    """Seed-314 tridiagonal generator (generator_banded.py:10-26) + binary writer (:70-172)."""
    np.random.seed(314)
    A = sp.diags(np.random.rand(n) + 1.0, 0) + sp.diags(np.random.rand(n - 1), 1) + \
        sp.diags(np.random.rand(n - 1), -1)
    A = A.tocsr()
    blob = np.array([A.shape[0], A.nnz, 0], np.int32).tobytes() + A.data.astype(np.float64).tobytes() + \
        A.indices.astype(np.int32).tobytes() + A.indptr.astype(np.int32).tobytes()
    return hashlib.sha256(blob).hexdigest()


def main():
    g = {}
    rng = np.random.RandomState(20261202)
    real = {
        "tridiag10": "solve_small/tridiagonal.bin", "penta10": "spike/penta_10.bin",
        "penta15": "spike/penta_15.bin", "hepta20": "heptadiagonal/small.bin",
        "hepta40": "heptadiagonal/medium.bin", "banded10": "bandwidth/banded.bin",
        "penta20k": "pentadiagonal/large.bin",
    }
    for name, rel in real.items():
        A = load(os.path.join(HERE, "ref", rel)).tocsc()
        n = A.shape[0]
        lu = spla.splu(A)
        b1 = np.ones((n, 1))
        b3 = rng.uniform(-1, 1, (n, 3))
        g["b_%s_rand" % name] = b3
        g["x_%s_ones" % name] = lu.solve(b1)
        g["x_%s_rand" % name] = lu.solve(b3)
    for name, P in (("hepta40", 4), ("banded10", 2), ("penta15", 3), ("tridiag10", 4), ("penta20k", 8)):
        A = load(os.path.join(HERE, "ref", real[name])).tocsr()
        if A.shape[0] > 2000:
            continue
        t = tips(A, P, g["b_%s_rand" % name])
        for k, v in t.items():
            g["tips_%s_P%d_%s" % (name, P, k)] = v
    for name, rel in (("zpenta15", "complex16/penta_15.z"), ("zpenta1k", "complex16/penta_1k.z")):
        A = load(os.path.join(HERE, "ref", rel)).tocsc()
        n = A.shape[0]
        b = rng.uniform(-1, 1, (n, 2)) + 1j * rng.uniform(-1, 1, (n, 2))
        g["b_%s" % name] = b
        g["x_%s" % name] = spla.splu(A).solve(b)
    for (n, ku, kl, seed) in ((300, 5, 3, 11), (257, 16, 16, 12), (1000, 2, 7, 13)):
        A = seeded_banded(n, ku, kl, seed)
        D = A.toarray()
        ab = np.zeros((ku + kl + 1, n))
        for j in range(n):
            for i in range(max(0, j - ku), min(n, j + kl + 1)):
                ab[ku + i - j, j] = D[i, j]
        b = rng.uniform(-1, 1, (n, 2))
        key = "band_%d_%d_%d" % (n, ku, kl)
        g[key + "_rowptr"] = A.indptr.astype(np.int32)
        g[key + "_colind"] = A.indices.astype(np.int32)
        g[key + "_aij"] = A.data
        g[key + "_b"] = b
        g[key + "_x"] = sla.solve_banded((kl, ku), ab, b)
    g["sha_tridiag_100000"] = np.frombuffer(regen_tridiagonal_sha(100000).encode(), np.uint8)
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **g)
    print("wrote %d arrays" % len(g))


if __name__ == "__main__":
    sys.exit(main())
