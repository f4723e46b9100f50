"""CPU-only checks of the library: it loads, exports every symbol include/*.h declares, and its
host-side logic (CSR IO, validation, schedule, reduced-system pattern, generators, RCM) is right.
No GPU compute is called here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, ref_path


@pytest.fixture(scope="module")
def L():
    import hspike_amd
    return hspike_amd.load()


def declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*[\s\*]+\**([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt, re.M):
            if m.group(1) not in ("if", "while", "return", "sizeof"):
                names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported(L):
    from hspike_amd import _lib
    names = declared_symbols()
    assert len(names) > 50
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(_lib.EXPORTED) <= set(names) | {"spike_last_error_message"}


def test_reference_signatures_kept():
    txt = open(os.path.join(ROOT, "include", "spike_interfaces.h")).read()
    for p in "sdcz":
        assert re.search(r"Error_t %sspike_core_host\(const integer_t n, const integer_t nnz, const integer_t nrhs," % p, txt)
    assert "dspike_core_host_blocking" in txt


def test_load_csr_and_compare(L, tmp_path):
    import hspike_amd as hs
    import oracle
    A = oracle.load_csr(ref_path("split/matrix.bin"))
    M = L.matrix_LoadCSR(ref_path("split/matrix.bin").encode())
    assert M
    out = str(tmp_path / "m.bin")
    assert L.matrix_SaveCSR(out.encode(), M) == 0
    assert open(out, "rb").read() == open(ref_path("split/matrix.bin"), "rb").read()
    # ExtractMatrix == the split fixture blocks (exact, OpenMP/split.c:43-75)
    for (r0, rf), name in (((0, 5), "block1"), ((5, 10), "block2")):
        E = L.matrix_ExtractMatrix(M, r0, rf, r0, rf, 1)
        G = L.matrix_LoadCSR(ref_path("split/%s.bin" % name).encode())
        assert L.matrix_AreEqual(E, G) == 1
        L.matrix_Deallocate(E)
        L.matrix_Deallocate(G)
    L.matrix_Deallocate(M)
    assert A.n == 10
    assert hs.CSRMatrix(A.n, A.rowptr, A.colind, A.aij).bandwidth() == (1, 1)


def test_bandwidth_fixture(L):
    import hspike_amd as hs
    import oracle
    A = oracle.load_csr(ref_path("bandwidth/banded.bin"))
    assert hs.CSRMatrix(A.n, A.rowptr, A.colind, A.aij).bandwidth() == (2, 4)


def test_complex_loader_detects_planar(L):
    class MZ(ctypes.Structure):
        _fields_ = [("n", ctypes.c_int), ("nnz", ctypes.c_uint64), ("ku", ctypes.c_int), ("kl", ctypes.c_int),
                    ("type", ctypes.c_int), ("colind", ctypes.c_void_p), ("rowptr", ctypes.c_void_p),
                    ("aij", ctypes.POINTER(ctypes.c_double)), ("planar_file", ctypes.c_int)]
    import oracle
    for f in ("complex16/penta_15.z", "complex16/penta_1k.z"):
        p = L.matrix_LoadCSR_z(ref_path(f).encode())
        assert p
        M = ctypes.cast(p, ctypes.POINTER(MZ)).contents
        assert M.planar_file == 1
        ref = oracle.load_csr(ref_path(f))
        vals = np.ctypeslib.as_array(M.aij, (2 * M.nnz,)).copy()
        np.testing.assert_array_equal(vals[0::2] + 1j * vals[1::2], ref.aij)
        L.matrix_z_Deallocate(p)


def test_malformed_csr_rejected(L):
    rp = np.array([0, 2, 1, 3], np.int32)         # not monotone
    ci = np.array([0, 1, 2], np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert L.matrix_CheckCSR(3, 3, p(ci), p(rp)) == -2
    rp = np.array([0, 1, 2, 3], np.int32)
    ci = np.array([0, 5, 2], np.int32)            # column out of range
    assert L.matrix_CheckCSR(3, 3, p(ci), p(rp)) == -2
    ci = np.array([0, 1, 2], np.int32)
    assert L.matrix_CheckCSR(3, 3, p(ci), p(rp)) == 0
    # the reference aborts on nnz < 0 (spike_interfaces.c:53-56); here: an error code, no abort
    x = np.zeros(3)
    assert L.dspike_core_host(3, -1, 1, p(ci), p(rp), p(x), p(x), p(x)) == -1
    assert L.dspike_core_host(3, 3, 1, p(np.array([0, 1, 7], np.int32)), p(rp), p(x), p(x), p(x)) == -2


def test_reduced_pattern_matches_oracle(L):
    """matrix_CreateEmptyReducedSystem + AssembleReducedSystem vs the oracle's R (spike_matrix.c:657-740)."""
    import oracle
    rng = np.random.RandomState(3)
    import scipy.sparse as sp
    n, ku, kl, P = 300, 4, 3, 5
    D = np.zeros((n, n))
    for d in range(-kl, ku + 1):
        idx = np.arange(max(0, -d), min(n, n - d))
        D[idx, idx + d] = rng.uniform(-1, 1, idx.size)
    D[np.arange(n), np.arange(n)] += np.abs(D).sum(1) * 1.1
    S = sp.csr_matrix(D)
    A = oracle.CSR(n, S.indptr, S.indices, S.data)
    _, info = oracle.spike_solve(A, np.ones((n, 2)), P)
    kua = np.full(P, ku, np.int32)
    kla = np.full(P, kl, np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    R = L.matrix_CreateEmptyReducedSystem(P, p(kua), p(kla))
    cat = {k: np.ascontiguousarray(info[k].transpose(0, 2, 1)).ravel() for k in ("Vt", "Vb", "Wt", "Wb", "yt", "yb")}
    assert L.matrix_AssembleReducedSystem(R, P, ku, kl, p(cat["Vt"]), p(cat["Vb"]), p(cat["Wt"]), p(cat["Wb"])) == 0

    class M(ctypes.Structure):
        _fields_ = [("n", ctypes.c_int), ("nnz", ctypes.c_uint64), ("ku", ctypes.c_int), ("kl", ctypes.c_int),
                    ("type", ctypes.c_int), ("colind", ctypes.POINTER(ctypes.c_int)),
                    ("rowptr", ctypes.POINTER(ctypes.c_int)), ("aij", ctypes.POINTER(ctypes.c_double))]
    m = ctypes.cast(R, ctypes.POINTER(M)).contents
    rp = np.ctypeslib.as_array(m.rowptr, (m.n + 1,))
    ci = np.ctypeslib.as_array(m.colind, (m.nnz,))
    av = np.ctypeslib.as_array(m.aij, (m.nnz,))
    Ro = info["R"]
    np.testing.assert_array_equal(rp, Ro.rowptr)
    np.testing.assert_array_equal(ci, Ro.colind)
    np.testing.assert_allclose(av, Ro.aij, rtol=0, atol=0)
    g = np.zeros(P * (ku + kl) * 2)
    assert L.block_AssembleReducedRHS(p(g), P, ku, kl, 2, p(cat["yt"]), p(cat["yb"])) == 0
    np.testing.assert_array_equal(g.reshape(2, -1).T, info["g"])
    L.matrix_Deallocate(R)


def test_generator_properties():
    import hspike_amd as hs
    A = hs.gen_banded(5000, 7, 5, fill=0.3, dom=(1.01, 2.0), diag_floor=1.0, seed=9)
    assert A.bandwidth() == (7, 5)
    S = A.to_scipy()
    d = np.abs(S.diagonal())
    off = np.asarray(abs(S).sum(1)).ravel() - d
    scaled = d > 1.0                                     # rows not lifted to the diagonal floor
    ratio = d[scaled] / off[scaled]
    assert ratio.min() >= 1.01 - 1e-12 and ratio.max() <= 2.0 + 1e-12
    assert (d[~scaled] == 1.0).all() and (off[~scaled] * 1.01 <= 1.0 + 1e-12).all()  # floor only when needed
    assert (d[off == 0] == 1.0).all()                 # diagonal floor for rows without off-diagonals
    frac = (A.nnz - 5000) / (5000 * 12.0)
    assert 0.27 < frac < 0.33                           # in-band fill
    B = hs.gen_banded(5000, 7, 5, fill=0.3, dom=(1.01, 2.0), diag_floor=1.0, seed=9)
    np.testing.assert_array_equal(A.aij, B.aij)         # deterministic
    Z = hs.gen_banded(300, 3, 3, fill=1.0, dom=(1.2, 1.2), seed=4, complex=True)
    Sz = Z.to_scipy()
    dz = np.abs(Sz.diagonal())
    np.testing.assert_allclose(dz, (np.asarray(abs(Sz).sum(1)).ravel() - dz) * 1.2)
    X = hs.gen_dense(100, 3, lo=-1, hi=1, seed=1)
    assert X.shape == (100, 3) and X.min() >= -1 and X.max() <= 1


def test_rcm_recovers_band():
    import hspike_amd as hs
    A = hs.gen_banded(3000, 4, 4, fill=0.6, dom=(1.1, 1.1), seed=2)
    perm = np.random.RandomState(1).permutation(A.n).astype(np.int32)
    Ap = hs.permute(A, perm)
    assert max(Ap.bandwidth()) > 1000
    order, ku, kl = hs.rcm(Ap)
    Ar = hs.permute(Ap, order)
    assert Ar.bandwidth() == (ku, kl)
    assert max(ku, kl) <= 16
    assert sorted(order.tolist()) == list(range(A.n))
    # P A P^T preserves the matrix: same entries, renumbered
    Ds = Ap.to_scipy().toarray()
    Dr = Ar.to_scipy().toarray()
    np.testing.assert_array_equal(Dr, Ds[np.ix_(order, order)])


def test_autotuner_is_pure():
    import hspike_amd as hs
    a = hs.autotune_partitions(1_000_000, 64, 64, 16)
    assert a == hs.autotune_partitions(1_000_000, 64, 64, 16)
    assert 32 <= a <= 8192
    assert hs.autotune_partitions(1000, 2, 2, 1) >= 1
    assert hs.autotune_partitions(100, 60, 60, 1) == 1
