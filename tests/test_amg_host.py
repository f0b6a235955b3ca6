"""Host side of the SA-AMG set-up (no GPU): the library's row-parallel
sparse product and matvec reproduce scipy.sparse's values bit for bit
(same per-row accumulation order, exact zeros dropped), independently of
the thread count, so the hierarchy equals the scipy recipe it replaces."""

import os

import numpy as np
import pytest
import scipy.sparse as sp

from up3d_amd.solve import amg


def _rand_csr(rng, m, n, density):
    M = sp.random(m, n, density=density, random_state=rng, format="csr")
    M.data = rng.standard_normal(M.nnz)
    M.sort_indices()
    return M


def _canon(M):
    M = M.tocsr().copy()
    M.sort_indices()
    return M


@pytest.mark.parametrize("threads", ["1", "3", "8"])
def test_spgemm_bitwise_scipy(threads, monkeypatch):
    monkeypatch.setenv("FP3D_HOST_THREADS", threads)
    rng = np.random.default_rng(7)
    for m, k, n, dens in ((1, 1, 1, 1.0), (50, 40, 30, 0.2), (200000, 3000, 500, 0.002)):
        A = _rand_csr(rng, m, k, dens)
        B = _rand_csr(rng, k, n, dens)
        C = amg.spgemm(A, B, force_native=True)
        R = _canon(A @ B)
        assert np.array_equal(C.indptr, R.indptr)
        assert np.array_equal(C.indices, R.indices)
        assert np.array_equal(C.data, R.data)


def test_spgemm_drops_exact_cancellation():
    # row 0 of A B is exactly zero in column 0: scipy drops it, so must we
    A = sp.csr_matrix(np.array([[1.0, 1.0], [2.0, 0.0]]))
    B = sp.csr_matrix(np.array([[1.0, 3.0], [-1.0, 0.0]]))
    C = amg.spgemm(A, B, force_native=True)
    R = _canon(A @ B)
    assert C.nnz == R.nnz == 3
    assert np.array_equal(C.toarray(), (A @ B).toarray())
    with pytest.raises(ValueError):
        amg.spgemm(A, sp.csr_matrix(np.ones((3, 2))))


def test_csr_matvec_bitwise_scipy(monkeypatch):
    monkeypatch.setenv("FP3D_HOST_THREADS", "4")
    rng = np.random.default_rng(3)
    A = _rand_csr(rng, 100000, 100000, 0.0002)   # 2M entries: 7 threads
    x = rng.standard_normal(100000)
    assert np.array_equal(amg.csr_matvec(A, x, force_native=True), A @ x)


def test_sa_setup_matches_scipy_recipe():
    from .helpers import jittered_cube

    from up3d_amd.mesh.metrics import precompute_element_geometry

    nodes, el = jittered_cube(10, 0.2, seed=5)
    B, V = precompute_element_geometry(nodes, el)
    n = len(nodes)
    rows, cols, vals = [], [], []
    for e in range(len(el)):
        for a in range(4):
            for b in range(4):
                rows.append(el[e, a])
                cols.append(el[e, b])
                vals.append(V[e] * float(B[e, a] @ B[e, b]))
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    old = amg._HOST_PARALLEL_NNZ
    amg._HOST_PARALLEL_NNZ = 0          # every product through the library
    try:
        levels, _ = amg.sa_setup(A, max_coarse=10)
    finally:
        amg._HOST_PARALLEL_NNZ = old
    # the scipy recipe, level by level
    Ar = A
    for lv in levels[:-1]:
        d = Ar.diagonal()
        dinv = np.where(d != 0.0, 1.0 / np.where(d != 0.0, d, 1.0), 0.0)
        agg, na = amg.standard_aggregation(Ar)
        r = np.nonzero(agg >= 0)[0]
        cnt = np.bincount(agg[r], minlength=na).astype(np.float64)
        T = sp.csr_matrix((1.0 / np.sqrt(cnt[agg[r]]), (r, agg[r])), shape=(Ar.shape[0], na))
        P = (T - sp.diags((4.0 / 3.0 / lv.rho) * dinv) @ (Ar @ T)).tocsr()
        P.sort_indices()
        Ac = _canon(P.T.tocsr() @ (Ar @ P))
        assert np.array_equal(_canon(lv.P).data, P.data)
        Ar = Ac
    assert np.array_equal(levels[-1].A.data, Ar.data)
    assert len(levels) >= 2


@pytest.mark.parametrize("shape_a,shape_b", [((0, 5), (5, 3)), ((4, 5), (5, 0)),
                                             ((3, 0), (0, 2))])
def test_spgemm_empty_shapes(shape_a, shape_b):
    A = sp.csr_matrix(shape_a)
    B = sp.csr_matrix(shape_b)
    C = amg.spgemm(A, B, force_native=True)
    assert C.shape == (shape_a[0], shape_b[1]) and C.nnz == 0
    assert np.array_equal(amg.csr_matvec(A, np.ones(shape_a[1]), force_native=True),
                          np.zeros(shape_a[0]))


def test_spgemm_all_cancelling_rows():
    # every product entry cancels exactly: no entries, like scipy
    A = sp.csr_matrix(np.array([[1.0, -1.0], [2.0, -2.0]]))
    B = sp.csr_matrix(np.array([[3.0, 1.0], [3.0, 1.0]]))
    C = amg.spgemm(A, B, force_native=True)
    assert C.nnz == (A @ B).nnz == 0
