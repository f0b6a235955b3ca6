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
    old = amg._HOST_PARALLEL_NNZ, amg._HOST_PROLONG_NNZ, amg._HOST_PARALLEL_FLOPS
    # everything through the library
    amg._HOST_PARALLEL_NNZ = amg._HOST_PROLONG_NNZ = amg._HOST_PARALLEL_FLOPS = 0
    try:
        levels, _ = amg.sa_setup(A, max_coarse=10)
    finally:
        amg._HOST_PARALLEL_NNZ, amg._HOST_PROLONG_NNZ, amg._HOST_PARALLEL_FLOPS = old
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


def _lap1d(n, dirichlet=True):
    """Tridiagonal 1-D Laplacian: Dirichlet ends (SPD) or Neumann (singular,
    constants in the kernel)."""
    main = np.full(n, 2.0)
    if not dirichlet:
        main[0] = main[-1] = 1.0
    return sp.diags([-np.ones(n - 1), main, -np.ones(n - 1)], [-1, 0, 1]).tocsr()


def test_coarse_inverse_spd_is_the_inverse():
    """An SPD coarsest operator goes through its Cholesky factor: the
    inverse (A X = I to rounding), equal to the pseudo-inverse."""
    A = _lap1d(40)
    X = amg.coarse_inverse(A)
    assert np.allclose(A.toarray() @ X, np.eye(40), rtol=0, atol=1e-11)
    assert np.allclose(X, np.linalg.pinv(A.toarray()), rtol=0, atol=1e-10)
    assert np.array_equal(X, X.T)


def test_coarse_inverse_singular_and_nonsymmetric_fall_back_to_pinv():
    """Singular (pure Neumann) and non-symmetric operators keep the
    pseudo-inverse: X A X = X, A X A = A, and X annihilates the kernel."""
    A = _lap1d(30, dirichlet=False)
    X = amg.coarse_inverse(A)
    Ad = A.toarray()
    assert np.array_equal(X, np.linalg.pinv(Ad))
    assert np.allclose(Ad @ X @ Ad, Ad, atol=1e-10)
    assert np.allclose(X @ np.ones(30), 0.0, atol=1e-10)
    N = _lap1d(12).toarray()
    N[0, 1] = -1.5                      # not symmetric
    assert np.array_equal(amg.coarse_inverse(sp.csr_matrix(N)), np.linalg.pinv(N))
    assert amg.coarse_inverse(sp.csr_matrix((0, 0))).shape == (0, 0)


def test_sa_setup_takes_a_given_finest_rho():
    """rho0 replaces only the finest level's power iteration."""
    A = _lap1d(3000)
    lv_a, _ = amg.sa_setup(A, max_coarse=10)
    lv_b, _ = amg.sa_setup(A, max_coarse=10, rho0=lv_a[0].rho)
    assert [l.rho for l in lv_a] == [l.rho for l in lv_b]
    lv_c, _ = amg.sa_setup(A, max_coarse=10, rho0=1.5 * lv_a[0].rho)
    assert lv_c[0].rho == 1.5 * lv_a[0].rho
    assert not np.array_equal(lv_c[0].P.data, lv_a[0].P.data)


@pytest.mark.parametrize("threads", ["1", "5"])
def test_sa_prolongator_bitwise_scipy_recipe(threads, monkeypatch):
    """The library's P = T - diag(s) (A T) and R = P^T equal the scipy
    recipe bit for bit, with isolated rows (aggregate -1, no T entry), rows
    whose damping s is zero, empty aggregates' neighbours and cancelling
    sums, at any thread count."""
    monkeypatch.setenv("FP3D_HOST_THREADS", threads)
    rng = np.random.default_rng(int(threads))
    n = 60000
    A = _rand_csr(rng, n, n, 6.0 / n)
    A = (A + sp.eye(n)).tocsr()
    A.sort_indices()
    n_agg = 7000
    agg = rng.integers(0, n_agg, n).astype(np.int32)
    agg[rng.random(n) < 0.05] = -1                   # isolated rows
    rows = np.nonzero(agg >= 0)[0]
    cnt = np.bincount(agg[rows], minlength=n_agg).astype(np.float64)
    cnt[cnt == 0] = 1.0
    tval = 1.0 / np.sqrt(cnt)
    s = rng.standard_normal(n)
    s[rng.random(n) < 0.05] = 0.0                    # zero damping
    # an exact cancellation t - x = 0 in one row: x = s_i (A T)_{i, agg_i}
    P, R = amg.sa_prolongator(A, agg, n_agg, tval, s)
    T = sp.csr_matrix((tval[agg[rows]], (rows, agg[rows])), shape=(n, n_agg))
    Pr = (T - sp.diags(s) @ (A @ T)).tocsr()
    Pr.sort_indices()
    Rr = Pr.T.tocsr()
    Rr.sort_indices()
    for got, ref in ((P, Pr), (R, Rr)):
        assert np.array_equal(got.indptr, ref.indptr)
        assert np.array_equal(got.indices, ref.indices)
        assert np.array_equal(got.data, ref.data)
    # exact zeros of t - x are dropped: force one and compare again
    i = int(rows[0])
    at = (A @ T)[i, agg[i]]
    if at != 0.0:
        s2 = s.copy()
        s2[i] = tval[agg[i]] / at
        if s2[i] * at == tval[agg[i]]:
            P2, _ = amg.sa_prolongator(A, agg, n_agg, tval, s2)
            Pr2 = (T - sp.diags(s2) @ (A @ T)).tocsr()
            Pr2.sort_indices()
            assert np.array_equal(P2.indices, Pr2.indices)
            assert np.array_equal(P2.data, Pr2.data)
            assert P2.nnz == Pr2.nnz
