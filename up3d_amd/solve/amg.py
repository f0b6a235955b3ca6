"""Smoothed-aggregation AMG: hierarchy set up on the host, V-cycle applied
on the device.

The reference builds `pyamg.smoothed_aggregation_solver(A)` with pyamg's
defaults and a pinned RNG seed (pyfp3d/solve/linear.py:61-81): symmetric
strength (theta = 0), standard aggregation, constant near-nullspace,
tentative prolongator smoothed by damped Jacobi with
omega = (4/3) / rho(D^-1 A), Galerkin coarse operators R A P with
R = P^T, at most 10 levels, coarsest size <= 10 solved by a pseudo-inverse.
This module follows that recipe for the set-up (aggregation and the
sparse products in the library's row-parallel host code, which reproduces
scipy.sparse's values) and replaces pyamg's
sequential Gauss-Seidel smoother by Chebyshev polynomial smoothing on
D^-1 A, which is fully parallel on the GPU and keeps the V-cycle symmetric
(same polynomial before and after the coarse correction), so it remains a
valid CG preconditioner.  The coarsest level is applied as a dense
float64 inverse matvec (through a Cholesky factor for SPD operators, the
pseudo-inverse otherwise).  Matrix values can be held in float32
(precision="f32"); vectors and accumulation stay float64.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional

import numpy as np
import scipy.sparse as sp
import torch

from .. import _native
from ..device import empty, ptr, stream_handle, to_dev


def standard_aggregation(A: sp.csr_matrix):
    """(agg int32[n], n_agg): pyamg-style three-pass standard aggregation on
    the nonzero off-diagonal pattern of A (host C++); -1 = isolated node."""
    A = A.tocsr()
    ip = np.ascontiguousarray(A.indptr, dtype=np.int32)
    ix = np.ascontiguousarray(A.indices, dtype=np.int32)
    dv = np.ascontiguousarray(A.data, dtype=np.float64)
    agg = np.empty(A.shape[0], dtype=np.int32)
    na = C.c_int64(0)
    _native.check(_native.lib().fp3d_host_standard_aggregation(
        A.shape[0], ip.ctypes.data, ix.ctypes.data, dv.ctypes.data,
        agg.ctypes.data, C.byref(na)), "standard_aggregation")
    return agg, int(na.value)


def _csr32(M: sp.csr_matrix):
    return (np.ascontiguousarray(M.indptr, dtype=np.int32),
            np.ascontiguousarray(M.indices, dtype=np.int32),
            np.ascontiguousarray(M.data, dtype=np.float64))


# below this many stored entries of A the single-pass scipy matvec is the
# faster of the two (the values are the same)
_HOST_PARALLEL_NNZ = 1 << 21
# sparse products with fewer estimated multiply-adds (A's entries times the
# mean length of B's rows) stay in scipy; above, the library's threaded
# product is faster (the same values)
_HOST_PARALLEL_FLOPS = 1 << 20
# the smoothed prolongator and its transpose go through the library above
# this many entries of A (one pass instead of scipy's chain of products,
# scalings, differences and transposes; the same values)
_HOST_PROLONG_NNZ = 1 << 14


def spgemm(A: sp.csr_matrix, B: sp.csr_matrix, force_native: bool = False
           ) -> sp.csr_matrix:
    """A @ B with the library's row-parallel host product: the values of
    scipy's product (same per-row accumulation order, exact zeros dropped),
    columns sorted.  Small products go to scipy itself."""
    A, B = A.tocsr(), B.tocsr()
    if A.shape[1] != B.shape[0]:
        raise ValueError(f"spgemm: shapes {A.shape} and {B.shape} do not align")
    flops = A.nnz * max(1.0, B.nnz / max(B.shape[0], 1))
    if flops < _HOST_PARALLEL_FLOPS and not force_native:
        out = (A @ B).tocsr()
        out.sort_indices()
        return out
    ap, aj, ax = _csr32(A)
    bp, bj, bx = _csr32(B)
    args = (A.shape[0], B.shape[1], ap.ctypes.data, aj.ctypes.data, ax.ctypes.data,
            bp.ctypes.data, bj.ctypes.data, bx.ctypes.data)
    L = _native.lib()
    ip = np.empty(A.shape[0] + 1, dtype=np.int32)
    nnz = C.c_int64(0)
    _native.check(L.fp3d_host_spgemm(*args, ip.ctypes.data, C.byref(nnz)),
                  "host_spgemm")
    ix = np.empty(int(nnz.value), dtype=np.int32)
    dv = np.empty(int(nnz.value), dtype=np.float64)
    _native.check(L.fp3d_host_spgemm_fill(*args, ip.ctypes.data, ix.ctypes.data,
                                          dv.ctypes.data), "host_spgemm_fill")
    out = sp.csr_matrix((dv, ix, ip), shape=(A.shape[0], B.shape[1]))
    out.has_sorted_indices = True
    return out


def sa_prolongator(A: sp.csr_matrix, agg: np.ndarray, n_agg: int, tval: np.ndarray,
                   s: np.ndarray):
    """(P, R = P^T) of smoothed aggregation, P = T - diag(s) (A T) with T the
    tentative prolongator (row j: tval[agg[j]] in column agg[j]), by the
    library's row-parallel host code: the values of the scipy recipe
    (T - sp.diags(s) @ (A @ T), transposed), columns sorted."""
    ap, aj, ax = _csr32(A)
    n = A.shape[0]
    ag = np.ascontiguousarray(agg, dtype=np.int32)
    tv = np.ascontiguousarray(tval, dtype=np.float64)
    sv = np.ascontiguousarray(s, dtype=np.float64)
    L = _native.lib()
    args = (n, ap.ctypes.data, aj.ctypes.data, ax.ctypes.data, ag.ctypes.data, int(n_agg),
            tv.ctypes.data, sv.ctypes.data)
    pp = np.empty(n + 1, dtype=np.int32)
    nnz = C.c_int64(0)
    _native.check(L.fp3d_host_sa_prolongator(*args, pp.ctypes.data, C.byref(nnz)),
                  "host_sa_prolongator")
    pj = np.empty(int(nnz.value), dtype=np.int32)
    px = np.empty(int(nnz.value), dtype=np.float64)
    _native.check(L.fp3d_host_sa_prolongator_fill(*args, pp.ctypes.data, pj.ctypes.data,
                                                  px.ctypes.data), "host_sa_prolongator_fill")
    P = sp.csr_matrix((px, pj, pp), shape=(n, int(n_agg)))
    P.has_sorted_indices = True
    rp = np.empty(int(n_agg) + 1, dtype=np.int32)
    rj = np.empty(len(pj), dtype=np.int32)
    rx = np.empty(len(pj), dtype=np.float64)
    _native.check(L.fp3d_host_csr_transpose(n, int(n_agg), pp.ctypes.data, pj.ctypes.data,
                                            px.ctypes.data, rp.ctypes.data, rj.ctypes.data,
                                            rx.ctypes.data), "host_csr_transpose")
    R = sp.csr_matrix((rx, rj, rp), shape=(int(n_agg), n))
    R.has_sorted_indices = True
    return P, R


def csr_matvec(A: sp.csr_matrix, x: np.ndarray, force_native: bool = False
               ) -> np.ndarray:
    """A @ x with the library's row-parallel host matvec (scipy's values);
    small matrices go to scipy itself."""
    if A.nnz < _HOST_PARALLEL_NNZ and not force_native:
        return A @ x
    ap, aj, ax = _csr32(A)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(A.shape[0], dtype=np.float64)
    _native.check(_native.lib().fp3d_host_csr_matvec(
        A.shape[0], ap.ctypes.data, aj.ctypes.data, ax.ctypes.data, x.ctypes.data,
        y.ctypes.data), "host_csr_matvec")
    return y


def spectral_radius_dinv_a(A: sp.csr_matrix, dinv: np.ndarray, iters: int = 12,
                           seed: int = 0) -> float:
    """Power-iteration estimate of rho(D^-1 A) from a seeded start vector."""
    rng = np.random.default_rng(seed)
    x = rng.random(A.shape[0]) + 0.5
    lam = 1.0
    for _ in range(iters):
        y = dinv * csr_matvec(A, x)
        ny = _norm2(y)
        if ny == 0.0:
            return 1.0
        lam = ny / _norm2(x)
        x = y / ny
    return float(lam)


def _norm2(v: np.ndarray) -> float:
    """2-norm by numpy's pairwise summation: one thread and a fixed order
    (a BLAS dot splits the sum over its thread pool, whose wake-ups cost
    about a millisecond per call on short vectors)."""
    return float(np.sqrt(np.add.reduce(v * v)))


@dataclass
class HostLevel:
    A: sp.csr_matrix
    P: Optional[sp.csr_matrix]
    R: Optional[sp.csr_matrix]
    dinv: np.ndarray
    rho: float


def spectral_radius_dinv_a_dev(n: int, nnz: int, indptr: torch.Tensor,
                               indices: torch.Tensor, data: torch.Tensor,
                               dinv: torch.Tensor, iters: int = 12, seed: int = 0
                               ) -> float:
    """spectral_radius_dinv_a on a float64 device CSR (int32 indices): the
    same seeded start vector and iteration, with device SpMVs and
    deterministic device norms; one host read at the end."""
    rng = np.random.default_rng(seed)
    x = to_dev(rng.random(n) + 0.5)
    y = empty(n)
    work = empty(int(_native.lib().fp3d_reduce_work_doubles()))
    norms = empty((iters, 2, 2))           # per iteration: (||y||, max), (||x||, max)
    s = stream_handle()
    for it in range(iters):
        _native.call("fp3d_csr_spmv", n, nnz, ptr(indptr), ptr(indices), ptr(data),
                     ptr(x), ptr(y), s)
        y.mul_(dinv)
        _native.call("fp3d_norms", n, ptr(y), ptr(work), ptr(norms[it, 0]), s)
        _native.call("fp3d_norms", n, ptr(x), ptr(work), ptr(norms[it, 1]), s)
        _native.call("fp3d_scale_inv_dev", n, ptr(norms[it, 0]), ptr(y), ptr(x), s)
    h = norms.cpu().numpy()
    if np.any(h[:, 0, 0] == 0.0):
        return 1.0
    return float(h[-1, 0, 0] / h[-1, 1, 0])


def coarse_inverse(Ac: sp.csr_matrix) -> np.ndarray:
    """Dense inverse of the coarsest operator: through a Cholesky factor when
    the matrix is symmetric to rounding and positive definite with a
    well-scaled factor (diagonal ratio above 1e-6), else the pseudo-inverse
    (singular or indefinite operators, e.g. pure Neumann problems)."""
    M = Ac.toarray()
    n = M.shape[0]
    if n:
        scale = np.abs(M).max()
        if scale > 0.0 and np.abs(M - M.T).max() <= 1e-12 * scale:
            from scipy.linalg import lapack

            # LAPACK potrf (lower factor) then potri (the inverse from it)
            L, info = lapack.dpotrf(0.5 * (M + M.T), lower=1, clean=1)
            if info == 0:
                dl = np.diag(L)
                if dl.min() > 1e-6 * dl.max():
                    X, info = lapack.dpotri(L, lower=1)
                    if info == 0:
                        X = np.tril(X)
                        return X + np.tril(X, -1).T
    return np.linalg.pinv(M)


def sa_setup(A: sp.csr_matrix, max_levels: int = 10, max_coarse: int = 10,
             omega: float = 4.0 / 3.0, seed: int = 0, rho0: Optional[float] = None):
    """Host SA hierarchy: list of HostLevel plus the dense coarse inverse.
    rho0: rho(D^-1 A) of the finest level when already known (estimated on
    the device from the same seeded start vector)."""
    A = A.tocsr()
    levels: List[HostLevel] = []
    while True:
        n = A.shape[0]
        d = A.diagonal()
        dinv = np.where(d != 0.0, 1.0 / np.where(d != 0.0, d, 1.0), 0.0)
        if rho0 is not None and not levels:
            rho = float(rho0)
        else:
            rho = spectral_radius_dinv_a(A, dinv, seed=seed)
        if n <= max_coarse or len(levels) >= max_levels - 1:
            levels.append(HostLevel(A, None, None, dinv, rho))
            break
        agg, n_agg = standard_aggregation(A)
        if n_agg == 0 or n_agg >= n:
            levels.append(HostLevel(A, None, None, dinv, rho))
            break
        rows = np.nonzero(agg >= 0)[0]
        cnt = np.bincount(agg[rows], minlength=n_agg).astype(np.float64)
        if A.nnz >= _HOST_PROLONG_NNZ:
            P, R = sa_prolongator(A, agg, n_agg, 1.0 / np.sqrt(cnt), (omega / rho) * dinv)
        else:
            T = sp.csr_matrix((1.0 / np.sqrt(cnt[agg[rows]]), (rows, agg[rows])),
                              shape=(n, n_agg))
            AT = spgemm(A, T)
            P = (T - sp.diags((omega / rho) * dinv) @ AT).tocsr()
            P.sort_indices()
            R = P.T.tocsr()
            R.sort_indices()
        Ac = spgemm(R, spgemm(A, P))
        levels.append(HostLevel(A, P, R, dinv, rho))
        A = Ac
    return levels, coarse_inverse(levels[-1].A)


class DeviceCSR:
    """CSR matrix resident in HBM (int32 indices, float64 and/or float32
    values; with compact_index a square matrix whose bandwidth fits also
    keeps its columns as int16 offsets from the row, which the V-cycle
    kernels read instead)."""

    def __init__(self, M: sp.csr_matrix, f32: bool = False, keep_f64: bool = True,
                 compact_index: bool = False):
        M = M.tocsr()
        self.shape = M.shape
        self.n = M.shape[0]
        self.nnz = int(M.nnz)
        self.indptr = to_dev(M.indptr.astype(np.int32), torch.int32)
        self.indices = to_dev(M.indices.astype(np.int32), torch.int32)
        self.data = to_dev(M.data.astype(np.float64), torch.float64) \
            if (keep_f64 or not f32) else None
        self.data32 = to_dev(M.data.astype(np.float32), torch.float32) if f32 else None
        self.is_f32 = 1 if f32 else 0
        self.indices16 = None
        if compact_index and M.shape[0] == M.shape[1] and self.nnz:
            from ..kernels.store import csr_offsets16

            self.indices16 = csr_offsets16(M.indptr, M.indices)

    @classmethod
    def on_device(cls, n: int, indptr: torch.Tensor, indices: torch.Tensor,
                  data: torch.Tensor, f32: bool = False, keep_f64: bool = True,
                  indices16: Optional[torch.Tensor] = None) -> "DeviceCSR":
        """A square CSR already in HBM (the pattern tensors are shared, the
        float64 values copied), equal to constructing it from its host copy
        with compact_index when indices16 is the pattern's int16 offsets."""
        M = cls.__new__(cls)
        M.shape = (n, n)
        M.n = n
        M.nnz = int(indices.numel())
        M.indptr = indptr
        M.indices = indices
        M.data = data.clone() if (keep_f64 or not f32) else None
        M.data32 = data.to(torch.float32) if f32 else None
        M.is_f32 = 1 if f32 else 0
        M.indices16 = indices16
        return M

    @property
    def values(self):
        return self.data32 if self.is_f32 else self.data

    @property
    def fmt(self) -> int:
        """Matrix format bits of the V-cycle kernels (1 float32 values,
        4 int16 row-relative columns)."""
        return self.is_f32 | (4 if self.indices16 is not None else 0)

    @property
    def index_array(self):
        return self.indices16 if self.indices16 is not None else self.indices

    def matvec(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        if self.is_f32:
            _native.call("fp3d_csr_spmv_f32", self.n, self.nnz, ptr(self.indptr),
                         ptr(self.indices), ptr(self.data32), ptr(x), ptr(y),
                         stream_handle())
        else:
            _native.call("fp3d_csr_spmv", self.n, self.nnz, ptr(self.indptr),
                         ptr(self.indices), ptr(self.data), ptr(x), ptr(y),
                         stream_handle())
        return y


class _DevLevel:
    def __init__(self, h: HostLevel, f32: bool, degree: int, eig_ratio: float,
                 v32: bool = False, compact_index: bool = True,
                 A_dev: Optional[DeviceCSR] = None):
        self.n = h.A.shape[0]
        self.A = A_dev if A_dev is not None else \
            DeviceCSR(h.A, f32=f32, keep_f64=not f32, compact_index=compact_index)
        self.P = DeviceCSR(h.P, f32=f32, keep_f64=not f32) if h.P is not None else None
        self.R = DeviceCSR(h.R, f32=f32, keep_f64=not f32) if h.R is not None else None
        vt = torch.float32 if v32 else torch.float64
        self.dinv = to_dev(h.dinv, torch.float64).to(vt)
        self.rho = h.rho
        self.lmax = 1.1 * h.rho
        self.lmin = self.lmax / eig_ratio
        self.degree = degree
        n = self.n
        self.x = empty(n, vt)
        self.b = empty(n, vt)
        self.r = empty(n, vt)
        self.d0 = empty(n, vt)
        self.d1 = empty(n, vt)


# side stream for V-cycle graph capture (default priority), made on first use
_CAPTURE_STREAM = None


class AMGHierarchy:
    """Device V-cycle of a host-built SA hierarchy (module docstring).

    `precision` "f64" or "f32" selects the matrix value type of every level
    and, by default, the type of the level vectors (`vectors`); the
    arithmetic of every step and the coarsest dense solve are float64 in
    both.  compact_index: square levels whose bandwidth fits keep int16
    row-relative columns for the V-cycle (the same sums, half the index
    bytes)."""

    def __init__(self, A: sp.csr_matrix, precision: str = "f64",
                 degree: int = 2, eig_ratio: float = 20.0, max_levels: int = 10,
                 max_coarse: int = 10, seed: int = 0, use_graph: bool = True,
                 vectors: Optional[str] = None, compact_index: bool = True,
                 device_pattern=None, device_values: Optional[torch.Tensor] = None,
                 rho0: Optional[float] = None, host_setup=None):
        """device_pattern / device_values: A's CSR pattern (an object with
        n_nodes, nnz, indptr, indices and optionally indices16 device
        tensors) and float64 values already in HBM; the finest level is then
        built from them (no upload) and its rho(D^-1 A) estimated there.
        rho0: the finest level's rho(D^-1 A), when already known.
        host_setup: the (levels, coarse inverse) of sa_setup for these
        arguments, when already computed (AMGPlan runs it apart)."""
        if precision not in ("f64", "f32"):
            raise ValueError(f"unknown AMG precision {precision!r}")
        vectors = precision if vectors is None else vectors
        if vectors not in ("f64", "f32"):
            raise ValueError(f"unknown AMG vector precision {vectors!r}")
        f32 = precision == "f32"
        A0 = None
        if device_values is not None:
            st = device_pattern
            if rho0 is None:
                d = A.diagonal()
                dinv = to_dev(np.where(d != 0.0, 1.0 / np.where(d != 0.0, d, 1.0), 0.0))
                rho0 = spectral_radius_dinv_a_dev(st.n_nodes, st.nnz, st.indptr,
                                                  st.indices, device_values, dinv,
                                                  seed=seed)
            i16 = getattr(st, "indices16", None) if compact_index else None
            A0 = DeviceCSR.on_device(st.n_nodes, st.indptr, st.indices, device_values,
                                     f32=f32, keep_f64=not f32, indices16=i16)
        if host_setup is None:
            host_setup = sa_setup(A, max_levels=max_levels, max_coarse=max_coarse,
                                  seed=seed, rho0=rho0)
        host, coarse = host_setup
        self.precision = precision
        self.v32 = vectors == "f32"
        self.vfmt = 2 if self.v32 else 0
        self.levels = [_DevLevel(h, f32, degree, eig_ratio, self.v32, compact_index,
                                 A_dev=A0 if i == 0 else None)
                       for i, h in enumerate(host)]
        self.coarse = to_dev(np.ascontiguousarray(coarse), torch.float64)
        self.shapes = [lv.n for lv in self.levels]
        self.nnzs = [lv.A.nnz for lv in self.levels]
        # HIP graph of the whole V-cycle on fixed buffers (captured lazily on
        # the first apply, after one eager V-cycle has loaded the kernels)
        self.use_graph = use_graph
        self._graph = None
        self._graphs = {}
        self._exec_graphs = {}
        self._gb = empty(self.levels[0].n)
        self._gx = empty(self.levels[0].n)
        if self.v32:
            # float32 copies of the float64 right-hand side and solution
            self._b32 = empty(self.levels[0].n, torch.float32)
            self._x32 = empty(self.levels[0].n, torch.float32)
        self._eager_calls = 0

    @property
    def operator_complexity(self) -> float:
        return float(sum(self.nnzs)) / float(self.nnzs[0])

    def _vcycle(self, l: int, b: torch.Tensor, x: torch.Tensor,
                started: bool = False) -> None:
        """V-cycle from level l on b into x (x = 0 initially; both of the
        level vector type); started: the restriction into this level already
        took the smoother's first step."""
        L = self.levels[l]
        s = stream_handle()
        vf = self.vfmt
        if l == len(self.levels) - 1:
            _native.call("fp3d_dense_matvec_v", vf, L.n, ptr(self.coarse), ptr(b),
                         ptr(x), s)
            return
        A = L.A
        _native.call("fp3d_amg_cheb_smooth", L.n, A.nnz, ptr(A.indptr),
                     ptr(A.index_array), ptr(A.values), A.fmt | vf, ptr(L.dinv),
                     ptr(b), ptr(x), ptr(L.r), ptr(L.d0), ptr(L.d1), L.degree,
                     L.lmin, L.lmax, 2 if started else 1, s)
        _native.call("fp3d_amg_residual", L.n, A.nnz, ptr(A.indptr),
                     ptr(A.index_array), ptr(A.values), A.fmt | vf, ptr(b), ptr(x),
                     ptr(L.r), s)
        Lc = self.levels[l + 1]
        R = L.R
        coarse_next = l + 1 == len(self.levels) - 1
        if coarse_next:
            _native.call("fp3d_amg_restrict", R.n, R.nnz, ptr(R.indptr),
                         ptr(R.index_array), ptr(R.values), R.fmt | vf, ptr(L.r),
                         ptr(Lc.b), s)
        else:
            # the restriction's epilogue also takes the coarse smoother's
            # first step from x = 0 (one launch fewer per level)
            _native.call("fp3d_amg_restrict_start", R.n, R.nnz, ptr(R.indptr),
                         ptr(R.index_array), ptr(R.values), R.fmt | vf, ptr(L.r),
                         ptr(Lc.b), ptr(Lc.dinv), Lc.degree, Lc.lmin, Lc.lmax,
                         ptr(Lc.x), ptr(Lc.r), ptr(Lc.d0), s)
        self._vcycle(l + 1, Lc.b, Lc.x, started=not coarse_next)
        P = L.P
        _native.call("fp3d_amg_prolong_add", P.n, P.nnz, ptr(P.indptr),
                     ptr(P.index_array), ptr(P.values), P.fmt | vf, ptr(Lc.x), ptr(x),
                     s)
        _native.call("fp3d_amg_cheb_smooth", L.n, A.nnz, ptr(A.indptr),
                     ptr(A.index_array), ptr(A.values), A.fmt | vf, ptr(L.dinv),
                     ptr(b), ptr(x), ptr(L.r), ptr(L.d0), ptr(L.d1), L.degree,
                     L.lmin, L.lmax, 0, s)

    def _vcycle_top(self, b: torch.Tensor, x: torch.Tensor) -> None:
        """One V-cycle on the float64 b into the float64 x (through the
        float32 copies when the level vectors are float32)."""
        if not self.v32:
            self._vcycle(0, b, x)
            return
        n = self.levels[0].n
        s = stream_handle()
        _native.call("fp3d_scale_to_f32", n, 1.0, ptr(b), ptr(self._b32), s)
        self._vcycle(0, self._b32, self._x32)
        _native.call("fp3d_f32_to_f64", n, ptr(self._x32), ptr(x), s)

    def apply_eager(self, b: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
        """x = M b with one launch per kernel (no graph)."""
        self._vcycle_top(b, x)
        return x

    def _capture(self, b: torch.Tensor, x: torch.Tensor):
        """The V-cycle from b into x captured as a graph, on a side stream of
        default priority (capture needs a non-default stream).  Not through
        torch.cuda.graph, whose set-up collects Python garbage and empties
        the allocator cache before every capture (about a millisecond each,
        hundreds of captures in a sweep): the V-cycle allocates nothing."""
        global _CAPTURE_STREAM
        if _CAPTURE_STREAM is None:
            _CAPTURE_STREAM = torch.cuda.Stream()
        side = _CAPTURE_STREAM
        cur = torch.cuda.current_stream()
        side.wait_stream(cur)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(side):
            g.capture_begin()
            try:
                self._vcycle_top(b, x)
            finally:
                g.capture_end()
        cur.wait_stream(side)
        return g

    @staticmethod
    def _launch(g) -> None:
        """Replay a captured V-cycle: one hipGraphLaunch on the current
        stream through the library (the V-cycle draws no random numbers, so
        torch's replay bookkeeping has nothing to do)."""
        h = g.raw_cuda_graph_exec()
        if not isinstance(h, int):
            g.replay()
            return
        _native.call("fp3d_graph_launch", h, stream_handle())

    # graphs captured directly on callers' (b, x) buffers; a Krylov solve
    # passes one source and two alternating destinations, and the caching
    # allocator hands the same buffers to the next solve of the same size.
    # apply() fills its own slots; graph_execs() (Krylov workspaces that
    # launch the graphs themselves) has slots of its own, so one-off apply
    # calls on other buffers cannot use them up
    _MAX_DIRECT_GRAPHS = 4
    _MAX_EXEC_GRAPHS = 8

    def apply(self, b: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
        """x = M b (one V-cycle from a zero initial guess).  After the first
        call the V-cycle is replayed as a HIP graph: one captured on the
        caller's own b and x for up to four buffer pairs, else one on fixed
        buffers with copies in and out."""
        if not self.use_graph:
            return self.apply_eager(b, x)
        if self._eager_calls == 0:
            self._eager_calls += 1
            return self.apply_eager(b, x)
        direct = (b.is_contiguous() and x.is_contiguous() and b.dtype == torch.float64
                  and x.dtype == torch.float64 and b.numel() == x.numel() == self._gb.numel()
                  and b.data_ptr() != x.data_ptr())
        if direct:
            key = (b.data_ptr(), x.data_ptr())
            g = self._graphs.get(key) or self._exec_graphs.get(key)
            if g is None and len(self._graphs) < self._MAX_DIRECT_GRAPHS:
                g = self._graphs[key] = self._capture(b, x)
            if g is not None:
                self._launch(g)
                return x
        if self._graph is None:
            self._graph = self._capture(self._gb, self._gx)
        self._gb.copy_(b)
        self._launch(self._graph)
        x.copy_(self._gx)
        return x

    def graph_execs(self, b: torch.Tensor, xs) -> Optional[list]:
        """Raw hipGraphExec handles of this hierarchy's V-cycle from b into
        each x of xs (graphs on the callers' buffers, captured here when
        missing), for a caller that launches them itself; None when graphs
        are off, before the first eager V-cycle (which loads the kernels),
        or when the buffers cannot take a direct graph."""
        if not self.use_graph or self._eager_calls == 0:
            return None
        out = []
        for x in xs:
            if not (b.is_contiguous() and x.is_contiguous() and b.dtype == torch.float64
                    and x.dtype == torch.float64 and b.numel() == x.numel() == self._gb.numel()
                    and b.data_ptr() != x.data_ptr()):
                return None
            key = (b.data_ptr(), x.data_ptr())
            g = self._exec_graphs.get(key) or self._graphs.get(key)
            if g is None:
                if len(self._exec_graphs) >= self._MAX_EXEC_GRAPHS:
                    return None
                g = self._exec_graphs[key] = self._capture(b, x)
            h = g.raw_cuda_graph_exec()
            if not isinstance(h, int) or h == 0:
                return None
            out.append(h)
        return out

    def program(self, b: torch.Tensor, x: torch.Tensor) -> "VCycleProgram":
        """The V-cycle from b into x as device-operation records (cached per
        buffer pair), for grouped launches with other hierarchies."""
        key = (ptr(b), ptr(x))
        progs = self.__dict__.setdefault("_programs", {})
        p = progs.get(key)
        if p is None:
            p = progs[key] = VCycleProgram(self, b, x)
        return p

    def aspreconditioner(self):
        return self


# ------------------------------------------------------- grouped V-cycles
# Record layout of fp3d_grp_vcycle_op (16 int64 words per solve and
# operation): n, row pointer, columns, values, SpMV operand, kind,
# five pointers, two float64 scalars (as their bits), flag, two spare.
_GRP_WORDS = 16
(_K_STORE, _K_RESIDUAL, _K_ADD, _K_CHEB_FIRST, _K_CHEB_ITER, _K_CHEB_LAST,
 _K_RESTRICT_START) = range(7)
_K_CHEB_ZERO, _K_ADD_VEC, _K_TO_F32, _K_TO_F64, _K_DENSE = 10, 11, 12, 13, 14
_OP_SPMV, _OP_ELEM, _OP_DENSE = 0, 1, 2


def _lanes_for(n_rows: int, nnz: int) -> int:
    """Lanes per CSR row of the library's SpMV launches (its lanes_for)."""
    avg = nnz / n_rows if n_rows > 0 else 1.0
    if avg <= 6:
        return 4
    if avg <= 20:
        return 8
    if avg <= 48:
        return 16
    if avg <= 128:
        return 32
    return 64


def _bits(v: float) -> int:
    return int(np.float64(v).view(np.int64))


class VCycleProgram:
    """One V-cycle of an AMGHierarchy from b into x as the list of device
    operations AMGHierarchy.apply_eager launches, in the same order with
    the same parameters, each written as a record for fp3d_grp_vcycle_op.

    `signature` holds every operation's launch shape (operation, lanes per
    row, format bits, row / entry count); programs with equal signatures
    run grouped, one launch per operation for all of them (run_grouped),
    and each gets the bits of its own V-cycle."""

    def __init__(self, amg: "AMGHierarchy", b: torch.Tensor, x: torch.Tensor):
        self.amg = amg
        self._ops = []
        self._recs = []
        vf = amg.vfmt
        if amg.v32:
            n0 = amg.levels[0].n
            self._elem(_K_TO_F32, n0, vf, (ptr(b), ptr(amg._b32)), d0=1.0)
            self._level(0, amg._b32, amg._x32, False)
            self._elem(_K_TO_F64, n0, vf, (ptr(amg._x32), ptr(x)))
        else:
            self._level(0, b, x, False)
        self.signature = tuple(self._ops)
        self.table = np.stack(self._recs) if self._recs else np.zeros((0, _GRP_WORDS),
                                                                       np.int64)

    # -- record builders
    def _rec(self, op, G, fmt, n, words):
        self._ops.append((op, G, fmt, int(n)))
        w = np.zeros(_GRP_WORDS, dtype=np.int64)
        for i, v in words.items():
            w[i] = v
        self._recs.append(w)

    def _spmv(self, M: DeviceCSR, xin, kind, p=(), d0=0.0, d1=0.0, flag=0):
        fmt = M.fmt | self.amg.vfmt
        words = {0: M.n, 1: ptr(M.indptr), 2: ptr(M.index_array), 3: ptr(M.values),
                 4: ptr(xin), 5: kind, 11: _bits(d0), 12: _bits(d1), 13: int(flag)}
        for i, q in enumerate(p):
            words[6 + i] = ptr(q) if isinstance(q, torch.Tensor) else int(q)
        self._rec(_OP_SPMV, _lanes_for(M.n, M.nnz), fmt, M.n, words)

    def _elem(self, kind, n, fmt, p, d0=0.0, flag=0):
        words = {0: n, 5: kind, 11: _bits(d0), 13: int(flag)}
        for i, q in enumerate(p):
            words[6 + i] = ptr(q) if isinstance(q, torch.Tensor) else int(q)
        self._rec(_OP_ELEM, 0, fmt & 2, n, words)

    # -- the V-cycle (AMGHierarchy._vcycle and the library's cheb_smooth_t)
    def _cheb(self, L, b, x, zero_init):
        A = L.A
        theta = 0.5 * (L.lmax + L.lmin)
        delta = 0.5 * (L.lmax - L.lmin)
        sigma = theta / delta
        rho0 = 1.0 / sigma
        if zero_init == 1:
            self._elem(_K_CHEB_ZERO, L.n, self.amg.vfmt, (b, L.dinv, x, L.r, L.d0),
                       d0=1.0 / theta, flag=1 if L.degree == 1 else 0)
        elif zero_init == 0:
            self._spmv(A, x, _K_CHEB_FIRST, (b, L.dinv, L.r, L.d0), d0=1.0 / theta)
        dold, dnew = L.d0, L.d1
        for k in range(1, L.degree):
            rho1 = 1.0 / (2.0 * sigma - rho0)
            c1 = rho1 * rho0
            c2 = 2.0 * rho1 / delta
            kind = _K_CHEB_LAST if k == L.degree - 1 else _K_CHEB_ITER
            self._spmv(A, dold, kind, (L.dinv, dold, x, L.r, dnew), d0=c1, d1=c2)
            rho0 = rho1
            dold, dnew = dnew, dold
        if L.degree == 1 and zero_init == 0:
            self._elem(_K_ADD_VEC, L.n, self.amg.vfmt, (dold, x))

    def _level(self, l, b, x, started):
        amg = self.amg
        L = amg.levels[l]
        if l == len(amg.levels) - 1:
            words = {0: L.n, 5: _K_DENSE, 6: ptr(amg.coarse), 7: ptr(b), 8: ptr(x)}
            self._rec(_OP_DENSE, 0, amg.vfmt, L.n, words)
            return
        self._cheb(L, b, x, 2 if started else 1)
        self._spmv(L.A, x, _K_RESIDUAL, (b, L.r))
        Lc = amg.levels[l + 1]
        coarse_next = l + 1 == len(amg.levels) - 1
        if coarse_next:
            self._spmv(L.R, L.r, _K_STORE, (Lc.b,))
        else:
            theta_c = 0.5 * (Lc.lmax + Lc.lmin)
            self._spmv(L.R, L.r, _K_RESTRICT_START, (Lc.b, Lc.dinv, Lc.x, Lc.r, Lc.d0),
                       d0=1.0 / theta_c, flag=1 if Lc.degree == 1 else 0)
        self._level(l + 1, Lc.b, Lc.x, not coarse_next)
        self._spmv(L.P, Lc.x, _K_ADD, (x,))
        self._cheb(L, b, x, 0)


# (op, max_n, G, fmt) words of fp3d_grp_vcycle_run per program signature
_SIGNATURE_OPS: dict = {}


def run_grouped(programs) -> None:
    """One V-cycle per program, all of equal signature and on distinct
    hierarchies, queued as one fp3d_grp_vcycle_op launch per operation."""
    p0 = programs[0]
    if any(p.signature != p0.signature for p in programs[1:]):
        raise ValueError("run_grouped: programs of different signatures")
    if len({id(p.amg) for p in programs}) != len(programs):
        raise ValueError("run_grouped: two programs share a hierarchy's buffers")
    nb = len(programs)
    n_ops = len(p0.signature)
    table = np.ascontiguousarray(np.stack([p.table for p in programs], axis=1))
    host = torch.from_numpy(table.reshape(-1)).pin_memory()
    dev = torch.empty_like(host, device=p0.amg.coarse.device)
    dev.copy_(host, non_blocking=True)
    ops = _SIGNATURE_OPS.get(p0.signature)
    if ops is None:
        ops = _SIGNATURE_OPS[p0.signature] = np.array(
            [(op, n, G, fmt) for op, G, fmt, n in p0.signature], dtype=np.int64).reshape(-1)
    _native.call("fp3d_grp_vcycle_run", nb, ptr(dev), n_ops, ops.ctypes.data,
                 stream_handle())


def apply_preconditioners(calls) -> int:
    """psolve(b, x) for each (psolve, b, x) of one Krylov round.  AMG
    V-cycles (psolve = AMGHierarchy.apply) whose programs have equal
    signatures run grouped (run_grouped, one launch per operation for all
    of them); any other preconditioner, and a V-cycle with no partner, runs
    by itself.  Every V-cycle gives the bits it gives alone.  Returns the
    number of V-cycles that ran grouped."""
    groups: dict = {}
    for ps, b, x in calls:
        amg = getattr(ps, "__self__", None)
        if (isinstance(amg, AMGHierarchy)
                and getattr(ps, "__func__", None) is AMGHierarchy.apply):
            prog = amg.program(b, x)
            slots = groups.setdefault(prog.signature, [[]])
            # two V-cycles of one hierarchy share its level buffers: never
            # in the same grouped launch
            for g in slots:
                if all(q[0].amg is not amg for q in g):
                    g.append((prog, b, x))
                    break
            else:
                slots.append([(prog, b, x)])
        else:
            ps(b, x)
    n_grouped = 0
    for slots in groups.values():
        for g in slots:
            if len(g) == 1:
                prog, b, x = g[0]
                prog.amg.apply(b, x)
            else:
                run_grouped([q[0] for q in g])
                n_grouped += len(g)
    return n_grouped


class AMGPlan:
    """An AMGHierarchy built in three parts, so the host set-up can run on
    another thread while the caller's thread keeps the device busy:

      plan = AMGPlan(A, ..., device_pattern=..., device_values=...)  # caller's
          thread: the finest level's rho(D^-1 A) on the device
      h = plan.host()        # any thread: sa_setup (host only, no device
                             # or torch call)
      M = plan.finish(h)     # caller's thread: the device levels

    The hierarchy equals AMGHierarchy(A, ...) with the same arguments."""

    def __init__(self, A: sp.csr_matrix, max_levels: int = 10, max_coarse: int = 10,
                 seed: int = 0, device_pattern=None,
                 device_values: Optional[torch.Tensor] = None, **kw):
        self.A = A.tocsr()
        self.kw = dict(kw, max_levels=max_levels, max_coarse=max_coarse, seed=seed,
                       device_pattern=device_pattern, device_values=device_values)
        self.rho0 = kw.pop("rho0", None)
        if device_values is not None and self.rho0 is None:
            st = device_pattern
            d = self.A.diagonal()
            dinv = to_dev(np.where(d != 0.0, 1.0 / np.where(d != 0.0, d, 1.0), 0.0))
            self.rho0 = spectral_radius_dinv_a_dev(st.n_nodes, st.nnz, st.indptr,
                                                   st.indices, device_values, dinv,
                                                   seed=seed)
        self.kw["rho0"] = self.rho0

    def host(self):
        return sa_setup(self.A, max_levels=self.kw["max_levels"],
                        max_coarse=self.kw["max_coarse"], seed=self.kw["seed"],
                        rho0=self.rho0)

    def finish(self, host_setup) -> "AMGHierarchy":
        return AMGHierarchy(self.A, host_setup=host_setup, **self.kw)
