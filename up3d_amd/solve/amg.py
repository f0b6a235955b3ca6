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
float64 pseudo-inverse matvec.  Matrix values can be held in float32
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


# below this many stored entries of A the single-pass scipy product is the
# faster of the two (the values are the same)
_HOST_PARALLEL_NNZ = 1 << 21


def spgemm(A: sp.csr_matrix, B: sp.csr_matrix, force_native: bool = False
           ) -> sp.csr_matrix:
    """A @ B with the library's row-parallel host product: the values of
    scipy's product (same per-row accumulation order, exact zeros dropped),
    columns sorted.  Small products go to scipy itself."""
    A, B = A.tocsr(), B.tocsr()
    if A.shape[1] != B.shape[0]:
        raise ValueError(f"spgemm: shapes {A.shape} and {B.shape} do not align")
    if A.nnz < _HOST_PARALLEL_NNZ and not force_native:
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
        ny = np.linalg.norm(y)
        if ny == 0.0:
            return 1.0
        lam = ny / np.linalg.norm(x)
        x = y / ny
    return float(lam)


@dataclass
class HostLevel:
    A: sp.csr_matrix
    P: Optional[sp.csr_matrix]
    R: Optional[sp.csr_matrix]
    dinv: np.ndarray
    rho: float


def sa_setup(A: sp.csr_matrix, max_levels: int = 10, max_coarse: int = 10,
             omega: float = 4.0 / 3.0, seed: int = 0):
    """Host SA hierarchy: list of HostLevel plus the dense coarse pinv."""
    A = A.tocsr()
    levels: List[HostLevel] = []
    while True:
        n = A.shape[0]
        d = A.diagonal()
        dinv = np.where(d != 0.0, 1.0 / np.where(d != 0.0, d, 1.0), 0.0)
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
    coarse = np.linalg.pinv(levels[-1].A.toarray())
    return levels, coarse


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
                 v32: bool = False, compact_index: bool = True):
        self.n = h.A.shape[0]
        self.A = DeviceCSR(h.A, f32=f32, keep_f64=not f32, compact_index=compact_index)
        self.P = DeviceCSR(h.P, f32=f32, keep_f64=not f32) if h.P is not None else None
        self.R = DeviceCSR(h.R, f32=f32, keep_f64=not f32) if h.R is not None else None
        vt = torch.float32 if v32 else torch.float64
        self.dinv = to_dev(h.dinv, torch.float64).to(vt)
        self.lmax = 1.1 * h.rho
        self.lmin = self.lmax / eig_ratio
        self.degree = degree
        n = self.n
        self.x = empty(n, vt)
        self.b = empty(n, vt)
        self.r = empty(n, vt)
        self.d0 = empty(n, vt)
        self.d1 = empty(n, vt)


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
                 vectors: Optional[str] = None, compact_index: bool = True):
        if precision not in ("f64", "f32"):
            raise ValueError(f"unknown AMG precision {precision!r}")
        vectors = precision if vectors is None else vectors
        if vectors not in ("f64", "f32"):
            raise ValueError(f"unknown AMG vector precision {vectors!r}")
        host, coarse = sa_setup(A, max_levels=max_levels, max_coarse=max_coarse,
                                seed=seed)
        f32 = precision == "f32"
        self.precision = precision
        self.v32 = vectors == "f32"
        self.vfmt = 2 if self.v32 else 0
        self.levels = [_DevLevel(h, f32, degree, eig_ratio, self.v32, compact_index)
                       for h in host]
        self.coarse = to_dev(np.ascontiguousarray(coarse), torch.float64)
        self.shapes = [lv.n for lv in self.levels]
        self.nnzs = [lv.A.nnz for lv in self.levels]
        # HIP graph of the whole V-cycle on fixed buffers (captured lazily on
        # the first apply, after one eager V-cycle has loaded the kernels)
        self.use_graph = use_graph
        self._graph = None
        self._graphs = {}
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
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._vcycle_top(b, x)
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
    # allocator hands the same buffers to the next solve of the same size
    _MAX_DIRECT_GRAPHS = 4

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
            g = self._graphs.get(key)
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

    def aspreconditioner(self):
        return self
