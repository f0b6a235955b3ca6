"""Device assembly of the level-set wake's sheet terms
(pyfp3d/kernels/cut_assembly.py:119-556, pyfp3d/wake/multivalued.py:276-379).

Everything that touches a special element (the two copies of each cut
element, the lower copy of each below-TE fan element) or an aux row lands in
an extended CSR pattern that the cut census fixes.  `SheetMatrix` builds that
pattern once on the host from a list of contributions (row, column, kind,
source element, local pair, static coefficient) and each assembly fills its
values on the device with one lane per slot (`fp3d_sheet_fill`), summing the
slot's contributions in their fixed order.  The Newton Term 3 of the special
copies couples rows to the DOFs of the element the upwind walk selects, so
its columns move with the selection; `SpecialTerm3` applies it as an
operator (`fp3d_sheet_t3_prepare` / `fp3d_sheet_t3_apply`) rebuilt per
linearisation.  Only O(sheet) index lists and the TE Kutta values (one row
per TE node) cross between host and device per call.
"""

from __future__ import annotations

from typing import Optional

import numpy as np
import scipy.sparse as sp
import torch

from .. import _native
from ..device import empty, ptr, stream_handle, to_dev, to_host

# contribution kinds of fp3d_sheet_fill
UPPER, LOWER, STATIC, HOSTVAL, TERM2_UPPER, TERM2_LOWER = range(6)


def _i32(a):
    return to_dev(np.ascontiguousarray(a, dtype=np.int32), torch.int32)


def _u8(a):
    return to_dev(np.ascontiguousarray(a, dtype=np.uint8), torch.uint8)


class SheetMatrix:
    """Fixed-pattern (n x n) CSR on the device whose values are sums of
    per-call contributions (see the module docstring)."""

    def __init__(self, n: int, rows, cols, kind, src, ab, coef):
        rows = np.asarray(rows, dtype=np.int64)
        cols = np.asarray(cols, dtype=np.int64)
        if rows.shape != cols.shape or (len(rows) and (rows.min() < 0 or cols.min() < 0
                                                       or rows.max() >= n or cols.max() >= n)):
            raise ValueError("sheet contributions out of range")
        self.n = int(n)
        key = rows * n + cols
        order = np.argsort(key, kind="stable")
        ks = key[order]
        uniq, first = np.unique(ks, return_index=True)
        self.n_slots = int(len(uniq))
        srow = uniq // n
        scol = uniq % n
        self.indptr_host = np.searchsorted(srow, np.arange(n + 1)).astype(np.int32)
        self.indices_host = scol.astype(np.int32)
        self.srow_host = srow.astype(np.int32)
        self.n_contrib = int(len(rows))
        self.indptr = _i32(self.indptr_host)
        self.indices = _i32(self.indices_host)
        self.srow = _i32(self.srow_host)
        self.seg = _i32(np.append(first, len(ks)))
        self.kind = _u8(np.asarray(kind)[order])
        self.src = _i32(np.asarray(src)[order])
        self.ab = _u8(np.asarray(ab)[order])
        self.coef = to_dev(np.ascontiguousarray(np.asarray(coef, dtype=np.float64)[order]))
        self.data = empty(max(self.n_slots, 1))
        self._blocks = {}
        self.kinds = frozenset(int(k) for k in np.unique(np.asarray(kind)))
        src = np.asarray(src, dtype=np.int64)
        host = src[np.asarray(kind) == HOSTVAL]
        self.n_hostval = int(host.max()) + 1 if len(host) else 0

    def fill(self, wa=None, wb=None, kv=None, sa=None, sb=None, ga=None, gb=None,
             B=None, fmask=None) -> torch.Tensor:
        """Values of every slot into self.data (device) and return it."""
        need = {UPPER: (wa,), LOWER: (wb,), HOSTVAL: (kv,), TERM2_UPPER: (sa, ga, B),
                TERM2_LOWER: (sb, gb, B)}
        for k in self.kinds:
            if any(a is None for a in need.get(k, ())):
                raise ValueError(f"sheet fill: contributions of kind {k} need their inputs")
        if self.n_hostval and (kv is None or len(kv) < self.n_hostval):
            raise ValueError(f"sheet fill: {self.n_hostval} host values needed")
        kvd = None
        if kv is not None:
            kvd = kv if isinstance(kv, torch.Tensor) else to_dev(np.asarray(kv, np.float64))
        p = lambda t: None if t is None else ptr(t)  # noqa: E731
        _native.call("fp3d_sheet_fill", self.n_slots, ptr(self.seg), ptr(self.kind),
                     ptr(self.src), ptr(self.ab), ptr(self.coef), p(wa), p(wb),
                     p(kvd), p(sa), p(sb), p(ga), p(gb), p(B), ptr(self.srow),
                     ptr(self.indices), p(fmask), ptr(self.data), stream_handle())
        return self.data

    def add_matvec(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        """y += S x."""
        if self.n_slots:
            _native.call("fp3d_amg_prolong_add", self.n, self.n_slots, ptr(self.indptr),
                         ptr(self.indices), ptr(self.data), 0, ptr(x), ptr(y),
                         stream_handle())
        return y

    def rows_matvec(self, r0: int, x: torch.Tensor, y_rows: torch.Tensor) -> torch.Tensor:
        """y_rows = S[r0:, :] x (y_rows has n - r0 entries)."""
        nr = self.n - r0
        if nr <= 0:
            return y_rows
        nnz = int(self.indptr_host[-1] - self.indptr_host[r0])
        if nnz == 0:
            y_rows.zero_()
            return y_rows
        _native.call("fp3d_csr_spmv", nr, nnz, ptr(self.indptr) + 4 * r0,
                     ptr(self.indices), ptr(self.data), ptr(x), ptr(y_rows),
                     stream_handle())
        return y_rows

    def to_scipy(self) -> sp.csr_matrix:
        d = to_host(self.data)[: self.n_slots] if self.n_slots else np.zeros(0)
        return sp.csr_matrix((d, self.indices_host, self.indptr_host), shape=(self.n, self.n))

    def block_csc(self, r0: int, c0: int) -> sp.csc_matrix:
        """The block S[r0:, c0:] on the host (the aux block of the extended
        system: O(sheet) values copied per call, pattern cached)."""
        key = (r0, c0)
        if key not in self._blocks:
            a, b = int(self.indptr_host[r0]), int(self.indptr_host[-1])
            slots = np.arange(a, b)
            keep = self.indices_host[a:b] >= c0
            slots = slots[keep]
            rr = self.srow_host[slots] - r0
            cc = self.indices_host[slots] - c0
            self._blocks[key] = (_i32(slots).long(), rr, cc)
        sl, rr, cc = self._blocks[key]
        vals = to_host(self.data[sl]) if len(rr) else np.zeros(0)
        return sp.csc_matrix((vals, (rr, cc)), shape=(self.n - r0, self.n - c0))

    def diagonal_slots(self, n_rows: int) -> np.ndarray:
        """Slot of (i, i) for i < n_rows, -1 where the pattern has none."""
        out = np.full(n_rows, -1, dtype=np.int64)
        for_rows = self.srow_host < n_rows
        diag = for_rows & (self.srow_host == self.indices_host)
        out[self.srow_host[diag]] = np.flatnonzero(diag)
        return out


class SpecialTerm3:
    """Term 3 of the special copies and of plain elements whose upstream
    element is special, on the device (module docstring).  Static: each
    side's DOF table (n_tets, 4), its special element list and its own-side
    plain mask, the row map of the extended rows (-1 = dropped row)."""

    def __init__(self, mv, row_map):
        cm = mv.cm
        ru, rl = mv.read_dofs()
        self.mv = mv
        self.n = mv.n_total
        self.rd = (_i32(ru), _i32(rl))
        self.row_map = _i32(row_map)
        sp_ids = (np.asarray(cm.cut_elems, dtype=np.int64),
                  np.concatenate([cm.cut_elems, cm.te_lower_elems]).astype(np.int64))
        self.sp_ids = tuple(to_dev(a, torch.int64) for a in sp_ids)
        special = np.zeros(len(mv.el), dtype=bool)
        special[mv.special] = True
        self.special = to_dev(special.astype(np.uint8), torch.uint8).bool()
        plain = (~mv.is_cut) & (~mv.is_tel)
        own = (mv.plain_plus, plain & ~mv.plain_plus)
        self.own = tuple(to_dev(o.astype(np.uint8), torch.uint8).bool() for o in own)
        self.n_cand = 0

    def build(self, sides, inv_u2: float, fmask=None) -> None:
        """Candidates, entries and the row gather for the current selection
        (rows and columns where the device uint8 `fmask` is 0 left out)."""
        st = self.mv.op.store
        cands, csides = [], []
        for k, sd in enumerate(sides):
            up = sd["upstream"].long()
            extra = torch.nonzero(self.own[k] & self.special[up]).flatten()
            c = torch.cat([self.sp_ids[k], extra])
            cands.append(c)
            csides.append(torch.full((c.numel(),), k, dtype=torch.uint8, device=c.device))
        cand = torch.cat(cands).to(torch.int32).contiguous()
        cside = torch.cat(csides).contiguous()
        nc = int(cand.numel())
        self.n_cand = nc
        if nc == 0:
            return
        self.rows = empty(4 * nc, torch.int32)
        self.W = empty(4 * nc)
        self.gu = empty(4 * nc)
        self.cols = empty(4 * nc, torch.int32)
        s0, s1 = sides
        _native.call("fp3d_sheet_t3_prepare", nc, ptr(cand), ptr(cside), ptr(s0["grad"]),
                     ptr(s1["grad"]), ptr(s0["upstream"]), ptr(s1["upstream"]),
                     ptr(s0["s_u"]), ptr(s1["s_u"]), ptr(s0["lim"]), ptr(s1["lim"]),
                     ptr(self.rd[0]), ptr(self.rd[1]), ptr(self.rd[0]), ptr(self.rd[1]),
                     ptr(self.row_map), None if fmask is None else ptr(fmask),
                     ptr(st.B), ptr(st.V), float(inv_u2),
                     ptr(self.rows), ptr(self.W), ptr(self.gu), ptr(self.cols),
                     stream_handle())
        r = to_host(self.rows).astype(np.int64)
        ent = np.flatnonzero(r >= 0)
        order = np.argsort(r[ent], kind="stable")
        rent = ent[order]
        rptr = np.searchsorted(r[rent], np.arange(self.n + 1))
        self.rent = _i32(rent)
        self.rptr = _i32(rptr)
        self.t = empty(nc)
        self.n_entries = int(len(rent))

    def add_matvec(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        """y += T x."""
        if self.n_cand and self.n_entries:
            _native.call("fp3d_sheet_t3_apply", self.n_cand, self.n, ptr(self.gu),
                         ptr(self.cols), ptr(self.W), ptr(self.rptr), ptr(self.rent),
                         ptr(x), ptr(self.t), ptr(y), stream_handle())
        return y

    def to_coo(self):
        """(rows, cols, vals) on the host: entry (c, a, b) = W[c,a] gu[c,b]."""
        if not self.n_cand:
            z = np.empty(0, dtype=np.int64)
            return z, z, np.empty(0)
        r = to_host(self.rows).reshape(-1, 4).astype(np.int64)
        W = to_host(self.W).reshape(-1, 4)
        gu = to_host(self.gu).reshape(-1, 4)
        cols = to_host(self.cols).reshape(-1, 4).astype(np.int64)
        c_i, a_i = np.nonzero(r >= 0)
        rows = np.repeat(r[c_i, a_i], 4)
        cc = cols[c_i].reshape(-1)
        vals = (W[c_i, a_i][:, None] * gu[c_i]).reshape(-1)
        return rows, cc, vals
