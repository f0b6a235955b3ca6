"""ORACLE -- test infrastructure only (parity checker and CPU baseline).

A CPU restatement of the reference's hot path (swayli94/UP3D, pyfp3d/):
plain C kernels in fp3d_oracle.c (loaded with ctypes) plus the numpy /
scipy glue the reference itself uses around its numba kernels.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker or the timed CPU baseline;
the product package (up3d_amd) never imports it.

Each function cites the reference function it restates (file:line).

Pinning (what checks this restatement, in tests/test_oracle_*.py):
  physics        closed forms: rho(1)=1, rho(0)=(1+(g-1)/2 M^2)^(1/(g-1)),
                 M^2(1)=M^2_inf, M^2(q*^2)=1, rho==1 bitwise at M=0,
                 derivative FD checks, limiter identities
  geometry       sum_i B_i = 0, exact gradients of linear fields, sum V =
                 domain volume, degenerate guard
  colouring      Kuhn cube n=8 -> 31 colours / min class 3 / max class 153
                 (reference fixture cases/demo/p0_infrastructure/results/
                 g03_coloring.csv:2) and brute-force validity
  CSR pattern    nnz closed form for Kuhn grids + scipy COO->CSR definition
  residual/A     freestream phi = x -> interior R = 0; serial-reference
                 triplet assembly; Laplace MMS (harmonic quadratics)
  upwind walk    strictly upstream for uniform flow; nu == 0 subcritical;
                 brute-force walk restatement in pure Python
  sensitivities  central finite differences of the rho_tilde sweep
  Jacobian       J == A at s = 0; JVP vs central FD of the residual
  drivers        residual definition (||R||_inf < tol) and Laplace limit
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = Path(__file__).resolve().parent
_LIB = None

GAMMA = 1.4

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)
_i8 = C.POINTER(C.c_int8)


def _lib():
    global _LIB
    if _LIB is None:
        path = HERE / "lib" / "liboracle.so"
        if not path.exists():
            subprocess.run(["make", "-C", str(HERE)], check=True,
                           stdout=subprocess.DEVNULL)
        _LIB = C.CDLL(str(path))
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


def set_threads(n: int) -> int:
    """OpenMP threads of the C restatement's parallel loops; returns the
    previous count."""
    prev = int(_lib().or_get_threads())
    _lib().or_set_threads(C.c_int(int(n)))
    return prev


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32a(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _i64a(a):
    return np.ascontiguousarray(a, dtype=np.int64)


# ---------------------------------------------------------------- physics
# pyfp3d/physics/isentropic.py (scalar forms, float64)
def density_isentropic(q2, m, g=GAMMA):              # :29-58
    base = 1.0 + 0.5 * (g - 1.0) * (m * m) * (1.0 - q2)
    if base < 1e-10:
        base = 1e-10
    return base ** (1.0 / (g - 1.0))


def mach_number_squared(q2, m, g=GAMMA):             # :79-97
    return q2 * (m * m) / (1.0 + 0.5 * (g - 1.0) * (m * m) * (1.0 - q2))


def critical_speed_squared(m, g=GAMMA):              # :100-118
    return (2.0 + (g - 1.0) * m ** 2) / ((g + 1.0) * m ** 2)


def pressure_coefficient(q2, m, g=GAMMA):            # :121-143
    rho = density_isentropic(q2, m, g)
    return 2.0 / (g * m ** 2) * (rho ** g - 1.0)


def q2_at_mach(mach, m, g=GAMMA):                    # :261-289
    if m == 0.0:
        return np.inf
    m2 = mach * mach
    return m2 * (1.0 + 0.5 * (g - 1.0) * m ** 2) / (m ** 2 * (1.0 + 0.5 * (g - 1.0) * m2))


def density_field(q2, m, g=GAMMA):                   # :235-258
    q2 = _f64(q2)
    out = np.empty_like(q2)
    _lib().or_density_field(C.c_int64(len(q2)), _p(q2, _d), C.c_double(m),
                            C.c_double(g), _p(out, _d))
    return out


def mach_squared_field(q2, m, g=GAMMA):              # :317-334
    q2 = _f64(q2)
    out = np.empty_like(q2)
    _lib().or_mach2_field(C.c_int64(len(q2)), _p(q2, _d), C.c_double(m),
                          C.c_double(g), _p(out, _d))
    return out


def density_derivative_field(q2, m, g=GAMMA):        # :366-380
    q2 = _f64(q2)
    out = np.empty_like(q2)
    _lib().or_drho_field(C.c_int64(len(q2)), _p(q2, _d), C.c_double(m),
                         C.c_double(g), _p(out, _d))
    return out


def mach_squared_derivative_field(q2, m, g=GAMMA):   # :383-400
    q2 = _f64(q2)
    out = np.empty_like(q2)
    _lib().or_dmach2_field(C.c_int64(len(q2)), _p(q2, _d), C.c_double(m),
                           C.c_double(g), _p(out, _d))
    return out


def limit_q2_field(q2, m, m_cap, g=GAMMA):           # :292-314
    cap = q2_at_mach(m_cap, m, g)
    q2 = _f64(q2)
    return np.where(q2 < cap, q2, cap)


# ----------------------------------------------------------------- geometry
def precompute_element_geometry(nodes, elements):
    """pyfp3d/mesh/metrics.py:176-204 (+ :26-70, :129-173): V = |det J|/6,
    B = J^{-T} grad_ref via np.linalg.inv, degenerate guard -> ValueError."""
    nodes = _f64(nodes)
    el = np.asarray(elements, dtype=np.int64)
    x0 = nodes[el[:, 0]]
    J = np.stack([nodes[el[:, 1]] - x0, nodes[el[:, 2]] - x0,
                  nodes[el[:, 3]] - x0], axis=2)          # columns = edges
    v1, v2, v3 = J[:, :, 0], J[:, :, 1], J[:, :, 2]
    det3 = np.einsum("ij,ij->i", v1, np.cross(v2, v3))
    V = np.abs(det3) / 6.0
    scale = (np.linalg.norm(v1, axis=1) * np.linalg.norm(v2, axis=1)
             * np.linalg.norm(v3, axis=1))
    det = np.linalg.det(J)
    bad = (scale == 0.0) | (np.abs(det) < 1e-12 * scale)
    if bad.any():
        raise ValueError("element_gradients: degenerate tetrahedron "
                         "(|det J| < 1e-12 x edge-length^3 scale) -- fix the mesh")
    Jinv = np.linalg.inv(J)
    gref = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0],
                     [0.0, 0.0, 1.0]])
    B = np.einsum("nkd,ik->nid", Jinv, gref)               # (J^-T g)_d = sum_k Jinv[k,d] g_k
    return np.ascontiguousarray(B), V


def build_face_adjacency(elements):
    """pyfp3d/mesh/metrics.py:207-278, the dict restatement (first owner
    waits in the map; the next occurrence links and deletes the entry)."""
    el = np.asarray(elements)
    n = len(el)
    fd = ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2))
    fn = np.full((n, 4), -1, dtype=np.int32)
    first = {}
    for e in range(n):
        t = el[e]
        for f in range(4):
            key = tuple(sorted((int(t[fd[f][0]]), int(t[fd[f][1]]), int(t[fd[f][2]]))))
            if key in first:
                packed = first.pop(key)
                e1, f1 = divmod(packed, 4)
                fn[e1, f1] = e
                fn[e, f] = e1
            else:
                first[key] = e * 4 + f
    return fn


# --------------------------------------------------------------- colouring
def node_to_element_csr(elements):
    """pyfp3d/mesh/coloring.py:45-62."""
    n_tets = len(elements)
    n_nodes = int(np.max(elements)) + 1
    flat = np.asarray(elements, dtype=np.int64).reshape(-1)
    fe = np.repeat(np.arange(n_tets, dtype=np.int64), 4)
    order = np.argsort(flat, kind="stable")
    counts = np.bincount(flat, minlength=n_nodes)
    off = np.zeros(n_nodes + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    return off, fe[order]


def greedy_coloring(elements):
    """pyfp3d/mesh/coloring.py:65-126."""
    off, ne = node_to_element_csr(elements)
    el = _i32a(elements)
    colors = np.empty(len(el), dtype=np.int32)
    rc = _lib().or_greedy_coloring(C.c_int64(len(el)), _p(el, _i32),
                                   _p(off, _i64), _p(ne, _i64),
                                   _p(colors, _i32))
    if rc != 0:
        raise ValueError("greedy coloring exceeded 256 colors")
    return colors, int(colors.max()) + 1


def color_partition_csr(elements):
    """pyfp3d/mesh/coloring.py:129-145."""
    colors, n_colors = greedy_coloring(elements)
    order = np.argsort(colors, kind="stable")
    counts = np.bincount(colors, minlength=n_colors)
    off = np.zeros(n_colors + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    return off, order.astype(np.int64), n_colors


def validate_coloring(elements, colors):
    """pyfp3d/mesh/coloring.py:148-177 (vectorised: every node's incident
    elements carry pairwise distinct colours)."""
    off, ne = node_to_element_csr(elements)
    c = np.asarray(colors)[ne]
    node = np.repeat(np.arange(len(off) - 1), np.diff(off))
    key = node.astype(np.int64) * 1024 + c
    return len(np.unique(key)) == len(key)


# ----------------------------------------------------------- CSR + kernels
def build_csr_pattern(elements, n_nodes):
    """pyfp3d/kernels/jacobian.py:60-74 (scipy COO->CSR, sorted, deduped)."""
    el = np.asarray(elements, dtype=np.int64)
    rows = np.repeat(el, 4, axis=1).reshape(-1)
    cols = np.tile(el, (1, 4)).reshape(-1)
    pat = sp.coo_matrix((np.ones(len(rows)), (rows, cols)),
                        shape=(n_nodes, n_nodes)).tocsr()
    pat.sum_duplicates()
    pat.sort_indices()
    pat.data[:] = 0.0
    return pat


def build_elem_to_csr(elements, indptr, indices):
    """pyfp3d/kernels/jacobian.py:77-102 (returned int32, (n,4,4))."""
    el = _i32a(elements)
    ip = _i32a(indptr)
    ix = _i32a(indices)
    out = np.empty((len(el), 4, 4), dtype=np.int32)
    _lib().or_build_elem_to_csr(C.c_int64(len(el)), _p(el, _i32), _p(ip, _i32),
                                _p(ix, _i32), _p(out, _i32))
    return out


def element_velocity_q2(elements, B, phi):
    """pyfp3d/kernels/gradient.py:32-64."""
    el = _i32a(elements)
    B = _f64(B)
    phi = _f64(phi)
    grad = np.empty((len(el), 3))
    q2 = np.empty(len(el))
    _lib().or_element_velocity_q2(C.c_int64(len(el)), _p(el, _i32), _p(B, _d),
                                  _p(phi, _d), _p(grad, _d), _p(q2, _d))
    return grad, q2


def upstream_elements(face_neighbors, centroids, nodes, elements, grad,
                      max_hops=4):
    """pyfp3d/kernels/upwind.py:68-125 (int32 out)."""
    fn = _i32a(face_neighbors)
    cen = _f64(centroids)
    nd = _f64(nodes)
    el = _i32a(elements)
    g = _f64(grad)
    out = np.empty(len(el), dtype=np.int32)
    _lib().or_upstream_elements(C.c_int64(len(el)), _p(fn, _i32), _p(cen, _d),
                                _p(nd, _d), _p(el, _i32), _p(g, _d),
                                C.c_int(max_hops), _p(out, _i32))
    return out


def _sweep_args(q2, rho, upstream):
    return _f64(q2), _f64(rho), _i32a(upstream)


def rho_tilde_sweep(q2, rho, upstream, m, m_crit, Cc, g, rho_floor):
    """pyfp3d/kernels/upwind.py:456-500 -> (nu, rho_tilde)."""
    q2, rho, up = _sweep_args(q2, rho, upstream)
    nu = np.empty_like(q2)
    rt = np.empty_like(q2)
    _lib().or_rho_tilde_sweep(C.c_int64(len(q2)), _p(q2, _d), _p(rho, _d),
                              _p(up, _i32), C.c_double(m), C.c_double(m_crit),
                              C.c_double(Cc), C.c_double(g),
                              C.c_double(rho_floor), _p(nu, _d), _p(rt, _d))
    return nu, rt


def classify_upwind_branches(q2, rho, upstream, m, m_crit, Cc, g, rho_floor):
    """pyfp3d/kernels/upwind.py:503-541."""
    q2, rho, up = _sweep_args(q2, rho, upstream)
    br = np.empty(len(q2), dtype=np.int8)
    _lib().or_classify_upwind_branches(C.c_int64(len(q2)), _p(q2, _d),
                                       _p(rho, _d), _p(up, _i32), C.c_double(m),
                                       C.c_double(m_crit), C.c_double(Cc),
                                       C.c_double(g), C.c_double(rho_floor),
                                       _p(br, _i8))
    return br


def rho_tilde_frozen_sweep(q2, rho, upstream, branch, m, m_crit, Cc, g,
                           rho_floor):
    """pyfp3d/kernels/upwind.py:544-593 -> (nu, rho_tilde)."""
    q2, rho, up = _sweep_args(q2, rho, upstream)
    br = np.ascontiguousarray(branch, dtype=np.int8)
    nu = np.empty_like(q2)
    rt = np.empty_like(q2)
    _lib().or_rho_tilde_frozen_sweep(C.c_int64(len(q2)), _p(q2, _d), _p(rho, _d),
                                     _p(up, _i32), _p(br, _i8), C.c_double(m),
                                     C.c_double(m_crit), C.c_double(Cc),
                                     C.c_double(g), C.c_double(rho_floor),
                                     _p(nu, _d), _p(rt, _d))
    return nu, rt


def rho_tilde_frozen_sensitivities(q2, rho, upstream, branch, m, m_crit, Cc,
                                   g):
    """pyfp3d/kernels/upwind.py:596-643 -> (s_e, s_u)."""
    q2, rho, up = _sweep_args(q2, rho, upstream)
    br = np.ascontiguousarray(branch, dtype=np.int8)
    se = np.empty_like(q2)
    su = np.empty_like(q2)
    _lib().or_rho_tilde_frozen_sensitivities(
        C.c_int64(len(q2)), _p(q2, _d), _p(rho, _d), _p(up, _i32), _p(br, _i8),
        C.c_double(m), C.c_double(m_crit), C.c_double(Cc), C.c_double(g),
        _p(se, _d), _p(su, _d))
    return se, su


def rho_tilde_sensitivities(q2, rho, upstream, m, m_crit, Cc, g, rho_floor):
    """pyfp3d/kernels/upwind.py:646-730 -> (s_e, s_u)."""
    q2, rho, up = _sweep_args(q2, rho, upstream)
    se = np.empty_like(q2)
    su = np.empty_like(q2)
    _lib().or_rho_tilde_sensitivities(C.c_int64(len(q2)), _p(q2, _d),
                                      _p(rho, _d), _p(up, _i32), C.c_double(m),
                                      C.c_double(m_crit), C.c_double(Cc),
                                      C.c_double(g), C.c_double(rho_floor),
                                      _p(se, _d), _p(su, _d))
    return se, su


class OraclePicardOperator:
    """pyfp3d/kernels/jacobian.py:231-375 PicardOperator restated on the CPU
    (same set-up order: geometry, colouring, CSR pattern, elem_to_csr).
    order="serial" sums every residual entry and CSR slot in ascending
    element order (the reference's serial assembly,
    pyfp3d/kernels/residual.py:106-189); order="colored" in colour-class
    order (the reference's coloured kernels)."""

    def __init__(self, nodes, elements, order="serial"):
        if order not in ("serial", "colored"):
            raise ValueError(f"unknown order {order!r}")
        self.order = order
        self.nodes = _f64(nodes)
        self.elements = _i32a(elements)
        self.n_nodes = len(nodes)
        self.n_tets = len(elements)
        self.B, self.V = precompute_element_geometry(self.nodes, self.elements)
        self.color_offsets, self.color_elems, self.n_colors = \
            color_partition_csr(self.elements)
        self.pattern = build_csr_pattern(self.elements, self.n_nodes)
        self.indptr = _i32a(self.pattern.indptr)
        self.indices = _i32a(self.pattern.indices)
        self.elem_to_csr = build_elem_to_csr(self.elements, self.indptr,
                                             self.indices)
        self.rho_ones = np.ones(self.n_tets)

    def velocities(self, phi):
        return element_velocity_q2(self.elements, self.B, phi)

    def _csr(self, data):
        return sp.csr_matrix((data, self.pattern.indices, self.pattern.indptr),
                             shape=self.pattern.shape)

    def _serial(self, rt, grad, s_e, lim, inv_u2, clamp):
        data = np.empty(len(self.indices))
        rt = _f64(rt)
        g = None if grad is None else _f64(grad)
        se = None if s_e is None else _f64(s_e)
        lm = None if lim is None else np.ascontiguousarray(lim, dtype=np.uint8)
        _lib().or_assemble_jacobian_serial(
            C.c_int64(self.n_tets), _p(self.elem_to_csr, _i32), _p(self.B, _d),
            _p(self.V, _d), _p(rt, _d), None if g is None else _p(g, _d),
            None if se is None else _p(se, _d),
            None if lm is None else lm.ctypes.data_as(C.POINTER(C.c_uint8)),
            C.c_double(inv_u2), C.c_double(clamp), C.c_int64(len(data)),
            _p(data, _d))
        return data

    def jacobian_spd_terms12(self, grad, rho_tilde, s_e, inv_u2, clamp, lim=None):
        """The clamped (SPD) Terms 1+2 preconditioner matrix values."""
        return self._serial(rho_tilde, grad, s_e, lim, inv_u2, clamp)

    def assemble_matrix(self, rho_tilde=None):
        rt = self.rho_ones if rho_tilde is None else _f64(rho_tilde)
        if self.order == "serial":
            return self._csr(self._serial(rt, None, None, None, 0.0, 0.0))
        data = np.zeros(len(self.indices))
        _lib().or_assemble_matrix_colored(
            C.c_int(self.n_colors), _p(self.color_offsets, _i64),
            _p(self.color_elems, _i64), _p(self.elem_to_csr, _i32),
            _p(self.B, _d), _p(self.V, _d), _p(rt, _d), _p(data, _d))
        return self._csr(data)

    def assemble_residual(self, phi, rho_tilde=None, grad=None):
        rt = self.rho_ones if rho_tilde is None else _f64(rho_tilde)
        if grad is None:
            grad, _ = self.velocities(phi)
        grad = _f64(grad)
        R = np.empty(self.n_nodes)
        if self.order == "serial":
            _lib().or_assemble_residual_serial(
                C.c_int64(self.n_tets), _p(self.elements, _i32), _p(self.B, _d),
                _p(self.V, _d), _p(_f64(rt), _d), _p(grad, _d),
                C.c_int64(self.n_nodes), _p(R, _d))
            return R
        _lib().or_assemble_residual_colored(
            C.c_int(self.n_colors), _p(self.color_offsets, _i64),
            _p(self.color_elems, _i64), _p(self.elements, _i32), _p(self.B, _d),
            _p(self.V, _d), _p(rt, _d), _p(grad, _d), C.c_int64(self.n_nodes),
            _p(R, _d))
        return R

    def jacobian_terms12(self, grad, rho_tilde, s_e, inv_u2):
        if self.order == "serial":
            return self._serial(rho_tilde, grad, s_e, None, inv_u2, 0.0)
        data = np.zeros(len(self.indices))
        g = _f64(grad)
        rt = _f64(rho_tilde)
        se = _f64(s_e)
        _lib().or_assemble_jacobian_colored(
            C.c_int(self.n_colors), _p(self.color_offsets, _i64),
            _p(self.color_elems, _i64), _p(self.elem_to_csr, _i32),
            _p(self.B, _d), _p(self.V, _d), _p(rt, _d), _p(g, _d), _p(se, _d),
            C.c_double(inv_u2), _p(data, _d))
        return data

    def term3_coo(self, grad, upstream, s_u, inv_u2):
        active = np.nonzero(s_u)[0].astype(np.int64)
        n = len(active)
        rows = np.empty(16 * n, dtype=np.int64)
        cols = np.empty(16 * n, dtype=np.int64)
        vals = np.empty(16 * n)
        if n:
            g = _f64(grad)
            up = _i32a(upstream)
            su = _f64(s_u)
            _lib().or_fill_term3_coo(
                C.c_int64(n), _p(active, _i64), _p(self.elements, _i32),
                _p(self.B, _d), _p(self.V, _d), _p(g, _d), _p(up, _i32),
                _p(su, _d), C.c_double(inv_u2), _p(rows, _i64), _p(cols, _i64),
                _p(vals, _d))
        return active, rows, cols, vals

    def assemble_newton_jacobian(self, phi, rho_tilde, s_e, s_u, upstream,
                                 u_inf=1.0, lim_mask=None):
        """pyfp3d/kernels/jacobian.py:288-359."""
        up = np.asarray(upstream)
        if lim_mask is not None:
            s_e = s_e * lim_mask
            s_u = s_u * lim_mask[up]
        inv_u2 = 1.0 / (u_inf * u_inf)
        grad, _ = self.velocities(phi)
        J = self._csr(self.jacobian_terms12(grad, rho_tilde, s_e, inv_u2))
        active, rows, cols, vals = self.term3_coo(grad, up, s_u, inv_u2)
        self.n_term3_active = len(active)
        if len(active):
            J = (J + sp.coo_matrix((vals, (rows, cols)),
                                   shape=self.pattern.shape).tocsr()).tocsr()
        return J


class OracleUpwind:
    """pyfp3d/kernels/upwind.py:733-920 UpwindOperator (walk mode)."""

    def __init__(self, nodes, elements):
        self.nodes = _f64(nodes)
        self.elements = _i32a(elements)
        self.face_neighbors = build_face_adjacency(self.elements)
        self.centroids = self.nodes[self.elements].mean(axis=1)

    def upstream(self, grad):
        return upstream_elements(self.face_neighbors, self.centroids,
                                 self.nodes, self.elements, grad)

    def rho_tilde(self, grad, q2, rho, m, Cc, m_crit, g=GAMMA, rho_floor=0.05):
        up = self.upstream(grad)
        nu, rt = rho_tilde_sweep(q2, rho, up, m, m_crit, Cc, g, rho_floor)
        self.upstream_last = up
        self.nu = nu
        self.nu_max = float(nu.max())
        self.n_supersonic = int(np.count_nonzero(nu > 0.0))
        self.n_floored = int(np.count_nonzero(rt == rho_floor))
        return rt

    def rho_tilde_sensitivities(self, grad, q2, rho, m, Cc, m_crit, g=GAMMA,
                                rho_floor=0.05):
        up = self.upstream(grad)
        se, su = rho_tilde_sensitivities(q2, rho, up, m, m_crit, Cc, g,
                                         rho_floor)
        return se, su, up

    def freeze_upwind_state(self, grad, q2, rho, m, Cc, m_crit, g=GAMMA,
                            rho_floor=0.05):
        up = self.upstream(grad)
        br = classify_upwind_branches(q2, rho, up, m, m_crit, Cc, g, rho_floor)
        return up.copy(), br.copy()


# ------------------------------------------- serial reference assembly
def stiffness_matrix_reference(nodes, elements):
    """pyfp3d/kernels/residual.py:106-141: serial triplets, scipy COO->CSR."""
    B, V = precompute_element_geometry(nodes, elements)
    el = np.asarray(elements, dtype=np.int64)
    vals = np.einsum("n,nid,njd->nij", V, B, B).reshape(-1)
    rows = np.repeat(el, 4, axis=1).reshape(-1)
    cols = np.tile(el, (1, 4)).reshape(-1)
    A = sp.coo_matrix((vals, (rows, cols)), shape=(len(nodes), len(nodes))).tocsr()
    A.sort_indices()
    return A


def residual_reference(nodes, elements, phi):
    """pyfp3d/kernels/residual.py:144-189: serial Laplace residual."""
    B, V = precompute_element_geometry(nodes, elements)
    el = np.asarray(elements, dtype=np.int64)
    g = np.einsum("ni,nid->nd", np.asarray(phi)[el], B)
    contrib = V[:, None] * np.einsum("nd,nid->ni", g, B)
    R = np.zeros(len(nodes))
    np.add.at(R, el.reshape(-1), contrib.reshape(-1))
    return R


# ------------------------------------------------------------- wall force
def wall_force_coefficients(nodes, elements, wall_faces, phi, alpha_deg=0.0,
                            u_inf=1.0, s_ref=1.0, m_inf=0.0):
    """pyfp3d/post/surface.py:80-118, 192-231, 265-363 (smooth_passes=0)."""
    nodes = _f64(nodes)
    wf = np.asarray(wall_faces, dtype=np.int64)
    p0, p1, p2 = nodes[wf[:, 0]], nodes[wf[:, 1]], nodes[wf[:, 2]]
    edge1 = p1 - p0
    area_vec = np.cross(edge1, p2 - p0)
    twice = np.linalg.norm(area_vec, axis=1)
    tri_n = area_vec / twice[:, None]
    e1 = edge1 / np.linalg.norm(edge1, axis=1)[:, None]
    e2 = np.cross(tri_n, e1)
    u1, v1 = np.sum(edge1 * e1, axis=1), np.sum(edge1 * e2, axis=1)
    edge2 = p2 - p0
    u2, v2 = np.sum(edge2 * e1, axis=1), np.sum(edge2 * e2, axis=1)
    phi = np.asarray(phi)
    df1 = phi[wf[:, 1]] - phi[wf[:, 0]]
    df2 = phi[wf[:, 2]] - phi[wf[:, 0]]
    det = u1 * v2 - v1 * u2
    du = (v2 * df1 - v1 * df2) / det
    dv = (-u2 * df1 + u1 * df2) / det
    grad_tri = du[:, None] * e1 + dv[:, None] * e2
    area = 0.5 * twice
    # outward normal from the owning tet (fluid side)
    el = np.asarray(elements, dtype=np.int64)
    tf = np.concatenate([el[:, [1, 2, 3]], el[:, [0, 2, 3]], el[:, [0, 1, 3]],
                         el[:, [0, 1, 2]]])
    tf.sort(axis=1)
    owners = np.tile(np.arange(len(el)), 4)
    key = {tuple(r): o for r, o in zip(map(tuple, tf), owners)}
    owner = np.array([key[tuple(sorted(map(int, f)))] for f in wf])
    normal = tri_n.copy()
    fc = nodes[wf].mean(axis=1)
    oc = nodes[el[owner]].mean(axis=1)
    flip = np.sum(normal * (oc - fc), axis=1) < 0
    normal[flip] *= -1.0
    q2 = np.sum(grad_tri * grad_tri, axis=1) / u_inf ** 2
    if m_inf > 0.0:
        cp = np.array([pressure_coefficient(q, m_inf) for q in q2])
    else:
        cp = 1.0 - q2
    cf = -(cp * area) @ normal / s_ref
    a = np.deg2rad(alpha_deg)
    lift = np.array([-np.sin(a), np.cos(a), 0.0])
    drag = np.array([np.cos(a), np.sin(a), 0.0])
    return {"cf": cf, "cl": float(cf @ lift), "cd_pressure": float(cf @ drag),
            "cp_tri": cp}


# ------------------------------------------------------------------ drivers
def solve_subsonic(nodes, elements, dirichlet_nodes, dirichlet_values, m_inf,
                   gamma_air=1.4, u_inf=1.0, body_source_rhs=None,
                   n_picard_max=40, tol_rho=1e-10):
    """pyfp3d/solve/picard.py:271-417 with a direct sparse solve in place of
    CG+AMG (the converged iterate does not depend on the linear solver)."""
    op = OraclePicardOperator(nodes, elements)
    n = op.n_nodes
    is_dir = np.zeros(n, dtype=bool)
    is_dir[dirichlet_nodes] = True
    free = np.where(~is_dir)[0]
    vals_d = _f64(dirichlet_values)
    b_wall = None if body_source_rhs is None else _f64(body_source_rhs)
    phi = np.zeros(n)
    phi[dirichlet_nodes] = vals_d
    rho = np.ones(op.n_tets)
    res_hist, drho_hist = [], []
    converged = False
    for _ in range(n_picard_max):
        A = op.assemble_matrix(rho)
        A_free = A[free][:, free].tocsc()
        b_free = -A[free][:, dirichlet_nodes] @ vals_d
        if b_wall is not None:
            b_free = b_free + b_wall[free]
        x = spla.spsolve(A_free, b_free)
        phi = np.empty(n)
        phi[free] = x
        phi[dirichlet_nodes] = vals_d
        _, q2 = op.velocities(phi)
        rho_new = density_field(q2 / u_inf ** 2, m_inf, gamma_air)
        R = op.assemble_residual(phi, rho_new)
        if b_wall is not None:
            R = R - b_wall
        res_hist.append(float(np.max(np.abs(R[free]))))
        drho = float(np.max(np.abs(rho_new - rho)))
        drho_hist.append(drho)
        rho = rho_new
        if drho < tol_rho:
            converged = True
            break
    return {"phi": phi, "n_picard": len(res_hist), "converged": converged,
            "residual_history": res_hist, "drho_history": drho_hist, "rho": rho}


class OracleNewton:
    """pyfp3d/solve/newton.py:97-501 NewtonWorkspace.eval_residual /
    assemble_coupled for a non-lifting case (no wake stations), with the
    reference's precond="direct" linear step (splu of J_ff)."""

    def __init__(self, nodes, elements, dirichlet_nodes, dirichlet_values,
                 u_inf=1.0, gamma_air=1.4, external_rhs=None):
        self.op = OraclePicardOperator(nodes, elements)
        self.upw = OracleUpwind(nodes, elements)
        n = self.op.n_nodes
        is_dir = np.zeros(n, dtype=bool)
        is_dir[dirichlet_nodes] = True
        self.dir = np.asarray(dirichlet_nodes)
        self.free = np.where(~is_dir)[0]
        self.vals = _f64(dirichlet_values)
        self.u_inf = u_inf
        self.gamma_air = gamma_air
        self.b_ext = None if external_rhs is None else _f64(external_rhs)

    def eval_residual(self, phi_free, m_inf, upwind_c=1.5, m_crit=0.95,
                      m_cap=3.0, rho_floor=0.05, frozen=None):
        phi = np.empty(self.op.n_nodes)
        phi[self.free] = phi_free
        phi[self.dir] = self.vals
        grad, q2 = self.op.velocities(phi)
        q2n = q2 / self.u_inf ** 2
        q2l = limit_q2_field(q2n, m_inf, m_cap, self.gamma_air)
        lim = q2l == q2n
        rho = density_field(q2l, m_inf, self.gamma_air)
        if m_inf > 0.0 and frozen is not None:
            nu, rt = rho_tilde_frozen_sweep(q2l, rho, frozen[0], frozen[1],
                                            m_inf, m_crit, upwind_c,
                                            self.gamma_air, rho_floor)
            n_floored = int(np.count_nonzero(frozen[1] == 3))
        elif m_inf > 0.0:
            rt = self.upw.rho_tilde(grad, q2l, rho, m_inf, upwind_c, m_crit,
                                    self.gamma_air, rho_floor)
            n_floored = self.upw.n_floored
        else:
            rt = rho
            n_floored = 0
        R = self.op.assemble_residual(phi, rt, grad)
        if self.b_ext is not None:
            R = R - self.b_ext
        state = {"phi": phi, "grad": grad, "q2l": q2l, "lim": lim, "rho": rho,
                 "rho_t": rt, "n_limited": int(np.count_nonzero(~lim)),
                 "n_floored": n_floored}
        return R[self.free], state

    def jacobian(self, state, m_inf, upwind_c=1.5, m_crit=0.95, rho_floor=0.05,
                 frozen=None):
        n_tets = self.op.n_tets
        if m_inf > 0.0 and frozen is not None:
            up = frozen[0]
            se, su = rho_tilde_frozen_sensitivities(
                state["q2l"], state["rho"], frozen[0], frozen[1], m_inf, m_crit,
                upwind_c, self.gamma_air)
        elif m_inf > 0.0:
            se, su, up = self.upw.rho_tilde_sensitivities(
                state["grad"], state["q2l"], state["rho"], m_inf, upwind_c,
                m_crit, self.gamma_air, rho_floor)
        else:
            se = np.zeros(n_tets)
            su = np.zeros(n_tets)
            up = np.arange(n_tets, dtype=np.int32)
        lim = state["lim"]
        J = self.op.assemble_newton_jacobian(
            state["phi"], state["rho_t"], se, su, up, u_inf=self.u_inf,
            lim_mask=None if lim.all() else lim)
        return J[self.free][:, self.free].tocsc()

    def solve(self, m_inf, phi_init, upwind_c=1.5, m_crit=0.95, m_cap=3.0,
              rho_floor=0.05, n_newton_max=30, tol_residual=1e-10,
              freeze_tol=None, freeze_refresh_max=2):
        """pyfp3d/solve/newton.py:781-1087 control flow for n_st = 0: freeze
        on r < freeze_tol or a live stall (:866-881), honesty re-evaluation
        and refresh of the freeze (:893-947), safety line search
        (:1059-1087); direct sparse linear step."""
        phi_free = _f64(phi_init)[self.free].copy()
        R, st = self.eval_residual(phi_free, m_inf, upwind_c, m_crit, m_cap,
                                   rho_floor)
        merit = float(R @ R)
        hist = []
        frozen = None
        converged = False
        n_refresh = 0
        prev_live = None
        for _ in range(n_newton_max):
            r = float(np.max(np.abs(R)))
            hist.append(r)
            stalled = (len(hist) >= 7 and r < 1e-3
                       and r > 0.25 * min(hist[-7:-1])
                       and r < 2.0 * min(hist[-7:-1]))
            if (frozen is None and freeze_tol is not None and m_inf > 0.0
                    and (r < freeze_tol or stalled) and st["n_limited"] == 0
                    and st["n_floored"] == 0):
                frozen = self.upw.freeze_upwind_state(
                    st["grad"], st["q2l"], st["rho"], m_inf, upwind_c, m_crit,
                    self.gamma_air, rho_floor)
            if r < tol_residual and st["n_limited"] == 0 and st["n_floored"] == 0:
                if frozen is None:
                    converged = True
                    break
                R_live, st_live = self.eval_residual(phi_free, m_inf, upwind_c,
                                                     m_crit, m_cap, rho_floor)
                r_live = float(np.max(np.abs(R_live)))
                if r_live < tol_residual:
                    converged = True
                    st = st_live
                    break
                if (prev_live is not None and r_live > 0.5 * prev_live) \
                        or n_refresh >= freeze_refresh_max:
                    converged = True
                    break
                prev_live = r_live
                n_refresh += 1
                frozen = self.upw.freeze_upwind_state(
                    st_live["grad"], st_live["q2l"], st_live["rho"], m_inf,
                    upwind_c, m_crit, self.gamma_air, rho_floor)
                R, st = R_live, st_live
                merit = float(R @ R)
                continue
            J = self.jacobian(st, m_inf, upwind_c, m_crit, rho_floor, frozen)
            dphi = spla.splu(J).solve(-R)
            lam = 1.0
            best = None
            for _ls in range(5):
                Rt, stt = self.eval_residual(phi_free + lam * dphi, m_inf,
                                             upwind_c, m_crit, m_cap, rho_floor,
                                             frozen)
                mt = float(Rt @ Rt)
                if np.isfinite(mt) and mt <= merit:
                    best = (mt, lam, Rt, stt)
                    break
                if np.isfinite(mt) and (best is None or mt < best[0]):
                    best = (mt, lam, Rt, stt)
                lam *= 0.5
            merit, lam, R, st = best
            phi_free = phi_free + lam * dphi
        return {"phi": st["phi"], "converged": converged,
                "residual_history": hist, "state": st}
