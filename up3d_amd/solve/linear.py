"""Device Krylov solvers and the reference's linear-algebra entry points
(pyfp3d/solve/linear.py).

* gmres_dev: restarted GMRES with left preconditioning and the same
  restart / inner-tolerance logic as the SciPy routine the reference calls
  (scipy.sparse.linalg.gmres, pyfp3d/solve/linear.py:134-201): convergence
  is declared on the TRUE residual ||b - A x|| <= max(atol, rtol ||b||).
  The Arnoldi step uses classical Gram-Schmidt with a re-orthogonalisation
  pass where the first one cancelled (CGS2 with the DGKS test, decided on
  the device), so each orthogonalisation is one or two fused multi-dot /
  multi-axpy passes instead of k dependent dot products; dots and norms are
  deterministic device reductions in float64, Givens rotations run on the
  host in float64 (one device->host copy per Krylov step).
* pcg_dev: preconditioned CG (scipy.sparse.linalg.cg semantics) for the SPD
  Picard systems.
* NewtonOperator: J_ff x with J's Terms 1+2 in CSR (float64 or float32
  values) and Term 3 applied matrix-free, Dirichlet rows as identity.
"""

from __future__ import annotations

from typing import Callable, Optional, Tuple

import numpy as np
import scipy.sparse as sp
import torch
from scipy.linalg import get_lapack_funcs

from .. import _native
from ..device import empty, ptr, require_gpu, stream_handle, to_dev, to_host, zeros
from .amg import AMGHierarchy, DeviceCSR, apply_preconditioners

_lartg = get_lapack_funcs("lartg", dtype=np.float64)


class _Red:
    """Device scratch for the deterministic reductions."""

    def __init__(self, max_nv: int):
        L = _native.lib()
        self.work = empty(max(L.fp3d_multi_dot_work_doubles(max_nv),
                              L.fp3d_reduce_work_doubles()))
        self.out = empty(2 * max_nv + 10)

    def norm(self, x: torch.Tensor, at: int = 0) -> None:
        _native.call("fp3d_norms", x.numel(), ptr(x), ptr(self.work),
                     ptr(self.out[at:]), stream_handle())

    def dots(self, V: torch.Tensor, nv: int, w: torch.Tensor, at: int) -> None:
        n = w.numel()
        _native.call("fp3d_multi_dot", n, nv, ptr(V), V.stride(0), ptr(w),
                     ptr(self.work), ptr(self.out[at:]), stream_handle())


def norm2(x: torch.Tensor, red: Optional[_Red] = None) -> float:
    red = red or _Red(1)
    red.norm(x)
    return float(red.out[0].item())


def dot(a: torch.Tensor, b: torch.Tensor, red: Optional[_Red] = None) -> float:
    red = red or _Red(1)
    _native.call("fp3d_multi_dot", a.numel(), 1, ptr(a), a.numel(), ptr(b),
                 ptr(red.work), ptr(red.out), stream_handle())
    return float(red.out[0].item())


def axpby(a: float, x: torch.Tensor, b: float, y: torch.Tensor) -> torch.Tensor:
    _native.call("fp3d_axpby", x.numel(), float(a), ptr(x), float(b), ptr(y),
                 stream_handle())
    return y


def scale(a: float, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    _native.call("fp3d_scale", x.numel(), float(a), ptr(x), ptr(y),
                 stream_handle())
    return y


_GMRES_WS: "OrderedDict" = None


# the device multi-dot / multi-axpy take at most 256 basis vectors
MAX_RESTART = 256


def gmres_buffers(n: int, restart: int, f32: bool, device, V=None) -> dict:
    """A fresh set of gmres_dev work buffers (V: an optional preallocated
    (restart + 1, n) basis, e.g. a slice of a batch's bases)."""
    nout = 2 * restart + 4
    if V is None:
        V = torch.empty((restart + 1, n), dtype=torch.float32 if f32 else torch.float64,
                        device=device)
    hbuf = [torch.empty(nout, dtype=torch.float64, pin_memory=True) for _ in range(2)]
    evs = [torch.cuda.Event(), torch.cuda.Event()]
    # hbuf / evs hold, per parity, where the last queued step's coefficients
    # land and the event that marks them: the solve's own buffers, or a row
    # of a batched round's shared buffer (batched_arnoldi); *_own keep the
    # solve's own, which a step queued by itself returns to
    return {"V": V, "v64": empty(n) if f32 else None, "tmp": empty(n), "r": empty(n),
            "w": [empty(n), empty(n)], "outs": [empty(nout), empty(nout)],
            "hbuf": list(hbuf), "evs": list(evs), "hbuf_own": hbuf, "evs_own": evs,
            "red": _Red(restart + 1),
            "ybuf": torch.empty(restart + 1, dtype=torch.float64, pin_memory=True),
            "ydev": empty(restart + 1)}


def _gmres_workspace(n: int, restart: int, f32: bool, device):
    """Work buffers of gmres_dev for one (n, restart, basis type), kept
    between calls (at most four shapes): the same buffers every call, so a
    preconditioner that replays HIP graphs captured on its callers' buffers
    keeps replaying them instead of copying in and out."""
    global _GMRES_WS
    from collections import OrderedDict

    if _GMRES_WS is None:
        _GMRES_WS = OrderedDict()
    key = (str(device), int(n), int(restart), bool(f32))
    ws = _GMRES_WS.get(key)
    if ws is None:
        ws = gmres_buffers(n, restart, f32, device)
        _GMRES_WS[key] = ws
        while len(_GMRES_WS) > 4:
            _GMRES_WS.popitem(last=False)
    else:
        _GMRES_WS.move_to_end(key)
    return ws


def drive(gen):
    """Run a solver coroutine whose device steps need no outside executor
    (it yields nothing) and return its result."""
    try:
        req = next(gen)
    except StopIteration as stop:
        return stop.value
    raise RuntimeError(f"solver coroutine yielded {req!r} without an executor")


class ArnoldiStep:
    """One Arnoldi step of a stepped GMRES (gmres_steps with
    pipelined=False), handed to the executor: run() queues the step's device
    work (enqueue(col)), ending in the copy of its coefficients to pinned
    host memory and an event the coroutine waits on before it reads them."""

    __slots__ = ("enqueue", "col", "matvec", "psolve", "gw", "f32")

    def __init__(self, enqueue, col, matvec, psolve, gw, f32):
        self.enqueue, self.col = enqueue, col
        self.matvec, self.psolve, self.gw, self.f32 = matvec, psolve, gw, f32

    def run(self):
        self.enqueue(self.col)


_JOB_FIELDS = 16       # int64 words per fp3d_batch job (include/fp3d.h)


def batch_key(step: ArnoldiStep):
    """Arnoldi steps that one batched launch can serve share this key: a
    Newton-Jacobian matvec on the same mesh store and value format, the
    same basis type and shape.  None for any other operator."""
    op = getattr(step.matvec, "__self__", None)
    if not isinstance(op, NewtonOperator):
        return None
    V = step.gw["V"]
    return (id(op.store), op.fmt(), step.f32, tuple(V.shape))


def batched_arnoldi(steps) -> int:
    """Queue the Arnoldi steps of several stepped GMRES solves that share a
    batch_key: the Jacobian matvecs in one batched launch per kernel, the
    solves' preconditioners (AMG V-cycles of equally shaped hierarchies in
    one launch per V-cycle operation, amg.apply_preconditioners), then the
    orthogonalisation, norm and normalisation in one batched launch per
    kernel, then each step's coefficient copy to pinned host memory.  Per
    solve this is the device work ArnoldiStep.run queues, in the same
    order, with the same kernel bodies and grids.  Returns the number of
    V-cycles that ran grouped."""
    s0 = steps[0]
    op0 = s0.matvec.__self__
    st = op0.store
    n = st.n_nodes
    f32 = s0.f32
    nb = len(steps)
    # the round's coefficients: one device block (row i = step i's out), one
    # copy to one pinned block, one event; each solve reads its row
    nout = s0.gw["outs"][0].numel()
    rdev = torch.empty((nb, nout), dtype=torch.float64, device=st.device)
    rhost = torch.empty((nb, nout), dtype=torch.float64, pin_memory=True)
    jobs = np.zeros((nb, _JOB_FIELDS), dtype=np.int64)
    max_active = 0
    max_nv = 0
    for i, stp in enumerate(steps):
        op = stp.matvec.__self__
        gw = stp.gw
        col = stp.col
        V = gw["V"]
        o = rdev[i]
        w = gw["w"][col % 2]
        na = op.n_active
        max_active = max(max_active, na)
        max_nv = max(max_nv, col + 1)
        x = gw["v64"] if f32 else V[col]
        jobs[i] = (ptr(V), V.stride(0), col + 1, ptr(w), ptr(o), ptr(gw["red"].work),
                   ptr(V[col + 1]), ptr(V[col]) if f32 else 0, ptr(x), ptr(gw["tmp"]),
                   ptr(op.data32 if op.is_f32 else op.data),
                   ptr(op.upstream) if na > 0 else 0, ptr(op.t3) if na > 0 else 0, na,
                   ptr(op.free_mask), 0)
    host = torch.from_numpy(jobs).pin_memory()
    dev = torch.empty_like(host, device=st.device)
    dev.copy_(host, non_blocking=True)
    sh = stream_handle()
    i16 = getattr(st, "indices16", None)
    _native.call("fp3d_batch_jacobian_matvec", len(steps), ptr(dev), max_active, n, st.nnz,
                 ptr(st.indptr), ptr(i16 if i16 is not None else st.indices), op0.fmt(),
                 ptr(st.el), st.n_tets, 1 if f32 else 0, sh)
    pre = []
    for stp in steps:
        gw = stp.gw
        w = gw["w"][stp.col % 2]
        if stp.psolve is None:
            w.copy_(gw["tmp"])
        else:
            pre.append((stp.psolve, gw["tmp"], w))
    # the solves' V-cycles, grouped where their hierarchies have the same
    # shape (one launch per V-cycle operation for all of them)
    n_grouped = apply_preconditioners(pre)
    _native.call("fp3d_batch_orthonormalise", len(steps), ptr(dev), max_nv, n,
                 1 if f32 else 0, 0.5, stream_handle())
    rhost.copy_(rdev, non_blocking=True)
    ev = torch.cuda.Event()
    ev.record()
    for i, stp in enumerate(steps):
        gw = stp.gw
        p = stp.col % 2
        gw["hbuf"][p] = rhost[i]
        gw["evs"][p] = ev
    return n_grouped


def _arnoldi_args(matvec, psolve, gw, n, f32):
    """(fp3d_arnoldi_step argument words, stream) for a Newton-Jacobian or a
    plain float64 CSR matvec and an AMG (or no) preconditioner, else None
    (the step is then issued launch by launch)."""
    op = getattr(matvec, "__self__", None)
    func = getattr(matvec, "__func__", None)
    plain = isinstance(op, DeviceCSR) and func is DeviceCSR.matvec and not op.is_f32
    if not plain and (not isinstance(op, NewtonOperator) or func is not NewtonOperator.matvec):
        return None
    gexec, pre = [0, 0], 0
    if psolve is not None:
        amg = getattr(psolve, "__self__", None)
        if not isinstance(amg, AMGHierarchy) or \
                getattr(psolve, "__func__", None) is not AMGHierarchy.apply:
            return None
        hs = amg.graph_execs(gw["tmp"], gw["w"])
        if hs is None:
            return None
        gexec, pre = hs, 1
    V = gw["V"]
    head = [n, 1 if f32 else 0, ptr(V), V.stride(0), ptr(gw["v64"]) if f32 else 0,
            ptr(gw["tmp"]), ptr(gw["w"][0]), ptr(gw["w"][1]), ptr(gw["outs"][0]),
            ptr(gw["outs"][1]), ptr(gw["red"].work), gw["hbuf_own"][0].data_ptr(),
            gw["hbuf_own"][1].data_ptr()]
    if plain:
        # y = A x, no Term 3, no Dirichlet rows (the matrix carries them)
        mv = [op.nnz, ptr(op.indptr), ptr(op.indices), ptr(op.data), 0, 0, 0, 0, 0, 0, 0]
    else:
        st = op.store
        na = op.n_active
        use_t3 = na > 0
        i16 = getattr(st, "indices16", None)
        vals = op.data32 if op.is_f32 else op.data
        mv = [st.nnz, ptr(st.indptr), ptr(i16 if i16 is not None else st.indices), ptr(vals),
              op.fmt(), ptr(st.el), ptr(op.upstream) if use_t3 else 0,
              ptr(op.t3) if use_t3 else 0, na, st.n_tets, ptr(op.free_mask)]
    words = head + mv + [gexec[0], gexec[1], pre]
    return np.array(words, dtype=np.int64), torch.cuda.current_stream()


def gmres_dev(matvec: Callable, b: torch.Tensor, psolve: Optional[Callable] = None,
              x0: Optional[torch.Tensor] = None, rtol: float = 1e-5,
              atol: float = 0.0, restart: int = 20, maxiter: Optional[int] = None,
              callback: Optional[Callable] = None, basis: str = "f64"):
    """Solve A x = b on the device.  matvec(x, y) / psolve(x, y) write into y.
    Returns (x, info, n_inner) with SciPy's info convention (left
    preconditioning, restart / ptol logic of scipy.sparse.linalg.gmres).

    The Arnoldi loop runs one step ahead of the host: each step's
    orthogonalisation (CGS plus a device-side DGKS-conditional second pass)
    and its normalisation (the norm read on the device) are queued, the
    step's coefficients go to pinned host memory asynchronously, and the
    host applies the Givens rotations of step j while the device already
    computes step j + 1 (a step queued past convergence is discarded).
    basis="f32" stores the Krylov basis in float32 (half the
    orthogonalisation traffic); dot products, the Hessenberg matrix, the
    Givens rotations and the solution update stay float64."""
    return drive(gmres_steps(matvec, b, psolve, x0, rtol, atol, restart, maxiter,
                             callback, basis))


def gmres_steps(matvec: Callable, b: torch.Tensor, psolve: Optional[Callable] = None,
                x0: Optional[torch.Tensor] = None, rtol: float = 1e-5,
                atol: float = 0.0, restart: int = 20, maxiter: Optional[int] = None,
                callback: Optional[Callable] = None, basis: str = "f64",
                gw: Optional[dict] = None, pipelined: bool = True):
    """gmres_dev as a coroutine.  pipelined=True runs the one-step-ahead
    loop itself and yields nothing; pipelined=False yields an ArnoldiStep
    wherever that loop queues a step, and resumes once its executor has
    queued it (an executor can queue the steps of many solves, batched, and
    then let each solve wait for its own previous step, so the device works
    on the new steps while the host does the Givens updates).  Both queue
    the same device work in the same order per solve, so they return the
    same bits.  gw: work buffers from gmres_buffers (default: the shared
    per-shape set)."""
    if basis not in ("f64", "f32"):
        raise ValueError(f"basis={basis!r}: 'f64' or 'f32'")
    if not 1 <= int(restart) <= MAX_RESTART:
        raise ValueError(f"restart={restart}: 1 to {MAX_RESTART} (the basis vectors one "
                         "multi-dot launch orthogonalises against)")
    f32 = basis == "f32"
    n = b.numel()
    if maxiter is None:
        maxiter = 10 * n
    restart = int(min(restart, n))
    if gw is None:
        gw = _gmres_workspace(n, restart, f32, b.device)
    red = gw["red"]
    tmp = gw["tmp"]
    r = gw["r"]

    def prec(src, dst):
        if psolve is None:
            dst.copy_(src)
        else:
            psolve(src, dst)
        return dst

    bnrm2 = norm2(b, red)
    atol = max(float(atol), float(rtol) * bnrm2)
    x = zeros(n) if x0 is None else x0.clone()
    if bnrm2 == 0.0:
        return zeros(n), 0, 0
    eps = np.finfo(np.float64).eps
    Mb_nrm2 = norm2(prec(b, tmp), red)
    ptol_max_factor = 1.0
    ptol = Mb_nrm2 * min(ptol_max_factor, atol / bnrm2)
    presid = 0.0
    V = gw["V"]
    v64 = gw["v64"]
    w_buf = gw["w"]
    outs = gw["outs"]
    hbuf = gw["hbuf"]
    evs = gw["evs"]
    h = np.zeros((restart, restart + 1))
    givens = np.zeros((restart, 2))
    # the rotation chain of each new column runs in the library's host code
    apply_givens = _native.lib().fp3d_host_apply_givens
    h_ptr, h_row, g_ptr = h.ctypes.data, h.strides[0], givens.ctypes.data
    inner = 0
    rnorm = np.inf
    x_nonzero = x0 is not None and bool(torch.any(x != 0).item())

    # Newton-Jacobian matvec with an AMG V-cycle (or none): each step in one
    # library call (fp3d_arnoldi_step, the same launches as below), set up at
    # the first step, once the preconditioner has run eagerly
    fast = [None, pipelined]
    arnoldi = _native.lib().fp3d_arnoldi_step

    def enqueue(col):
        """Device work of Arnoldi step col: w = M^-1 A v_col orthogonalised
        against v_0..v_col and normalised into v_{col+1}; its coefficients
        (out[2:2+k] = h, out[2+k] = ||w||^2 before, out[3+k:3+2k] = second
        pass, out[0] = ||w|| after) copied to pinned host memory."""
        if fast[1]:
            fast[0] = _arnoldi_args(matvec, psolve, gw, n, f32)
            fast[1] = False
        p = col % 2
        hbuf[p] = gw["hbuf_own"][p]
        evs[p] = gw["evs_own"][p]
        if fast[0] is not None:
            a, stream = fast[0]
            _native.check(arnoldi(a.ctypes.data, col, stream.cuda_stream), "arnoldi_step")
            evs[p].record(stream)
            return
        k = col + 1
        o = outs[col % 2]
        w = w_buf[col % 2]
        if f32:
            _native.call("fp3d_f32_to_f64", n, ptr(V[col]), ptr(v64), stream_handle())
            matvec(v64, tmp)
        else:
            matvec(V[col], tmp)
        prec(tmp, w)
        sh = stream_handle()
        if f32:
            dot_, reo_, axpy_ = ("fp3d_multi_dot_norm_f32v", "fp3d_multi_dot_reorth_f32v",
                                 "fp3d_multi_axpy_f32v")
        else:
            dot_, reo_, axpy_ = ("fp3d_multi_dot_norm", "fp3d_multi_dot_reorth",
                                 "fp3d_multi_axpy")
        # h = V^T w into out[2:2+k] and ||w||^2 into out[2+k], one pass over w
        _native.call(dot_, n, k, ptr(V), V.stride(0), ptr(w), ptr(red.work),
                     ptr(o[2:]), sh)
        _native.call(axpy_, n, k, ptr(V), V.stride(0), ptr(o[2:]), -1.0, ptr(w), sh)
        # second CGS pass only where the first one cancelled (device-side DGKS
        # test on h and ||w||^2; a skipped pass leaves exact zeros and w as is)
        _native.call(reo_, n, k, ptr(V), V.stride(0), ptr(w), ptr(o[2:]), 0.5,
                     ptr(red.work), ptr(o[3 + k:]), sh)
        _native.call(axpy_, n, k, ptr(V), V.stride(0), ptr(o[3 + k:]), -1.0, ptr(w), sh)
        _native.call("fp3d_norms", n, ptr(w), ptr(red.work), ptr(o), sh)
        _native.call("fp3d_scale_inv_dev_to_f32" if f32 else "fp3d_scale_inv_dev", n,
                     ptr(o), ptr(w), ptr(V[col + 1]), sh)
        hbuf[col % 2][: 3 + 2 * k].copy_(o[: 3 + 2 * k], non_blocking=True)
        evs[col % 2].record()

    for iteration in range(maxiter):
        if iteration == 0:
            if x_nonzero:
                matvec(x, tmp)
                r.copy_(b)
                axpby(-1.0, tmp, 1.0, r)
            else:
                r.copy_(b)
            if norm2(r, red) < atol:
                return x, 0, 0
        w0 = w_buf[1]
        prec(r, w0)
        t0 = norm2(w0, red)
        if f32:
            _native.call("fp3d_scale_to_f32", n, 1.0 / t0, ptr(w0), ptr(V[0]),
                         stream_handle())
        else:
            scale(1.0 / t0, w0, V[0])
        S = np.zeros(restart + 1)
        S[0] = t0
        breakdown = False
        col = 0
        if pipelined:
            enqueue(0)
        else:
            yield ArnoldiStep(enqueue, 0, matvec, psolve, gw, f32)
        for col in range(restart):
            if col + 1 < restart:           # runs while the host does step col
                if pipelined:
                    enqueue(col + 1)
                else:
                    yield ArnoldiStep(enqueue, col + 1, matvec, psolve, gw, f32)
            evs[col % 2].synchronize()
            k = col + 1
            host = hbuf[col % 2][: 3 + 2 * k].numpy()
            h0 = float(np.sqrt(max(host[2 + k], 0.0)))
            h1 = float(host[0])
            hv = host[2:2 + k] + host[3 + k:3 + 2 * k]
            h[col, :k] = hv
            h[col, col + 1] = h1
            if h1 <= eps * h0:
                h[col, col + 1] = 0.0
                breakdown = True
            if col:
                _native.check(apply_givens(col, h_ptr + col * h_row, g_ptr),
                              "apply_givens")
            c, s_, mag = _lartg(h[col, col], h[col, col + 1])
            givens[col, 0], givens[col, 1] = c, s_
            h[col, col], h[col, col + 1] = mag, 0.0
            t = -s_ * S[col]
            S[col], S[col + 1] = c * S[col], t
            presid = abs(t)
            inner += 1
            if callback is not None:
                callback(presid / bnrm2)
            if presid <= ptol or breakdown:
                break
        if h[col, col] == 0.0:
            S[col] = 0.0
        y = np.zeros(col + 1)
        y[:] = S[:col + 1]
        for k2 in range(col, 0, -1):
            if y[k2] != 0.0:
                y[k2] /= h[k2, k2]
                y[:k2] -= y[k2] * h[k2, :k2]
        if y[0] != 0.0:
            y[0] /= h[0, 0]
        # the solution coefficients through the workspace's pinned buffer
        # (rewritten only at the next restart, after the residual norm below
        # has synchronised the stream)
        yh = gw["ybuf"][: col + 1]
        yh.numpy()[:] = y
        yd = gw["ydev"][: col + 1]
        yd.copy_(yh, non_blocking=True)
        _native.call("fp3d_multi_axpy_f32v" if f32 else "fp3d_multi_axpy", n, col + 1,
                     ptr(V), V.stride(0), ptr(yd), 1.0, ptr(x), stream_handle())
        matvec(x, tmp)
        r.copy_(b)
        axpby(-1.0, tmp, 1.0, r)
        rnorm = norm2(r, red)
        if rnorm <= atol:
            break
        elif breakdown:
            break
        elif presid <= ptol:
            ptol_max_factor = max(eps, 0.25 * ptol_max_factor)
        else:
            ptol_max_factor = min(1.0, 1.5 * ptol_max_factor)
        ptol = presid * min(ptol_max_factor, atol / rnorm)
    info = 0 if rnorm <= atol else maxiter
    return x, info, inner


def pcg_dev(matvec: Callable, b: torch.Tensor, psolve: Optional[Callable] = None,
            x0: Optional[torch.Tensor] = None, rtol: float = 1e-5,
            atol: float = 0.0, maxiter: Optional[int] = None):
    """Preconditioned CG (scipy.sparse.linalg.cg semantics: stop when
    ||r|| <= max(atol, rtol ||b||)).  Returns (x, info, n_iter)."""
    n = b.numel()
    maxiter = 10 * n if maxiter is None else maxiter
    red = _Red(1)
    bnrm2 = norm2(b, red)
    if bnrm2 == 0.0:
        return zeros(n), 0, 0
    atol = max(float(atol), float(rtol) * bnrm2)
    x = zeros(n) if x0 is None else x0.clone()
    r = b.clone()
    q = empty(n)
    if x0 is not None:
        matvec(x, q)
        axpby(-1.0, q, 1.0, r)
    z = empty(n)
    p = empty(n)
    rho_prev = None
    for it in range(maxiter):
        if norm2(r, red) < atol:
            return x, 0, it
        if psolve is None:
            z.copy_(r)
        else:
            psolve(r, z)
        rho_cur = dot(r, z, red)
        if rho_prev is None:
            p.copy_(z)
        else:
            axpby(1.0, z, rho_cur / rho_prev, p)
        matvec(p, q)
        alpha = rho_cur / dot(p, q, red)
        axpby(alpha, p, 1.0, x)
        axpby(-alpha, q, 1.0, r)
        rho_prev = rho_cur
    return x, (0 if norm2(r, red) < atol else maxiter), maxiter


# ----------------------------------------------------- Newton Jacobian operator
class NewtonOperator:
    """x -> J_ff x on full-length vectors with zero Dirichlet entries: the
    Terms 1+2 CSR (Dirichlet rows/cols masked to identity, float64 or
    float32 values) plus the matrix-free Term 3 of the active elements."""

    def __init__(self, store, free_mask: torch.Tensor):
        self.store = store
        self.free_mask = free_mask
        self.n = store.n_nodes
        self.data = empty(store.nnz)
        self.data32 = None
        self.is_f32 = 0
        nbytes = _native.lib().fp3d_term3_workspace_bytes(store.n_tets,
                                                          store.n_nodes)
        self.t3 = torch.empty(int(nbytes), dtype=torch.uint8,
                              device=store.device)
        self._n_active: Optional[int] = 0
        self.upstream = None

    def finalize(self, f32: bool = False) -> None:
        """Mask Dirichlet rows/cols in place; optionally keep a float32 copy
        that the matvec then uses."""
        st = self.store
        if f32 and self.data32 is None:
            self.data32 = torch.empty(st.nnz, dtype=torch.float32,
                                      device=st.device)
        _native.call("fp3d_csr_mask_dirichlet", st.n_nodes, ptr(st.indptr),
                     ptr(st.indices), ptr(self.free_mask), ptr(self.data),
                     ptr(self.data32) if f32 else None, stream_handle())
        self.is_f32 = 1 if f32 else 0

    @property
    def n_active(self) -> int:
        """Active Term-3 elements of the last build (read from the device on
        first use after a build: one stream synchronisation)."""
        if self._n_active is None:
            import ctypes as C

            st = self.store
            na = C.c_int64(0)
            _native.check(_native.lib().fp3d_term3_active_count(
                ptr(self.t3), st.n_tets, st.n_nodes, C.byref(na), stream_handle()),
                "term3_active_count")
            self._n_active = int(na.value)
        return self._n_active

    @n_active.setter
    def n_active(self, value: int) -> None:
        self._n_active = int(value)

    def build_term3(self, grad, upstream, s_u, lim, inv_u2, gdotb=None) -> None:
        """Select the active elements and build the Term-3 row table on the
        stream (no synchronisation; n_active is read when first needed).
        gdotb: grad . B of the same state (state sweep output), optional."""
        st = self.store
        _native.check(_native.lib().fp3d_term3_build(
            st.n_tets, st.n_nodes, ptr(st.el), ptr(st.inc_ptr), ptr(st.inc_ent),
            ptr(st.inc_pos), ptr(st.B), ptr(st.V), ptr(grad),
            ptr(gdotb) if gdotb is not None else None, ptr(upstream), ptr(s_u),
            ptr(lim), ptr(self.free_mask), float(inv_u2), ptr(self.t3), None,
            stream_handle()), "term3_build")
        self._n_active = None
        self.upstream = upstream

    def fmt(self) -> int:
        """Value / index format bits of the Terms 1+2 CSR the matvec reads
        (1: float32 values, 4: int16 row-relative columns)."""
        return self.is_f32 | (4 if getattr(self.store, "indices16", None) is not None else 0)

    def matvec(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        st = self.store
        vals = self.data32 if self.is_f32 else self.data
        use_t3 = self.n_active > 0
        i16 = getattr(st, "indices16", None)
        fmt = self.fmt()
        _native.call("fp3d_csr_spmv_term3", st.n_nodes, st.nnz, ptr(st.indptr),
                     ptr(i16 if i16 is not None else st.indices), ptr(vals), fmt,
                     ptr(st.el),
                     ptr(self.upstream) if use_t3 else None,
                     ptr(self.t3) if use_t3 else None, self.n_active,
                     st.n_tets, ptr(self.free_mask), ptr(x), ptr(y),
                     stream_handle())
        return y


# ------------------------------------------- reference-compatible API
def apply_dirichlet(A: sp.spmatrix, b: np.ndarray, dirichlet_nodes: np.ndarray,
                    dirichlet_values: np.ndarray):
    """Free-dof principal-submatrix reduction (pyfp3d/solve/linear.py:20-58);
    host bookkeeping, returns (A_free, b_free, free_idx, phi_dirichlet)."""
    n = A.shape[0]
    is_dir = np.zeros(n, dtype=bool)
    is_dir[dirichlet_nodes] = True
    free = np.where(~is_dir)[0]
    phi_d = np.zeros(n)
    phi_d[dirichlet_nodes] = dirichlet_values
    A = A.tocsr()
    b_free = b[free] - A[free][:, dirichlet_nodes] @ dirichlet_values
    return A[free][:, free].tocsr(), b_free, free, phi_d


class _HostOperatorOnDevice:
    """Wrap a scipy matrix (or an object with .matvec) as a device matvec."""

    def __init__(self, A, f32: bool = False):
        if sp.issparse(A):
            self.csr = DeviceCSR(A, f32=f32)
            self.mv = self.csr.matvec
            self.shape = A.shape
        elif hasattr(A, "matvec") and hasattr(A, "shape"):
            self.csr = None
            self.shape = A.shape
            op = A

            def mv(x, y):
                y.copy_(to_dev(np.asarray(op.matvec(to_host(x)), dtype=np.float64)))
                return y

            self.mv = mv
        else:
            raise TypeError(f"unsupported operator {type(A)!r}")


def build_amg_preconditioner(A: sp.spmatrix, precision: str = "f64"):
    """(ml, M): SA hierarchy (host set-up, seed pinned) and its device
    preconditioner (pyfp3d/solve/linear.py:61-81)."""
    ml = AMGHierarchy(sp.csr_matrix(A), precision=precision, seed=0)
    return ml, ml


def _psolve_of(M):
    if M is None:
        return None
    if isinstance(M, AMGHierarchy):
        return M.apply
    if hasattr(M, "matvec"):
        def ps(x, y):
            y.copy_(to_dev(np.asarray(M.matvec(to_host(x)), dtype=np.float64)))
            return y
        return ps
    raise TypeError(f"unsupported preconditioner {type(M)!r}")


def solve_gmres(A, b, M=None, rtol: float = 1e-6, atol: float = 0.0,
                x0=None, restart: int = 60, maxiter: int = 50,
                on_fail: str = "retry") -> Tuple[np.ndarray, int, int]:
    """Restarted preconditioned GMRES on the device with the reference's
    retry-at-2x-restart policy (pyfp3d/solve/linear.py:134-201).
    A: scipy sparse matrix, an object with .matvec, or a device callable
    matvec(x, y).  Returns (x, n_inner_iterations, info); RuntimeError under
    on_fail='retry' when the retry also fails."""
    host = not isinstance(b, torch.Tensor)
    bd = to_dev(b, torch.float64) if host else b
    mv = A if callable(A) and not sp.issparse(A) and not hasattr(A, "matvec") \
        else _HostOperatorOnDevice(A).mv
    ps = _psolve_of(M)
    x0d = None if x0 is None else (to_dev(x0, torch.float64)
                                   if not isinstance(x0, torch.Tensor) else x0)
    x, info, n_it = gmres_dev(mv, bd, ps, x0d, rtol, atol, restart, maxiter)
    if info != 0 and on_fail == "retry":
        n_first = n_it
        x, info, n_it = gmres_dev(mv, bd, ps, x, rtol, atol, 2 * restart, maxiter)
        if info != 0:
            raise RuntimeError(
                f"GMRES did not converge (info={info}) after {n_first} + {n_it} "
                f"inner iterations (restart {restart} then {2 * restart}, "
                f"rtol={rtol:g})")
        n_it += n_first
    return (to_host(x) if host else x), n_it, info


def solve_cg_amg(A: sp.spmatrix, b, rtol: float = 1e-10, maxiter: int = 500):
    """CG preconditioned by SA-AMG on the device (pyfp3d/solve/linear.py:204-239).
    Returns (x, n_iterations); RuntimeError when CG does not converge."""
    host = not isinstance(b, torch.Tensor)
    _, M = build_amg_preconditioner(A)
    Ad = DeviceCSR(sp.csr_matrix(A))
    bd = to_dev(b, torch.float64) if host else b
    x, info, n_it = pcg_dev(Ad.matvec, bd, M.apply, rtol=rtol, maxiter=maxiter)
    if info != 0:
        raise RuntimeError(f"CG did not converge (info={info}) after {n_it} iterations")
    return (to_host(x) if host else x), n_it
