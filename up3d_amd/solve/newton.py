"""Newton solve of the non-lifting full-potential problem on the device
(pyfp3d/solve/newton.py:97-1307, specialised to a case without wake
stations: the Gamma unknowns, B, K and F of the coupled system vanish and
the eliminated step is J_ff dphi = -R_free).  The reference's conforming
driver needs a WakeCut and cannot run a non-lifting case as shipped, so
this is the same chain with n_st = 0, with identical keyword names,
defaults, control logic and result keys.

Per Newton step (all device-resident):
  eval_residual: fused grad/q2/limiter/density sweep -> walk + nu switch +
      rho_tilde (live) or frozen sweep -> node-gather residual -> minus
      the external (boundary-flux) RHS -> ||R_free||
  assemble: sensitivities (s_e, s_u) -> Terms 1+2 node-gathered into the
      fixed CSR -> Dirichlet embedding (+ float32 copy) -> Term 3 gather
      table
  precond: SA-AMG of the Picard matrix A(rho_tilde) every amg_rebuild_every
      steps (host set-up, device V-cycle)
  linsolve: device GMRES(restart) with Eisenstat-Walker forcing
  line search: safety-only backtracking on ||R||^2, best-of on failure.

precision="fp64" is the all-FP64 solve; precision="fp32" keeps the
residual, norms, the Arnoldi dot products / Hessenberg / Givens arithmetic
and the coarsest AMG solve in float64, and moves the GMRES Jacobian values,
every AMG matrix and the storage of the Krylov basis to float32; the outer
Newton iteration on the float64 residual acts as iterative refinement.
"""

from __future__ import annotations

import time
from typing import Dict, Optional

import numpy as np
import torch

from .. import _native
from ..device import empty, ptr, stream_handle, to_dev, to_host
from ..kernels.jacobian import PicardOperator
from ..kernels.upwind import UpwindOperator
from ..physics.isentropic import q2_at_mach
from .common import (Reducer, amg_plan_from_device, dirichlet_setup, mask_dirichlet,
                     masked)
from .continuation import mach_schedule
from .linear import (NewtonOperator, axpby, drive, gmres_buffers, gmres_steps,
                     scale)
from .timing import finalize, new_timings, snapshot, step_delta, sum_timings


class _State:
    """Device buffers of one evaluated state (phi, element state, R)."""

    def __init__(self, n, nt):
        self.phi = empty(n)
        self.grad = empty((nt, 3))
        self.gb = empty(4 * nt)    # grad . B_a per element (state sweep)
        self.q2n = empty(nt)
        self.q2l = empty(nt)
        self.rho = empty(nt)
        self.rt = empty(nt)
        self.lim = empty(nt, torch.uint8)
        self.up = empty(nt, torch.int32)   # upstream of the live walk
        self.se = empty(nt)        # exact s_e, s_u of that walk's selection
        self.su = empty(nt)
        self.sens_key = None       # (M, C, M_crit, gamma, floor) they are for
        self.R = empty(n)          # free-masked residual
        self.n_limited = 0
        self.n_floored = 0
        self.nu_max = 0.0
        self.n_nu_active = 0
        self.r_norm = 0.0
        self.merit = 0.0


class NewtonWorkspace:
    """Per-case device workspace (pyfp3d/solve/newton.py:97-501 for
    n_st = 0): operators, the free/Dirichlet split, the external RHS, state
    buffers, the Jacobian operator and the preconditioner."""

    def __init__(self, nodes, elements, dirichlet_nodes, dirichlet_values,
                 u_inf: float = 1.0, gamma_air: float = 1.4,
                 external_rhs: Optional[np.ndarray] = None,
                 precision: str = "fp64"):
        if precision not in ("fp64", "fp32"):
            raise ValueError(f"unknown precision {precision!r}")
        self.precision = precision
        self.op = PicardOperator(nodes, elements)
        self.upw = UpwindOperator(nodes, elements, weighted=False)
        st = self.st = self.op.store
        self.n = n = st.n_nodes
        self.nt = nt = st.n_tets
        self.u_inf = float(u_inf)
        self.gamma_air = float(gamma_air)
        self.fm, self.dn, self.dv, self.phi_d = dirichlet_setup(
            n, dirichlet_nodes, dirichlet_values, st.device)
        self.n_free = int(n - len(np.unique(np.asarray(dirichlet_nodes))))
        self.b_ext = None
        if external_rhs is not None:
            e = np.asarray(external_rhs, dtype=np.float64)
            if e.shape != (n,):
                raise ValueError(f"external_rhs must be ({n},), got {e.shape}")
            self.b_ext = to_dev(e)
        self.states = [_State(n, nt) for _ in range(3)]
        self.jac = NewtonOperator(st, self.fm)
        self.A_pic = empty(st.nnz)
        self.red = Reducer()
        self.work = empty(_native.lib().fp3d_reduce_work_doubles())
        self.cnt = empty(1)
        self.m_inf = None
        self.q2cap = np.inf
        self.precond = None
        self._ptc = None

    def ptc_weights(self):
        """(lumped nodal volume, CSR diagonal positions) on the device."""
        if self._ptc is None:
            st = self.st
            m = empty(self.n)
            _native.call("fp3d_lumped_volume", self.n, ptr(st.inc_ptr),
                         ptr(st.inc_ent), ptr(st.V), ptr(m), stream_handle())
            ip, ix = st.indptr_host, st.indices_host
            rows = np.repeat(np.arange(self.n), np.diff(ip))
            diag = np.flatnonzero(ix == rows)
            if len(diag) != self.n:
                raise RuntimeError("CSR pattern lacks a diagonal entry")
            self._ptc = (m, to_dev(diag.astype(np.int32), torch.int32))
        return self._ptc

    def _diag_shift(self, data, scale):
        if scale > 0.0:
            m, diag = self.ptc_weights()
            _native.call("fp3d_csr_diag_shift", self.n, ptr(diag), ptr(m),
                         float(scale), ptr(self.fm), ptr(data), stream_handle())

    def set_mach(self, m_inf: float, m_cap: float = 3.0) -> None:
        if not 0.0 <= m_inf < 1.0:
            raise ValueError(f"Newton driver needs 0 <= M_inf < 1, got {m_inf}")
        self.m_inf = float(m_inf)
        self.q2cap = float(q2_at_mach(m_cap, m_inf, self.gamma_air))

    def gmres_work(self, restart: int, f32: bool) -> dict:
        """GMRES buffers of this workspace alone (stepped solves of several
        workspaces run side by side, so they cannot share the per-shape
        set)."""
        key = (int(min(restart, self.n)), bool(f32))
        gws = getattr(self, "_gmres_gw", None)
        if gws is None or gws[0] != key:
            self._gmres_gw = gws = (key, gmres_buffers(self.n, key[0], key[1],
                                                       self.st.device))
        return gws[1]

    def free_state(self, *busy):
        for s in self.states:
            if all(s is not b for b in busy):
                return s
        raise RuntimeError("no free state buffer")

    # ------------------------------------------------------- residual path
    def eval_residual(self, s: _State, upwind_c, m_crit, rho_floor,
                      frozen=None) -> _State:
        """The single state-reconstruction + residual path; s.phi must hold
        the full potential (Dirichlet entries set)."""
        op, upw = self.op, self.upw
        op.state_dev(s.phi, self.u_inf ** 2, self.m_inf, self.gamma_air,
                     self.q2cap, s.grad, s.q2n, s.q2l, s.rho, s.lim, gdotb=s.gb)
        _native.call("fp3d_count_limited", self.nt, ptr(s.lim), ptr(self.work),
                     ptr(self.cnt), stream_handle())
        s.n_limited = int(self.cnt.item())
        s.sens_key = None
        if self.m_inf > 0.0 and frozen is not None:
            upw.frozen_dev(s.q2l, s.rho, frozen[0], frozen[1], self.m_inf,
                           upwind_c, m_crit, self.gamma_air, rho_floor)
            s.rt.copy_(upw.d_rt)
        elif self.m_inf > 0.0:
            # the walk's sweep also keeps the exact sensitivities of its
            # selection, which assemble_coupled reuses for this state
            upw.rho_tilde_dev(s.grad, s.q2l, s.rho, self.m_inf, upwind_c, m_crit,
                              self.gamma_air, rho_floor, up_out=s.up,
                              rt_out=s.rt, se_out=s.se, su_out=s.su)
            s.sens_key = (self.m_inf, upwind_c, m_crit, self.gamma_air, rho_floor)
        else:
            s.rt.copy_(s.rho)
            upw.nu_max, upw.n_supersonic, upw.n_floored = 0.0, 0, 0
        s.n_floored = upw.n_floored
        s.nu_max = upw.nu_max
        s.n_nu_active = upw.n_supersonic
        R = op.residual_dev(s.rt, s.grad, gdotb=s.gb)
        if self.b_ext is not None:
            axpby(-1.0, self.b_ext, 1.0, R)
        masked(self.fm, R, s.R)
        l2, linf = self.red.norms(s.R)
        s.r_norm = linf
        s.merit = l2 * l2
        return s

    # ------------------------------------------------------- linearisation
    def assemble_coupled(self, s: _State, upwind_c, m_crit, rho_floor,
                         frozen=None, ptc_scale: float = 0.0):
        """Exact Jacobian at state s into the NewtonOperator (Terms 1+2 in
        CSR, Dirichlet-embedded, Term 3 as a gather table); ptc_scale > 0
        adds ptc_scale * (lumped nodal volume) to the free diagonal."""
        upw = self.upw
        inv_u2 = 1.0 / (self.u_inf * self.u_inf)
        if self.m_inf > 0.0 and frozen is not None:
            upstream = frozen[0]
            se, su = upw.frozen_sensitivities_dev(
                s.q2l, s.rho, frozen[0], frozen[1], self.m_inf, upwind_c,
                m_crit, self.gamma_air, rho_floor)
        elif self.m_inf > 0.0:
            # the live walk on this state's grad already ran in eval_residual
            # (pyfp3d re-walks the same grad, which yields the same u(e))
            upstream = s.up
            if s.sens_key == (self.m_inf, upwind_c, m_crit, self.gamma_air, rho_floor):
                se, su = s.se, s.su
            else:
                se, su = upw.sensitivities_at_dev(
                    s.q2l, s.rho, upstream, self.m_inf, upwind_c, m_crit,
                    self.gamma_air, rho_floor)
        else:
            se = upw.d_se.zero_()
            su = upw.d_su.zero_()
            upstream = upw.walk_dev(s.grad)
        self.se_used = se          # the preconditioner's SPD matrix reuses it
        main = side = None
        if self.m_inf > 0.0:
            # Term 3 (its own table, read-only inputs) on a second stream,
            # overlapping Terms 1+2; the main stream waits for it below
            main = torch.cuda.current_stream()
            side = self._side_stream()
            side.wait_stream(main)
            with torch.cuda.stream(side):
                self.jac.build_term3(s.grad, upstream, su, s.lim, inv_u2, gdotb=s.gb)
        else:
            self.jac.n_active = 0
        self.op.jacobian_dev(s.rt, s.grad, se, s.lim, inv_u2, self.jac.data)
        self._diag_shift(self.jac.data, ptc_scale)
        self.jac.finalize(f32=self.precision == "fp32")
        if side is not None:
            main.wait_stream(side)
        return self.jac

    def _side_stream(self):
        """Second stream of this workspace's device (default priority)."""
        if getattr(self, "_side", None) is None:
            self._side = torch.cuda.Stream(device=self.st.device)
        return self._side

    def build_preconditioner(self, s: _State, precond_matrix: str = "picard",
                             clamp: float = 0.9, se=None, ptc_scale: float = 0.0):
        """AMG of the Picard matrix A(rho_tilde) (the reference's choice) or
        of the SPD clamped Jacobian (precond_matrix="jacobian_spd")."""
        return drive_inline(self.build_preconditioner_steps(s, precond_matrix, clamp, se,
                                                            ptc_scale, stepped=False))

    def build_preconditioner_steps(self, s: _State, precond_matrix: str = "picard",
                                   clamp: float = 0.9, se=None, ptc_scale: float = 0.0,
                                   stepped: bool = False):
        """build_preconditioner as a coroutine: stepped=True yields the host
        SA set-up as a HostTask (an executor may run it on another thread
        while other solves use the device) and builds the device levels
        when resumed with its result; the same hierarchy either way."""
        self._precond_matrix_dev(s, precond_matrix, clamp, se, ptc_scale)
        plan = amg_plan_from_device(
            self.st, self.A_pic, "f32" if self.precision == "fp32" else "f64",
            **getattr(self, "amg_kw", {}))
        if stepped:
            host = yield HostTask(plan.host)
        else:
            host = plan.host()
        self.precond = plan.finish(host)
        return self.precond

    def _precond_matrix_dev(self, s: _State, precond_matrix: str, clamp: float, se,
                            ptc_scale: float) -> None:
        """The masked preconditioner matrix into self.A_pic."""
        if precond_matrix == "jacobian_spd" and se is not None:
            self.op.jacobian_spd_dev(s.rt, s.grad, se, s.lim,
                                     1.0 / (self.u_inf * self.u_inf), clamp,
                                     self.A_pic)
        elif precond_matrix in ("picard", "jacobian_spd"):
            self.op.matrix_dev(s.rt, self.A_pic)
        else:
            raise ValueError(f"unknown precond_matrix {precond_matrix!r}")
        self._diag_shift(self.A_pic, ptc_scale)
        mask_dirichlet(self.st, self.fm, self.A_pic)

    def freeze(self, s: _State, upwind_c, m_crit, rho_floor):
        return self.upw.freeze_dev(s.grad, s.q2l, s.rho, self.m_inf, upwind_c,
                                   m_crit, self.gamma_air, rho_floor)


def _ew_forcing(r_norm, r_norm_prev, eta_prev, eta0=1e-2, gamma_ew=0.9,
                eta_max=1e-2, eta_floor=1e-10):
    """Eisenstat-Walker choice 2 with the standard safeguard
    (pyfp3d/solve/newton.py:504-513)."""
    if r_norm_prev is None:
        return eta0
    eta = gamma_ew * (r_norm / r_norm_prev) ** 2
    safeguard = gamma_ew * eta_prev ** 2
    if safeguard > 0.1:
        eta = max(eta, safeguard)
    return float(min(eta_max, max(eta, eta_floor)))


def solve_newton(nodes, elements, dirichlet_nodes, dirichlet_values,
                 m_inf: float, u_inf: float = 1.0, gamma_air: float = 1.4,
                 upwind_c: float = 1.5, m_crit: float = 0.95,
                 m_cap: float = 3.0, rho_floor: float = 0.05,
                 phi_init: Optional[np.ndarray] = None, n_picard_seed: int = 5,
                 n_newton_max: int = 30, tol_residual: float = 1e-10,
                 ew_eta0: float = 1e-2, ew_gamma: float = 0.9,
                 ew_eta_max: float = 1e-2, amg_rebuild_every: int = 2,
                 gmres_restart: int = 60, gmres_maxiter: int = 10,
                 line_search: bool = True, rtol_seed: float = 1e-7,
                 tol_residual_loose: Optional[float] = None,
                 tol_residual_rel: Optional[float] = None,
                 accept_on_stall: bool = False,
                 freeze_tol: Optional[float] = None,
                 freeze_refresh_max: int = 2, freeze_max_reverts: int = 3,
                 workspace: Optional[NewtonWorkspace] = None,
                 external_rhs: Optional[np.ndarray] = None,
                 precision: str = "fp64", precond_matrix: str = "picard",
                 precond_clamp: float = 0.9, ptc_dt0: Optional[float] = None,
                 ptc_dt_max: float = 1e6, amg_degree: Optional[int] = None,
                 amg_max_coarse: Optional[int] = None, verbose: bool = False
                 ) -> Dict[str, object]:
    """Newton at ONE Mach level (pyfp3d/solve/newton.py:516-1141 with
    n_st = 0): newton_steps run to completion; see newton_steps for the
    arguments."""
    kw = dict(locals())
    return drive(newton_steps(**kw))


class EvalRequest:
    """A residual evaluation of a stepped Newton solve (newton_steps with
    stepped=True), handed to the executor, which runs
    ws.eval_residual(s, *args) -- or a batched equivalent -- and resumes the
    solve."""

    __slots__ = ("ws", "s", "args")

    def __init__(self, ws, s, args):
        self.ws, self.s, self.args = ws, s, args


class HostTask:
    """Host-only work of a stepped solve (no device or torch call inside
    fn), handed to the executor, which may run fn() on a worker thread and
    resumes the solve with its return value."""

    __slots__ = ("fn",)

    def __init__(self, fn):
        self.fn = fn


def drive_inline(gen):
    """Run a coroutine that may yield HostTasks, running each inline."""
    try:
        req = next(gen)
        while True:
            if not isinstance(req, HostTask):
                raise TypeError(f"unexpected request {type(req).__name__}")
            req = gen.send(req.fn())
    except StopIteration as stop:
        return stop.value


def _eval(ws, stepped, s, *args):
    """ws.eval_residual(s, *args), inline or through the executor."""
    if stepped:
        yield EvalRequest(ws, s, args)
    else:
        ws.eval_residual(s, *args)
    return s


def newton_steps(nodes, elements, dirichlet_nodes, dirichlet_values,
                 m_inf: float, u_inf: float = 1.0, gamma_air: float = 1.4,
                 upwind_c: float = 1.5, m_crit: float = 0.95,
                 m_cap: float = 3.0, rho_floor: float = 0.05,
                 phi_init: Optional[np.ndarray] = None, n_picard_seed: int = 5,
                 n_newton_max: int = 30, tol_residual: float = 1e-10,
                 ew_eta0: float = 1e-2, ew_gamma: float = 0.9,
                 ew_eta_max: float = 1e-2, amg_rebuild_every: int = 2,
                 gmres_restart: int = 60, gmres_maxiter: int = 10,
                 line_search: bool = True, rtol_seed: float = 1e-7,
                 tol_residual_loose: Optional[float] = None,
                 tol_residual_rel: Optional[float] = None,
                 accept_on_stall: bool = False,
                 freeze_tol: Optional[float] = None,
                 freeze_refresh_max: int = 2, freeze_max_reverts: int = 3,
                 workspace: Optional[NewtonWorkspace] = None,
                 external_rhs: Optional[np.ndarray] = None,
                 precision: str = "fp64", precond_matrix: str = "picard",
                 precond_clamp: float = 0.9, ptc_dt0: Optional[float] = None,
                 ptc_dt_max: float = 1e6, amg_degree: Optional[int] = None,
                 amg_max_coarse: Optional[int] = None, verbose: bool = False,
                 stepped: bool = False):
    """Newton at ONE Mach level (pyfp3d/solve/newton.py:516-1141 with
    n_st = 0; precond is always the device AMG).  Converged when
    ||R_free||_inf < tol_residual with 0 limited and 0 floored elements.
    amg_rebuild_every = k > 0 rebuilds the AMG hierarchy every k Newton
    steps (the reference's rule); k <= 0 selects the adaptive rule (keep
    the hierarchy, also across Mach levels of a shared workspace, until a
    Krylov solve stalls or its iteration count doubles, at most every 3
    Newton steps).  precond_matrix
    "picard" builds the AMG on A(rho_tilde) as the reference does;
    "jacobian_spd" builds it on the Newton Terms 1+2 with the streamwise
    (Term 2) weight clamped at -precond_clamp * rho_tilde V / q2 (SPD, keeps
    the near-sonic anisotropy the Picard matrix lacks).

    ptc_dt0 (off by default; an addition without a reference counterpart)
    turns on pseudo-transient continuation: each linear step solves
    (J + M_lumped / dt) dphi = -R with the switched-evolution-relaxation
    rule dt = ptc_dt0 * ||R_0||_2 / ||R||_2 (R_0 the level's first
    residual); once dt reaches ptc_dt_max the steps are plain Newton.  The
    converged state is the same discrete solution (R = 0).

    amg_degree / amg_max_coarse (None: the hierarchy's defaults, Chebyshev
    degree 2 and a coarsest level of at most 10 rows) set the AMG smoother
    degree and the size at which the hierarchy stops and solves densely --
    the knobs `up3d_amd.tune` searches.

    As a coroutine: stepped=False runs every device step inline and yields
    nothing; stepped=True yields an EvalRequest per residual evaluation and
    an ArnoldiStep per Krylov step (GMRES on the workspace's own buffers),
    so an executor can advance many solves in lockstep.  The device work
    per solve, and so the result, is the same either way."""
    if m_inf > 0.0 and upwind_c <= 0.0:
        raise ValueError("the Newton driver runs with the walk upwind machinery "
                         "active for m_inf > 0; pass upwind_c > 0")
    ws = workspace
    if ws is None:
        ws = NewtonWorkspace(nodes, elements, dirichlet_nodes, dirichlet_values,
                             u_inf, gamma_air, external_rhs, precision)
    elif ws.precision != precision:
        raise ValueError(f"workspace precision {ws.precision!r} != {precision!r}")
    ws.set_mach(m_inf, m_cap)
    amg_kw = dict(getattr(ws, "amg_kw", {}))
    if amg_degree is not None:
        amg_kw["degree"] = int(amg_degree)
    if amg_max_coarse is not None:
        amg_kw["max_coarse"] = int(amg_max_coarse)
    if amg_kw != getattr(ws, "amg_kw", {}):
        ws.amg_kw = amg_kw
        ws.precond = None             # a kept hierarchy was built with other options
    t_wall0 = time.perf_counter()
    timings = new_timings()
    step_records = []

    def sync():
        torch.cuda.synchronize()

    # ---- initial state
    t0 = time.perf_counter()
    cur = ws.states[0]
    if phi_init is not None:
        cur.phi.copy_(to_dev(phi_init, torch.float64)
                      if not isinstance(phi_init, torch.Tensor) else phi_init)
        cur.phi.index_copy_(0, ws.dn, ws.dv)
    elif n_picard_seed > 0:
        from .picard import solve_subsonic

        seed = solve_subsonic(nodes, elements, dirichlet_nodes, dirichlet_values,
                              m_inf=m_inf, gamma_air=gamma_air, u_inf=u_inf,
                              body_source_rhs=external_rhs,
                              n_picard_max=n_picard_seed, tol_rho=1e-3,
                              rtol=rtol_seed)
        cur.phi.copy_(to_dev(seed["phi"]))
    else:
        from ..constraints.dirichlet import freestream_phi

        cur.phi.copy_(to_dev(freestream_phi(ws.st.nodes_host, 0.0, u_inf)))
        cur.phi.index_copy_(0, ws.dn, ws.dv)
    sync()
    timings["seed"] = time.perf_counter() - t0

    residual_history, clamp_history, eta_history, newton_orders = [], [], [], []
    n_gmres_total = 0
    n_gmres_stalled = 0
    converged = False
    accept_reason = None
    eta_prev = None
    t0 = time.perf_counter()
    yield from _eval(ws, stepped, cur, upwind_c, m_crit, rho_floor)
    sync()
    timings["residual"] += time.perf_counter() - t0

    frozen = None
    n_freeze_refresh = 0
    residual_unfrozen = None
    n_assignment_stale = 0
    assignment_cycle = False
    prev_live_residual = None
    freeze_point = None            # (phi copy, r_norm)
    freeze_cooldown = 0
    n_freeze_reverts = 0
    ptc_r2_ref = None
    r_level_best = np.inf
    freeze_armed = freeze_tol is not None
    prev_snap = snapshot(timings)
    prev_gm = [0]
    dphi = empty(ws.n)
    rhs = empty(ws.n)
    M_pre = ws.precond if amg_rebuild_every <= 0 else None
    # (info, n_iter) of the last solve, n_iter right after the last rebuild,
    # Newton step of the last rebuild
    last_lin = [0, 0, 0, -10]

    def close_step(rec):
        nonlocal prev_snap
        rec.update(step_delta(timings, prev_snap))
        rec["n_lin_iters"] = n_gmres_total - prev_gm[0]
        rec["wall_cum_s"] = time.perf_counter() - t_wall0
        prev_snap = snapshot(timings)
        prev_gm[0] = n_gmres_total

    for it in range(n_newton_max):
        r_norm = cur.r_norm
        residual_history.append(r_norm)
        clamp_history.append((cur.n_limited, cur.n_floored))
        if step_records:
            close_step(step_records[-1])
        rec = {"i": it, "residual": r_norm, "residual_l2": float(np.sqrt(cur.merit)),
               "n_limited": cur.n_limited,
               "n_floored": cur.n_floored, "frozen": frozen is not None,
               "n_lin_solves": 0, "eta": None, "lam": None}
        step_records.append(rec)
        if len(residual_history) >= 3:
            r2, r1, r0 = residual_history[-1], residual_history[-2], residual_history[-3]
            if r2 > 0.0 and r1 > 0.0 and r0 > 0.0 and r1 != r0:
                newton_orders.append(float(np.log(r2 / r1) / np.log(r1 / r0)))
        if verbose:
            print(f"  newton {it:2d}: |R|={r_norm:.3e} |R|2={np.sqrt(cur.merit):.3e} "
                  f"lim={cur.n_limited} "
                  f"flr={cur.n_floored}" + (" [frozen]" if frozen is not None else ""),
                  flush=True)
        r_level_best = min(r_level_best, r_norm)

        # frozen-phase safety net: revert to the freeze point when diverging
        if frozen is not None and freeze_point is not None:
            if (r_norm > 10.0 * freeze_point[1] or cur.n_limited > 0
                    or cur.n_floored > 0):
                cur.phi.copy_(freeze_point[0])
                frozen = None
                freeze_point = None
                freeze_cooldown = 5
                n_freeze_reverts += 1
                r_level_best = np.inf
                if n_freeze_reverts >= freeze_max_reverts:
                    freeze_armed = False
                t0 = time.perf_counter()
                yield from _eval(ws, stepped, cur, upwind_c, m_crit, rho_floor)
                sync()
                timings["residual"] += time.perf_counter() - t0
                continue
        if freeze_cooldown > 0:
            freeze_cooldown -= 1
        # level fail-fast
        if r_norm > 100.0 * r_level_best and r_norm > 1e-3:
            break
        live_stalled = (len(residual_history) >= 7 and r_norm < 1e-3
                        and r_norm > 0.25 * min(residual_history[-7:-1])
                        and r_norm < 2.0 * min(residual_history[-7:-1]))
        if (frozen is None and freeze_armed and freeze_tol is not None
                and freeze_cooldown == 0 and m_inf > 0.0
                and (r_norm < freeze_tol or live_stalled)
                and cur.n_limited == 0 and cur.n_floored == 0):
            frozen = ws.freeze(cur, upwind_c, m_crit, rho_floor)
            freeze_point = (cur.phi.clone(), r_norm)
            r_level_best = np.inf
        accept = None
        if cur.n_limited == 0 and cur.n_floored == 0:
            if r_norm < tol_residual:
                accept = "tol"
            elif frozen is None and it > 0:
                if tol_residual_loose is not None and r_norm < tol_residual_loose:
                    accept = "loose_tol"
                elif (tol_residual_rel is not None
                      and r_norm <= tol_residual_rel * residual_history[0]):
                    accept = "rel_drop"
                elif accept_on_stall and live_stalled:
                    accept = "stall"
        if accept is not None:
            if frozen is None:
                converged = True
                accept_reason = accept
                break
            # honesty check: the live residual at the frozen solution
            t0 = time.perf_counter()
            live = ws.free_state(cur)
            live.phi.copy_(cur.phi)
            yield from _eval(ws, stepped, live, upwind_c, m_crit, rho_floor)
            residual_unfrozen = live.r_norm
            up_new, br_new = ws.freeze(live, upwind_c, m_crit, rho_floor)
            sync()
            timings["residual"] += time.perf_counter() - t0
            n_assignment_stale = int(
                (up_new != frozen[0]).sum().item()
                + ((br_new != frozen[1]) & (up_new == frozen[0])).sum().item())
            if verbose:
                print(f"  frozen system converged; live |R|={residual_unfrozen:.3e}"
                      f", stale assignments={n_assignment_stale}", flush=True)
            if residual_unfrozen < tol_residual:
                converged = True
                cur = live
                break
            if (prev_live_residual is not None
                    and residual_unfrozen > 0.5 * prev_live_residual):
                assignment_cycle = True
                converged = True
                break
            if n_freeze_refresh >= freeze_refresh_max:
                converged = True
                break
            prev_live_residual = residual_unfrozen
            n_freeze_refresh += 1
            frozen = (up_new, br_new)
            freeze_point = (live.phi.clone(), residual_unfrozen)
            r_level_best = np.inf
            cur = live
            continue

        # ---- linearisation
        ptc_scale = 0.0
        if ptc_dt0 is not None:
            r2 = float(np.sqrt(cur.merit))
            if ptc_r2_ref is None:
                ptc_r2_ref = r2
            dt = ptc_dt0 * ptc_r2_ref / max(r2, 1e-300)
            ptc_scale = 0.0 if dt >= ptc_dt_max else 1.0 / dt
            rec["ptc_dt"] = dt
        t0 = time.perf_counter()
        ws.assemble_coupled(cur, upwind_c, m_crit, rho_floor, frozen,
                            ptc_scale=ptc_scale)
        sync()
        timings["assembly"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        if amg_rebuild_every > 0:
            rebuild = M_pre is None or it % amg_rebuild_every == 0
        else:
            # adaptive: keep the hierarchy while the Krylov solves stay cheap;
            # rebuild after a stalled solve or when the iteration count has
            # doubled since the last rebuild
            rebuild = (M_pre is None
                       or ((last_lin[0] != 0
                            or last_lin[1] > 2 * max(last_lin[2], 10))
                           and it - last_lin[3] >= 3))
        if rebuild:
            M_pre = yield from ws.build_preconditioner_steps(
                cur, precond_matrix, precond_clamp, se=ws.se_used, ptc_scale=ptc_scale,
                stepped=stepped)
            last_lin[2] = -1
            last_lin[3] = it
        sync()
        timings["precond"] += time.perf_counter() - t0
        eta = _ew_forcing(r_norm, residual_history[-2] if it > 0 else None,
                          eta_prev, ew_eta0, ew_gamma, ew_eta_max)
        eta_prev = eta
        eta_history.append(eta)
        rec["eta"] = eta
        rec["n_lin_solves"] = 1
        t0 = time.perf_counter()
        scale(-1.0, cur.R, rhs)
        f32b = ws.precision == "fp32"
        x, info, n_it = yield from gmres_steps(
            ws.jac.matvec, rhs, M_pre.apply, rtol=eta, restart=gmres_restart,
            maxiter=gmres_maxiter, basis="f32" if f32b else "f64",
            gw=ws.gmres_work(gmres_restart, f32b) if stepped else None,
            pipelined=not stepped)
        dphi.copy_(x)
        sync()
        timings["linsolve"] += time.perf_counter() - t0
        n_gmres_total += n_it
        rec["gmres_iters"] = n_it
        rec["gmres_info"] = info
        if verbose:
            print(f"      gmres eta={eta:.1e} its={n_it} info={info} "
                  f"t3_active={ws.jac.n_active}"
                  + (" [amg rebuilt]" if rebuild else ""), flush=True)
        if info != 0:
            n_gmres_stalled += 1
        last_lin[0], last_lin[1] = info, n_it
        if last_lin[2] < 0:
            last_lin[2] = n_it

        # ---- safety-only backtracking, least-bad on failure
        lam = 1.0
        best = None
        accepted = False
        trial = None
        t0 = time.perf_counter()
        for _ in range(5):
            trial = ws.free_state(cur, best[1] if best else None)
            trial.phi.copy_(cur.phi)
            axpby(lam, dphi, 1.0, trial.phi)
            yield from _eval(ws, stepped, trial, upwind_c, m_crit, rho_floor, frozen)
            mt = trial.merit
            if np.isfinite(mt) and (not line_search or mt <= cur.merit):
                accepted = True
                break
            if np.isfinite(mt) and (best is None or mt < best[0]):
                best = (mt, trial, lam)
            lam *= 0.5
        sync()
        timings["residual"] += time.perf_counter() - t0
        if not accepted:
            if best is None:
                raise RuntimeError(f"Newton line search failed at iteration {it}")
            _, trial, lam = best
        rec["lam"] = float(lam)
        if verbose:
            print(f"      step lam={lam:g}" + (f" dt={rec['ptc_dt']:.2e}"
                                               if "ptc_dt" in rec else ""), flush=True)
        cur = trial

    if step_records:
        close_step(step_records[-1])
    if converged and accept_reason is None:
        accept_reason = "tol"
    m2 = empty(ws.nt)
    _native.call("fp3d_mach_squared_field", ws.nt, ptr(cur.q2l), float(m_inf),
                 float(gamma_air), ptr(m2), stream_handle())
    mach2_max = ws.red.norms(m2)[1]
    timings["jacobian"] = timings["assembly"]
    timings["amg_setup"] = timings["precond"]
    timings["gmres"] = timings["linsolve"]
    finalize(timings, time.perf_counter() - t_wall0)
    return {
        "phi": to_host(cur.phi), "phi_dev": cur.phi.clone(),
        "gamma": np.zeros(0), "converged": converged,
        "accept_reason": accept_reason,
        "n_newton": len(residual_history) - (1 if converged else 0),
        "residual_history": residual_history,
        "F_history": [0.0] * len(residual_history),
        "clamp_history": clamp_history, "newton_orders": newton_orders,
        "eta_history": eta_history, "n_gmres_total": n_gmres_total,
        "n_gmres_stalled": n_gmres_stalled, "froze": frozen is not None,
        "n_freeze_refresh": n_freeze_refresh,
        "residual_unfrozen": residual_unfrozen,
        "n_assignment_stale": n_assignment_stale,
        "assignment_cycle": assignment_cycle,
        "n_freeze_reverts": n_freeze_reverts, "mach2_max": mach2_max,
        "nu_max": cur.nu_max, "n_nu_active": cur.n_nu_active,
        "n_limited": cur.n_limited, "n_floored": cur.n_floored,
        "jacobian_nnz": ws.st.nnz, "n_term3_active": ws.jac.n_active,
        "n_refactor": 0, "step_records": step_records, "timings": timings,
        "precision": ws.precision, "workspace": ws,
    }


def solve_newton_transonic(nodes, elements, dirichlet_nodes, dirichlet_values,
                           m_inf: float, m_start: float = 0.70, dm: float = 0.05,
                           dm_min: float = 0.01, upwind_c: float = 1.5,
                           m_crit: float = 0.95, upwind_c_post=None,
                           freeze_tol: Optional[float] = 1e-6,
                           intermediate_tol: Optional[float] = None,
                           newton_kw: Optional[dict] = None,
                           verbose: bool = False) -> Dict[str, object]:
    """Upward Mach continuation of solve_newton
    (pyfp3d/solve/newton.py:1144-1307): warm starts from the last CONVERGED
    level, Mach step halved on failure down to dm_min, optional loose
    intermediate levels and post-ramp upwind_c stages (transonic_steps run
    to completion)."""
    kw = dict(locals())
    return drive(transonic_steps(**kw))


def transonic_steps(nodes, elements, dirichlet_nodes, dirichlet_values,
                    m_inf: float, m_start: float = 0.70, dm: float = 0.05,
                    dm_min: float = 0.01, upwind_c: float = 1.5,
                    m_crit: float = 0.95, upwind_c_post=None,
                    freeze_tol: Optional[float] = 1e-6,
                    intermediate_tol: Optional[float] = None,
                    newton_kw: Optional[dict] = None,
                    verbose: bool = False, stepped: bool = False,
                    workspace: Optional[NewtonWorkspace] = None):
    """solve_newton_transonic as a coroutine (stepped as in newton_steps);
    workspace: reuse this workspace for every level."""
    if m_inf < m_start:
        raise ValueError(f"upward ramp only: m_inf {m_inf} < m_start {m_start}")
    kw = dict(newton_kw or {})
    kw.setdefault("upwind_c", upwind_c)
    kw.setdefault("m_crit", m_crit)
    kw.setdefault("freeze_tol", freeze_tol)
    ws = workspace
    result = None
    last_good = None
    level_history, level_results = [], []
    levels = mach_schedule(m_inf, m_start, dm)
    loose = [True] * (len(levels) - 1) + [False]

    def run_level(m, lvl_kw):
        t_lvl = time.perf_counter()
        if stepped:
            res = yield from newton_steps(nodes, elements, dirichlet_nodes,
                                          dirichlet_values, m_inf=m, workspace=ws,
                                          verbose=verbose, stepped=True, **lvl_kw)
        else:
            res = solve_newton(nodes, elements, dirichlet_nodes, dirichlet_values,
                               m_inf=m, workspace=ws, verbose=verbose, **lvl_kw)
        wall = time.perf_counter() - t_lvl
        level_history.append((m, res["n_newton"], res["residual_history"][-1]))
        level_results.append({
            "m": m, "upwind_c": lvl_kw["upwind_c"],
            "converged": res["converged"], "accept_reason": res["accept_reason"],
            "n_newton": res["n_newton"],
            "residual_history": res["residual_history"],
            "step_records": res["step_records"],
            "n_lin_iters": res["n_gmres_total"], "wall_s": wall,
            "timings": res["timings"]})
        return res

    i = 0
    while i < len(levels):
        m = levels[i]
        lvl_kw = dict(kw)
        if intermediate_tol is not None and loose[i]:
            lvl_kw.update(tol_residual_loose=intermediate_tol,
                          tol_residual_rel=1e-3, accept_on_stall=True,
                          freeze_tol=None)
        if last_good is not None:
            lvl_kw.update(phi_init=last_good["phi_dev"], n_picard_seed=0)
        if verbose:
            print(f"newton continuation level M = {m:.4f}", flush=True)
        result = yield from run_level(m, lvl_kw)
        ws = result["workspace"]
        if not result["converged"]:
            m_prev = m_start if i == 0 else levels[i - 1]
            dm_new = 0.5 * (m - m_prev)
            if i == 0 or dm_new < dm_min:
                break
            levels.insert(i, m_prev + dm_new)
            loose.insert(i, False)
            continue
        last_good = result
        i += 1
    if result["converged"] and upwind_c_post:
        for c_post in upwind_c_post:
            lvl_kw = dict(kw)
            lvl_kw.update(upwind_c=c_post, phi_init=result["phi_dev"],
                          n_picard_seed=0)
            result = yield from run_level(m_inf, lvl_kw)
            if not result["converged"]:
                break
    result = dict(result)
    result["level_history"] = level_history
    result["level_results"] = level_results
    result["timings_total"] = sum_timings([lr["timings"] for lr in level_results])
    result.pop("workspace", None)
    return result
