"""The plain drop-in: ``Algorithm(ROS3|RODAS3).solve(system, Y, dt, atol, rtol,
counts, memory)`` on one group, with the reference's signature, return values
and status strings (fastode/solvers.py:270-415).

The host keeps the reference's call structure (one protocol call per
operation, fastode/solvers.py:330-411) while every number lives on the
device: the stage arithmetic, the group error norm and the accept/reject
controller with its H211b memory and step quantization run in HIP kernels
(``fode_stage_*``, ``fode_norm_control``), and the host only reads the
group's small control block after each attempt to learn what comes next.
``project`` (fastode/solvers.py:100-178) runs its Newton iteration on the
device system; its host decisions read device reductions.
"""
from __future__ import annotations

import ctypes as ct
import math

import numpy as np
import torch

from . import _lib as L
from .counts import Counts, Memory
from .system import F64, as_device, device
from .tableaus import ROS3, Tableau


def quantize(h, interval):
    """Host copy of the step rounding (fastode/solvers.py:46-51), for callers
    and tests; the solver rounds on the device."""
    if h >= interval:
        return interval
    return interval / max(math.ceil(interval / h - 1.0e-9), 1)


def ctrl_params(alg, tab: Tableau, constrained: bool) -> L.CtrlParams:
    """The device parameter block of a solver: Algorithm's options, or
    RestartingLoop's / BackwardEuler's (``alg.loop``)."""
    loop = getattr(alg, "loop", L.LOOP_ALGORITHM)
    p = L.CtrlParams()
    p.loop = loop
    p.constrained = int(constrained)
    if loop == L.LOOP_BACKWARD_EULER:
        if len(alg.reductions) > 8:
            raise ValueError("backward Euler takes at most 8 reductions")
        p.stages = 1
        p.max_iterations = int(alg.max_iterations)
        p.n_reductions = len(alg.reductions)
        for i, r in enumerate(alg.reductions):
            p.reductions[i] = float(r)
        p.small = float(alg.small)
        return p
    S = tab.stages
    if loop == L.LOOP_RESTARTING:
        p.step_change_rows = int(alg.step_change_rows)
        p.max_steps = int(alg.max_steps)
    else:
        p.b, p.safety, p.fac_min, p.fac_max = alg.b, alg.safety, alg.fac_min, alg.fac_max
        p.h211b = int(alg.controller == "h211b")
        p.carry, p.quantize, p.carry_memory = int(alg.carry), int(alg.quantize), int(alg.carry_memory)
        p.max_steps = int(alg.max_steps)
    p.k_exp, p.gamma = tab.k, tab.gamma
    p.stages = S
    nf = tab.new_forcing
    p.n_new_forcing = int(sum(nf[1:]))
    a = np.zeros((4, 4))
    c = np.zeros((4, 4))
    a[:S, :S] = tab.a
    c[:S, :S] = tab.c
    for i in range(16):
        p.a[i] = a.ravel()[i]
        p.c[i] = c.ravel()[i]
    for i in range(S):
        p.m[i] = tab.m[i]
        p.e[i] = tab.e[i]
        p.new_forcing[i] = int(nf[i])
    return p


class GroupControl:
    """One group's device control block (fode_group_ctrl)."""

    def __init__(self, dt, memory: Memory):
        rec = np.zeros(1, dtype=L.CTRL_DTYPE)
        rec["dt"] = dt
        rec["h_next"], rec["err_prev"], rec["h_prev"] = memory.h_next, memory.err_prev, memory.h_prev
        self.dev = torch.from_numpy(rec.view(np.int64).copy()).to(device())
        self.ptr = ct.c_void_p(self.dev.data_ptr())

    def read(self):
        return self.dev.cpu().numpy().view(L.CTRL_DTYPE)[0]

    def alpha(self):
        """0-dim view of the device alpha = 1/(H gamma) of the next attempt."""
        return self.dev.view(F64)[L.ALPHA_OFFSET // 8: L.ALPHA_OFFSET // 8 + 1]


def stages(system, tab, params, Y, F0, LU, ctrl, K, work):
    """One Rosenbrock attempt (fastode/solvers.py:54-86) on the device; H is
    read by the kernels from the control block."""
    C, n = Y.shape
    s = L.stream_handle()
    dae = int(system.has_constraints)
    nf = tab.new_forcing
    forcing = F0
    for r in range(tab.stages):
        if r > 0 and nf[r]:
            Yr = work["Yr"]
            L.check(L.lib.fode_stage_point(ct.byref(params), r, C, n, L.ptr(Y), L.ptr_array(K[:r]),
                                           L.ptr(Yr), s), "fode_stage_point")
            forcing = system.rhs(Yr)
        rhs = work["rhs"]
        L.check(L.lib.fode_stage_forcing(ct.byref(params), r, C, n, ctrl.ptr, L.ptr(forcing),
                                         L.ptr_array(K[:r]), L.ptr(system.mass_dev), dae, L.ptr(rhs), s),
                "fode_stage_forcing")
        LU.solve(rhs, out=K[r])
    Ynew, E = work["Ynew"], work["E"]
    L.check(L.lib.fode_stage_final(ct.byref(params), C, n, L.ptr(Y), L.ptr_array(K[: tab.stages]),
                                   L.ptr(Ynew), L.ptr(E), s), "fode_stage_final")
    return Ynew, E


def clamp(system, Y):
    """max(Y, 0), differential columns only with constraints
    (fastode/solvers.py:89-97); in place on the device tensor."""
    C, n = Y.shape
    L.check(L.lib.fode_clamp(C, n, L.ptr(Y), L.ptr(system.mass_dev), int(system.has_constraints),
                             L.stream_handle()), "fode_clamp")
    return Y


def project(system, Y, atol, rtol, counts: Counts, max_updates=10, tol=0.1,
            max_backtracks=24, backtrack_factor=0.5, sigma=1.0e-4):
    """Consistent initialization by damped Newton on the algebraic rows
    (fastode/solvers.py:100-178).  Returns (Y, status); a failure restores
    Y.  A NumPy Y comes back as NumPy, a device tensor as a device tensor.

    The residual masking, the backtracking candidates and the finite /
    zero / weighted-max measures run as HIP kernels (``fode_proj_*``); the
    host takes the reference's decisions, reading two doubles per measure."""
    if not np.any(system.mass_np == 0.0):
        return Y, "converged"
    as_numpy = not isinstance(Y, torch.Tensor)
    Yp, status = _project_device(system, as_device(Y), atol, rtol, counts, max_updates, tol, max_backtracks,
                                 backtrack_factor, sigma)
    return _out(Yp, as_numpy), status


def _project_device(system, Y, atol, rtol, counts, max_updates, tol, max_backtracks, backtrack_factor, sigma):
    C, n = Y.shape
    s = L.stream_handle()
    mass = system.mass_dev
    atol_d = as_device(atol, (n,))
    rtol = float(rtol)
    meas = torch.empty(2, dtype=F64, device=Y.device)

    def measure(V, Z=None):
        """(finite, max): max |V| with Z None, else the weighted max of (Z, V)."""
        L.check(L.lib.fode_proj_measure(C, n, L.ptr(V), L.ptr(Z), L.ptr(mass), L.ptr(atol_d), rtol,
                                        L.ptr(meas), s), "fode_proj_measure")
        bad, m = meas.tolist()
        return bad == 0.0, m

    def residual(Z):
        G = system.rhs(Z).contiguous()
        counts.rhs += 1
        out = torch.empty_like(G)                  # differential rows exactly 0
        L.check(L.lib.fode_proj_mask(C, n, L.ptr(G), L.ptr(mass), L.ptr(out), s), "fode_proj_mask")
        return out

    def candidate(Z, D, step):
        out = torch.empty_like(Z)
        L.check(L.lib.fode_proj_candidate(C, n, L.ptr(Z), L.ptr(D), float(step), L.ptr(mass), L.ptr(out), s),
                "fode_proj_candidate")
        return out

    Y = Y.contiguous()
    saved = None
    for update in range(max_updates + 1):
        Gv = residual(Y)
        ok, gmax = measure(Gv)
        if not ok:
            break
        counts.proj_passes += 1
        if gmax == 0.0:
            return Y, "converged"
        J = system.jac(Y)
        counts.jac += 1
        LU = system.factor(1.0, J, algebraic_rows_only=True)
        counts.lu += 1
        delta = LU.solve(Gv).contiguous()
        counts.solves += 1
        ok, q = measure(delta, Y)
        if not ok:
            break
        if q <= tol:
            return Y, "converged"
        if update == max_updates:
            break
        if saved is None:
            saved = Y.clone()
        step = 1.0
        taken = None
        for _ in range(max_backtracks + 1):
            cand = candidate(Y, delta, step)
            used = step
            step *= backtrack_factor
            if not measure(cand)[0]:
                continue
            Gc = residual(cand)
            counts.proj_trials += 1
            ok, gcmax = measure(Gc)
            if not ok:
                continue
            if gcmax == 0.0:
                taken = cand
                break
            trial = LU.solve(Gc).contiguous()
            counts.solves += 1
            ok, qt = measure(trial, cand)
            if not ok:
                continue
            if qt <= tol or qt < (1.0 - sigma * used) * q:
                taken = cand
                break
        if taken is None:
            break
        Y = taken
    if saved is not None:
        Y = saved
    return Y, "projection failed"


class Algorithm:
    """The integrator of the reference's Section 4 (fastode/solvers.py:270-415)
    on the device.  ``trace``, when set to a list, receives one (H, err,
    accepted) tuple per attempt."""

    name = "algorithm"
    loop = L.LOOP_ALGORITHM

    def __init__(self, tableau: Tableau = ROS3, controller: str = "h211b", b: float = 4.0,
                 safety: float = 0.9, fac_min: float = 0.2, fac_max: float = 6.0,
                 carry: bool = True, quantize: bool = True, max_steps: int = 1000,
                 carry_memory: bool = True):
        self.tableau = tableau
        self.controller = controller
        self.b = b
        self.safety = safety
        self.fac_min = fac_min
        self.fac_max = fac_max
        self.carry = carry
        self.quantize = quantize
        self.max_steps = max_steps
        self.carry_memory = carry_memory
        self.trace = None

    def new_memory(self):
        return Memory()

    def solve(self, system, Y, dt, atol, rtol, counts: Counts, memory: Memory = None):
        return rosenbrock_loop(self, system, Y, dt, atol, rtol, counts, memory)


def rosenbrock_loop(self, system, Y, dt, atol, rtol, counts: Counts, memory: Memory = None):
    """The host side of Algorithm.solve and RestartingLoop.solve: one
    protocol call per operation, the step decisions on the device (the
    controller named by ``self.loop``)."""
    tab = self.tableau
    mem = memory if memory is not None else Memory()
    as_numpy = not isinstance(Y, torch.Tensor)
    Yd = as_device(Y).clone()
    C, n = Yd.shape
    atol_d = as_device(atol)
    constrained = system.has_constraints
    params = ctrl_params(self, tab, constrained)
    ctrl = GroupControl(float(dt), mem)
    s = L.stream_handle()
    L.check(L.lib.fode_ctrl_init(ct.byref(params), ctrl.ptr, 1, int(constrained), s), "ctrl_init")
    host_counts = Counts()
    if constrained:
        Yp, st = project(system, Yd, atol, rtol, host_counts)
        if st != "converged":
            g = ctrl.read()
            mem.err_prev, mem.h_prev = float(g["err_prev"]), float(g["h_prev"])
            counts.add_array(g["counts"])
            counts.add(host_counts)
            return _out(Yp, as_numpy), st
        Yd = Yp.contiguous()
        L.check(L.lib.fode_ctrl_begin(ct.byref(params), ctrl.ptr, 1, s), "ctrl_begin")
    work = {k: torch.empty_like(Yd) for k in ("Yr", "rhs", "Ynew", "E")}
    shares = torch.empty(C, dtype=F64, device=Yd.device)     # the cells' shares of the error norm
    K = [torch.empty_like(Yd) for _ in range(tab.stages)]
    g = ctrl.read()
    F0 = J = None
    attempted = False
    while g["action"] != L.FINISH:
        attempted = True
        if g["action"] == L.NEW_STEP:
            F0 = system.rhs(Yd)
            J = system.jac(Yd)
        H = float(g["H"])
        LU = system.factor(ctrl.alpha(), J)
        Ynew, E = stages(system, tab, params, Yd, F0, LU, ctrl, K, work)
        L.check(L.lib.fode_norm_control(ct.byref(params), ctrl.ptr, C, n, L.ptr(Yd), L.ptr(Ynew),
                                        L.ptr(E), L.ptr(atol_d), float(rtol),
                                        L.ptr(system.mass_dev), L.ptr(shares), s), "fode_norm_control")
        g = ctrl.read()
        if self.trace is not None:
            self.trace.append((H, float(g["err"]), bool(g["accepted"])))
    if not attempted:
        # no attempt ran (a zero budget, an empty interval): the norm/control
        # kernel that clamps at the end never launched (solvers.py:414)
        clamp(system, Yd)
    if self.loop == L.LOOP_ALGORITHM:
        mem.h_next = float(g["h_next"])
        mem.err_prev, mem.h_prev = float(g["err_prev"]), float(g["h_prev"])
    counts.add_array(g["counts"])
    counts.add(host_counts)
    return _out(Yd, as_numpy), L.STATUS[int(g["status"])]


def _out(Y, as_numpy):
    return Y.cpu().numpy() if as_numpy else Y


class RestartingLoop:
    """The baseline Rosenbrock loop (fastode/solvers.py:189-267) on the
    device: every call starts from 1e-6 of the interval with the elementary
    controller; no memory is carried.  ``step_change_rows`` replaces the
    algebraic rows of the error estimate by the step change."""

    name = "restarting"
    loop = L.LOOP_RESTARTING

    def __init__(self, tableau: Tableau = ROS3, max_steps: int = 1000, step_change_rows: bool = True):
        self.tableau = tableau
        self.max_steps = max_steps
        self.step_change_rows = step_change_rows
        self.trace = None

    def new_memory(self):
        return None

    def solve(self, system, Y, dt, atol, rtol, counts: Counts, memory=None):
        return rosenbrock_loop(self, system, Y, dt, atol, rtol, counts, None)


class BackwardEuler:
    """Backward Euler with Newton's iteration (fastode/solvers.py:418-491):
    each pass evaluates F and J at the iterate and factors I/H - J through the
    system, solves for the correction, and the device applies it, tests the
    group's convergence and takes the step decision (``fode_be_update``)."""

    name = "backward-euler"
    loop = L.LOOP_BACKWARD_EULER

    def __init__(self, max_iterations: int = 11, reductions=(0.5, 0.5, 0.5, 0.5, 0.1),
                 small: float = 1.0e-40):
        self.max_iterations = max_iterations
        self.reductions = tuple(reductions)
        self.small = small

    def new_memory(self):
        return None

    def solve(self, system, Y, dt, atol, rtol, counts: Counts, memory=None):
        as_numpy = not isinstance(Y, torch.Tensor)
        Yn = as_device(Y).clone()
        Yn1 = Yn.clone()
        C, n = Yn.shape
        atol_d = as_device(atol)
        params = ctrl_params(self, None, system.has_constraints)
        ctrl = GroupControl(float(dt), Memory())
        s = L.stream_handle()
        L.check(L.lib.fode_ctrl_init(ct.byref(params), ctrl.ptr, 1, 0, s), "ctrl_init")
        resid = torch.empty_like(Yn)
        flags = torch.empty(2 * C, dtype=F64, device=Yn.device)    # each cell's convergence flags
        g = ctrl.read()
        while g["action"] != L.FINISH:
            F = system.rhs(Yn1)
            J = system.jac(Yn1)
            LU = system.factor(ctrl.alpha(), J)
            L.check(L.lib.fode_be_residual(C, n, ctrl.ptr, L.ptr(F), L.ptr(Yn1), L.ptr(Yn), L.ptr(resid),
                                           L.stream_handle()), "fode_be_residual")
            delta = LU.solve(resid)
            L.check(L.lib.fode_be_update(ct.byref(params), ctrl.ptr, C, n, L.ptr(Yn1), L.ptr(Yn),
                                         L.ptr(delta), L.ptr(atol_d), float(rtol), L.ptr(flags),
                                         L.stream_handle()), "fode_be_update")
            g = ctrl.read()
        counts.add_array(g["counts"])
        return _out(Yn1, as_numpy), L.STATUS[int(g["status"])]


def fixed_steps(system, tab, Y, t_end, n_steps, counts: Counts = None):
    """n uncontrolled Rosenbrock steps of equal length (fastode/solvers.py:494-504)."""
    alg = Algorithm(tab)
    counts = counts or Counts()
    Yd = as_device(Y).clone()
    params = ctrl_params(alg, tab, system.has_constraints)
    H = t_end / n_steps
    ctrl = GroupControl(t_end, Memory())
    ctrl.dev.view(F64)[1] = H                        # the H field read by the stage kernels
    work = {k: torch.empty_like(Yd) for k in ("Yr", "rhs", "Ynew", "E")}
    K = [torch.empty_like(Yd) for _ in range(tab.stages)]
    for _ in range(n_steps):
        F0 = system.rhs(Yd)
        J = system.jac(Yd)
        LU = system.factor(1.0 / (H * tab.gamma), J)
        Ynew, _ = stages(system, tab, params, Yd, F0, LU, ctrl, K, work)
        Yd = Ynew.clone()
    return Yd
