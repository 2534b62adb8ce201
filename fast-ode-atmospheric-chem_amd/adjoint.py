"""Discrete adjoint of the batched integrator: gradients of a loss of the
final state with respect to log rate constants and the initial state.

The forward pass runs the batched persistent form interval by interval
(fastode/ts1.py:196-216's loop with per-group Memory), keeping one state
checkpoint per interval and recording each group's accepted steps.  The
backward pass walks the intervals last first with ``fode_adjoint_interval``,
which recomputes the interval's accepted steps from its checkpoint and runs
their exact discrete adjoint (stage solves with (alpha*M - J)^T).  Step sizes
are held fixed: they are quantized to divisors of the interval
(fastode/solvers.py:46-51), so they do not move under small perturbations of
the inputs, and the result is the derivative of the map the forward applies.

``integrate(adj, logk, Y0, kmuls, dt, atol, rtol)`` is the differentiable
entry: a ``torch.autograd.Function`` whose backward is the kernel above.
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L
from .batched import BatchedAlgorithm, BatchedRunner
from .device import device_mechanism, mechanism_plan, transposed_plan
from .system import F64, ChemSystem, as_device
from .tableaus import ROS3


@dataclass
class Tape:
    """What the backward pass needs: per interval the checkpoint (the state
    at the interval's start), the multipliers, and the accepted steps."""
    k: torch.Tensor                       # (C, R) rate constants (before multipliers)
    kmuls: torch.Tensor                   # (I, R)
    checkpoints: list = field(default_factory=list)   # I x (C, n)
    hlogs: list = field(default_factory=list)         # I x (ngroups, hcap)
    naccs: list = field(default_factory=list)         # I x (ngroups,)
    statuses: np.ndarray | None = None


class BatchedAdjoint:
    """Forward integration with a tape, and its adjoint."""

    def __init__(self, mech, tableau=ROS3, group_size=32, blocks_per_cu=0, **options):
        if len(mech.constraint_rows):
            raise NotImplementedError("the adjoint covers unconstrained mechanisms")
        self.mech = mech
        self.tableau = tableau
        self.balg = BatchedAlgorithm(tableau, group_size, blocks_per_cu, **options)
        self.plan = mechanism_plan(mech)
        self.plan_t = transposed_plan(mech)
        self.dmech = device_mechanism(mech)

    # ----------------------------------------------------------- forward
    def forward(self, k, Y0, kmuls, dt, atol, rtol, memory=None):
        """Integrate len(kmuls) intervals of dt from Y0 with rate constants
        k * kmuls[i] in interval i.  Returns (Y_final, tape)."""
        k = as_device(k)
        C = Y0.shape[0]
        if k.dim() == 1:
            k = k[None, :]
        if k.shape[0] == 1:
            k = k.expand(C, -1)
        k = k.contiguous()
        kmuls = as_device(kmuls)
        if kmuls.dim() != 2 or kmuls.shape[1] != self.mech.R:
            raise ValueError("kmuls must be (intervals, R)")
        system = ChemSystem(self.mech, k)
        run = BatchedRunner(self.balg, system, Y0, dt, atol, rtol, memory, graph=False, record=True)
        tape = Tape(k=k, kmuls=kmuls)
        self.dt, self.rtol = float(dt), float(rtol)
        for i in range(kmuls.shape[0]):
            tape.checkpoints.append(run.Y.clone())
            hlog = torch.empty((run.ngroups, run.hcap), dtype=F64, device=run.Y.device)
            nacc = torch.zeros(run.ngroups, dtype=torch.int32, device=run.Y.device)
            run.record_into(hlog, nacc)
            run.step(kmuls[i])
            tape.hlogs.append(hlog)
            tape.naccs.append(nacc)
        self.run = run
        return run.Y, tape

    # ---------------------------------------------------------- backward
    def backward(self, tape: Tape, Ybar):
        """dL/dlog k (C, R) and dL/dY0 (C, n) from dL/dY_final."""
        lam = as_device(Ybar).contiguous().clone()
        C, n = lam.shape
        R = self.mech.R
        kbar = torch.zeros((C, R), dtype=F64, device=lam.device)
        S = self.tableau.stages
        nmax = int(torch.stack(tape.naccs).max().item()) if tape.naccs else 0
        ws_stride = max(1, nmax) * (S + 1) * n
        blocks = ct.c_int32()
        L.check(L.lib.fode_adjoint_ws_blocks(self.plan.handle, self.plan_t.handle, self.dmech.handle, S, C,
                                             ct.byref(blocks)), "fode_adjoint_ws_blocks")
        ws = torch.empty(blocks.value * ws_stride, dtype=F64, device=lam.device)
        queue = torch.zeros(1, dtype=torch.int32, device=lam.device)
        params = self.run.params
        gs = self.balg.group_size
        mass = self.run.system.mass_dev
        s = L.stream_handle()
        for i in reversed(range(len(tape.checkpoints))):
            kmul = tape.kmuls[i].contiguous()
            d = L.AdjointDesc(
                C=C, group_size=gs, ngroups=-(-C // gs), k_stride=R, hcap=self.run.hcap,
                ws_stride=ws_stride, Y=tape.checkpoints[i].data_ptr(), k=tape.k.data_ptr(),
                kmul=kmul.data_ptr(), mass=mass.data_ptr(), hlog=tape.hlogs[i].data_ptr(),
                nacc=tape.naccs[i].data_ptr(), lam=lam.data_ptr(), kbar=kbar.data_ptr(),
                ws=ws.data_ptr(), queue=queue.data_ptr())
            # every array named in d is held by the tape or by this frame; ws and
            # queue return to torch's stream-ordered cache only after the launches
            L.check(L.lib.fode_adjoint_interval(self.plan.handle, self.plan_t.handle, self.dmech.handle,
                                                ct.byref(params), ct.byref(d), s), "fode_adjoint_interval")
        return kbar, lam


class _Integrate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, logk, Y0, adj, kmuls, dt, atol, rtol):
        k = torch.exp(logk)
        Y, tape = adj.forward(k, Y0.detach(), kmuls, dt, atol, rtol)
        ctx.adj, ctx.tape, ctx.shared = adj, tape, logk.dim() == 1 or logk.shape[0] == 1
        ctx.logk_shape = logk.shape
        return Y.clone()

    @staticmethod
    def backward(ctx, Ybar):
        kbar, y0bar = ctx.adj.backward(ctx.tape, Ybar)
        if ctx.shared:
            kbar = kbar.sum(0).reshape(ctx.logk_shape)
        return kbar, y0bar, None, None, None, None, None


def integrate(adj: BatchedAdjoint, logk, Y0, kmuls, dt, atol, rtol):
    """Final state after len(kmuls) intervals, differentiable with respect
    to logk ((C, R) per cell, or (R,) shared) and Y0 (C, n)."""
    return _Integrate.apply(logk, Y0, adj, kmuls, dt, atol, rtol)
