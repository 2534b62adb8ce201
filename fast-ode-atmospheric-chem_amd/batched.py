"""Batched persistent form: every group of a batch over one host interval in
one launch (``fode_batched_interval``), optionally replayed from a hipGraph.

It replaces the reference's group loop around ``Algorithm.solve``
(fastode/ts1.py:196-216, group-outer and interval-inner) by interval-outer,
group-parallel launches with the same per-group semantics: each group of
``group_size`` consecutive cells keeps its own controller state and Memory
on the device across calls, and its steps, Counts and status are those a
separate ``Algorithm.solve`` call per group would give.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np
import torch

from . import _lib as L
from .counts import COUNT_FIELDS, Counts, Memory
from .device import device_mechanism, mechanism_plan
from .solvers import Algorithm, BackwardEuler, RestartingLoop, ctrl_params
from .system import F64, as_device, device
from .tableaus import ROS3


class GroupMemories:
    """The device controller blocks of all groups (fode_group_ctrl each):
    the carried Memory, the running Counts and the failure record."""

    def __init__(self, ngroups):
        rec = np.zeros(ngroups, dtype=L.CTRL_DTYPE)
        self.dev = torch.from_numpy(rec.view(np.int64).copy()).to(device())
        self.ngroups = ngroups

    def read(self):
        return self.dev.cpu().numpy().view(L.CTRL_DTYPE)

    def memory(self, group) -> Memory:
        r = self.read()[group]
        return Memory(float(r["h_next"]), float(r["err_prev"]), float(r["h_prev"]))

    def counts(self) -> np.ndarray:
        return self.read()["counts"].copy()


class BatchedAlgorithm:
    """``Algorithm`` semantics (same options) over a whole batch of groups."""

    name = "batched-algorithm"

    def __init__(self, tableau=ROS3, group_size=32, blocks_per_cu=0, stage_tables=None, longest_first=True,
                 **options):
        """stage_tables: copy the plan's tables into LDS when they fit (True),
        never (False), or as the mechanism's tuned choice says (None; stage
        when untuned)."""
        self._setup(Algorithm(tableau, **options), tableau, group_size, blocks_per_cu, stage_tables,
                    longest_first)

    def _setup(self, solver, tableau, group_size, blocks_per_cu, stage_tables, longest_first):
        self.alg = solver
        self.tableau = tableau
        self.group_size = group_size
        self.blocks_per_cu = blocks_per_cu
        self.stage_tables = stage_tables
        self.longest_first = longest_first

    def new_memory(self, C):
        return GroupMemories(-(-C // self.group_size))

    def runner(self, system, Y, dt, atol, rtol, memory=None, graph=True, record=False):
        return BatchedRunner(self, system, Y, dt, atol, rtol, memory, graph, record)

    def solve(self, system, Y, dt, atol, rtol, counts: Counts, memory=None, kmul=None):
        """One host interval for every group. Returns (Y_new, statuses), the
        statuses a list of per-group strings."""
        r = BatchedRunner(self, system, Y, dt, atol, rtol, memory, graph=False)
        before = r.memory.counts()
        r.step(kmul)
        rec = r.memory.read()
        d = rec["counts"] - before
        for name, v in zip(COUNT_FIELDS, d.sum(axis=0)):
            setattr(counts, name, getattr(counts, name) + int(v))
        statuses = [L.STATUS[int(s)] for s in rec["status"]]
        out = r.Y.clone()
        return (out.cpu().numpy() if not isinstance(Y, torch.Tensor) else out), statuses


class BatchedRestartingLoop(BatchedAlgorithm):
    """``RestartingLoop`` semantics (fastode/solvers.py:189-267) over a whole
    batch of groups: each call restarts every group from 1e-6 of the
    interval; no memory is carried (the device blocks keep only Counts)."""

    name = "batched-restarting"

    def __init__(self, tableau=ROS3, group_size=32, blocks_per_cu=0, stage_tables=None, longest_first=True,
                 max_steps=1000, step_change_rows=True):
        self._setup(RestartingLoop(tableau, max_steps, step_change_rows), tableau, group_size, blocks_per_cu,
                    stage_tables, longest_first)


class BatchedBackwardEuler(BatchedAlgorithm):
    """``BackwardEuler`` semantics (fastode/solvers.py:418-491) over a whole
    batch of groups: every Newton pass of a group runs cell by cell on its
    persistent workgroup, and the group's convergence decides its step."""

    name = "batched-backward-euler"

    def __init__(self, group_size=32, blocks_per_cu=0, stage_tables=None, longest_first=True,
                 max_iterations=11, reductions=(0.5, 0.5, 0.5, 0.5, 0.1), small=1.0e-40):
        self._setup(BackwardEuler(max_iterations, reductions, small), None, group_size, blocks_per_cu,
                    stage_tables, longest_first)


class BatchedRunner:
    """Device buffers and (optionally) the captured interval graph of one
    batch; ``step(kmul)`` advances every group by one interval."""

    def __init__(self, balg: BatchedAlgorithm, system, Y, dt, atol, rtol, memory=None, graph=True,
                 record=False):
        loop = balg.alg.loop
        if system.has_constraints and loop != L.LOOP_BACKWARD_EULER and balg.tableau.name != "rodas3":
            raise ValueError("constrained systems need RODAS3")
        if record and loop != L.LOOP_ALGORITHM:
            raise ValueError("step records are kept for Algorithm's loop only")
        self.balg = balg
        self.system = system
        self.Y = as_device(Y).clone()
        C, n = self.Y.shape
        self.C, self.n = C, n
        gs = balg.group_size
        self.ngroups = -(-C // gs)
        self.memory = memory if memory is not None else balg.new_memory(C)
        if self.memory.ngroups != self.ngroups:
            raise ValueError("memory holds a different number of groups")
        k = system.k
        if k.shape[0] not in (1, C):
            raise ValueError("k must have 1 or C rows")
        self.k = k
        self.kmul = torch.ones(system.mech.R, dtype=F64, device=self.Y.device)
        self.atol = as_device(atol)
        # ynew per cell; the projection's corrections and saved state when constrained
        projected = system.has_constraints and loop != L.LOOP_BACKWARD_EULER
        self.work = torch.empty(((2 if projected else 1) * C, n), dtype=F64, device=self.Y.device)
        self.queue = torch.zeros(1, dtype=torch.int32, device=self.Y.device)
        # order the workgroups take groups in: by the previous call's attempts,
        # most first (no effect on results)
        self.order = torch.arange(self.ngroups, dtype=torch.int32, device=self.Y.device)
        self.params = ctrl_params(balg.alg, balg.tableau, system.has_constraints)
        self.plan = mechanism_plan(system.mech)
        self.dmech = device_mechanism(system.mech)
        self.desc = L.BatchDesc(
            C=C, group_size=gs, ngroups=self.ngroups, k_stride=(0 if k.shape[0] == 1 else k.shape[1]),
            dt=float(dt), rtol=float(rtol), Y=self.Y.data_ptr(), k=k.data_ptr(),
            kmul=self.kmul.data_ptr(), atol=self.atol.data_ptr(), mass=system.mass_dev.data_ptr(),
            g=self.memory.dev.data_ptr(), work=self.work.data_ptr(), queue=self.queue.data_ptr(),
            options=0 if self._stage(balg, system) else 1,
            order=self.order.data_ptr() if balg.longest_first else None)
        self.hcap = int(getattr(balg.alg, "max_steps", 0))
        if record:
            # accepted steps of the last call: H per group, and their number
            self.hlog = torch.zeros((self.ngroups, self.hcap), dtype=F64, device=self.Y.device)
            self.nacc = torch.zeros(self.ngroups, dtype=torch.int32, device=self.Y.device)
            self._point_record(self.hlog, self.nacc)
        self.graph = None
        if graph:
            h = ct.c_void_p()
            L.check(L.lib.fode_batched_graph_create(self.plan.handle, self.dmech.handle,
                                                    ct.byref(self.params), ct.byref(self.desc),
                                                    self._blocks_per_cu(balg, system), ct.byref(h)),
                    "fode_batched_graph_create")
            self.graph = h
            # held here: module globals may be gone when a runner is
            # collected at interpreter exit
            self._graph_destroy = L.lib.fode_graph_destroy

    @staticmethod
    def _blocks_per_cu(balg, system):
        """The caller's workgroups per CU, else the tuned choice's."""
        if balg.blocks_per_cu:
            return int(balg.blocks_per_cu)
        tuned = getattr(system.mech, "_tuned", None)
        return int(getattr(tuned, "blocks_per_cu", 0) or 0)

    @staticmethod
    def _stage(balg, system):
        if balg.stage_tables is not None:
            return bool(balg.stage_tables)
        tuned = getattr(system.mech, "_tuned", None)
        return True if tuned is None else bool(tuned.stage_tables)

    def record_into(self, hlog, nacc):
        """Record the accepted steps of the following calls into hlog
        (ngroups, max_steps) and nacc (ngroups,) (eager launches only: a
        captured graph keeps the arrays it was built with)."""
        if self.graph is not None:
            raise RuntimeError("re-pointing step records needs eager launches (graph=False)")
        self._point_record(hlog, nacc)

    def _point_record(self, hlog, nacc):
        if tuple(hlog.shape) != (self.ngroups, self.hcap) or tuple(nacc.shape) != (self.ngroups,):
            raise ValueError("record arrays have the wrong shape")
        if hlog.dtype != F64 or nacc.dtype != torch.int32 or not (hlog.is_contiguous() and nacc.is_contiguous()):
            raise ValueError("record arrays must be contiguous float64 / int32")
        self.desc.hlog = hlog.data_ptr()
        self.desc.nacc = nacc.data_ptr()
        self.desc.hcap = self.hcap
        self._rec = (hlog, nacc)

    def set_kmul(self, kmul):
        if kmul is None:
            self.kmul.fill_(1.0)
        else:
            self.kmul.copy_(torch.as_tensor(kmul, dtype=F64), non_blocking=False)

    def step(self, kmul=None):
        self.set_kmul(kmul)
        self.step_graph_only()

    def step_graph_only(self):
        """Advance one interval with the multipliers already in ``self.kmul``."""
        s = L.stream_handle()
        if self.graph is not None:
            L.check(L.lib.fode_graph_launch(self.graph, s), "fode_graph_launch")
        else:
            L.check(L.lib.fode_batched_interval(self.plan.handle, self.dmech.handle, ct.byref(self.params),
                                                ct.byref(self.desc), self._blocks_per_cu(self.balg, self.system),
                                                s),
                    "fode_batched_interval")
        if self.balg.longest_first:
            # attempts of this call per group (fode_group_ctrl.nsteps), most first
            nsteps = self.memory.dev.view(torch.int32).view(self.ngroups, L.CTRL_DTYPE.itemsize // 4)[
                :, L.NSTEPS_WORD]
            self.order.copy_(torch.argsort(nsteps, descending=True, stable=True).to(torch.int32))

    def lds_bytes(self):
        b = ct.c_int64()
        L.check(L.lib.fode_batched_lds_bytes(self.plan.handle, self.dmech.handle, self.balg.tableau.stages,
                                             ct.byref(b)), "fode_batched_lds_bytes")
        return b.value

    def __del__(self):
        g = getattr(self, "graph", None)
        if g:
            self._graph_destroy(g)
            self.graph = None
