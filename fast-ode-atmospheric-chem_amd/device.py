"""Device-resident mechanism tables and LU plans (created once, cached).

A ``DeviceMechanism`` uploads a ``Mechanism``'s gather tables through
``fode_mech_create``; a ``DevicePlan`` runs the host symbolic analysis of a
pattern (``symbolic.build_plan``) and uploads it through ``fode_plan_create``.
Both are cached on the host object they come from, as the reference shares one
mechanism and pattern across the per-interval systems (fastode/ts1.py:215).
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib as L
from .symbolic import build_plan


def _i32(a):
    return np.ascontiguousarray(np.asarray(a), dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(np.asarray(a), dtype=np.float64)


def _u32(a):
    return np.ascontiguousarray(np.asarray(a), dtype=np.uint32)


def _addr(a):
    return a.ctypes.data_as(ct.c_void_p) if a.size else None


class DeviceMechanism:
    def __init__(self, mech):
        self.mech = mech
        arrs = dict(
            react=_i32(mech.react), rhs_ptr=_i32(mech.rhs_ptr), rhs_rxn=_i32(mech.rhs_rxn),
            rhs_coef=_f64(mech.rhs_coef), jac_ptr=_i32(mech.jac_ptr), jac_src=_i32(mech.jac_src),
            jac_coef=_f64(mech.jac_coef), jac_const=_f64(mech.jac_const),
            con_rows=_i32(mech.constraint_rows), con_coef=_f64(mech.constraint_coef),
            con_total=_f64(mech.constraint_total))
        d = L.MechDesc(n=mech.n, R=mech.R, P=mech.P, nnz=mech.nnz, ncon=len(mech.constraint_rows),
                       **{k: _addr(v) for k, v in arrs.items()})
        h = ct.c_void_p()
        L.check(L.lib.fode_mech_create(ct.byref(d), ct.byref(h)), "fode_mech_create")
        self.handle = h
        self._keep = arrs
        self._destroy = L.lib.fode_mech_destroy   # module globals may be gone at interpreter exit

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self._destroy(h)
            self.handle = None


class DevicePlan:
    def __init__(self, n, rows, cols, **plan_kw):
        plan = build_plan(n, rows, cols, **plan_kw)
        self.plan = plan
        self.plan_kw = dict(plan_kw)
        a = plan.arrays
        unsigned = ("f_tasks", "f_pairs", "sf_tasks", "sb_tasks", "s_pairs", "sch_pairs", "sch_mask")
        arrs = {}
        for k, v in a.items():
            if k in unsigned:
                arrs[k] = _u32(v)
            elif k == "core_src":
                arrs[k] = np.ascontiguousarray(v, dtype=np.int16)
            else:
                arrs[k] = _i32(v)
        d = L.PlanDesc(
            n=plan.n, c=plan.c, m=plan.m, tile=plan.tile, ns=plan.ns, nsch=plan.nsch, nnz=plan.nnz,
            slot_stride=plan.slot_stride,
            n_f_tasks=len(a["f_tasks"]) // 2, f_nst=len(a["f_ptr"]) - 1, n_f_pairs=len(a["f_pairs"]),
            n_sf_tasks=len(a["sf_tasks"]) // 2, sf_nst=len(a["sf_ptr"]) - 1,
            n_sb_tasks=len(a["sb_tasks"]) // 2, sb_nst=len(a["sb_ptr"]) - 1,
            n_s_pairs=len(a["s_pairs"]), n_sch_pairs=len(a["sch_pairs"]),
            **{k: _addr(v) for k, v in arrs.items()})
        h = ct.c_void_p()
        lu_len = ct.c_int64()
        L.check(L.lib.fode_plan_create(ct.byref(d), ct.byref(h), ct.byref(lu_len)), "fode_plan_create")
        self.handle = h
        self.lu_len = lu_len.value
        self._keep = arrs
        self._destroy = L.lib.fode_plan_destroy

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self._destroy(h)
            self.handle = None


def device_mechanism(mech) -> DeviceMechanism:
    dm = getattr(mech, "_device", None)
    if dm is None:
        dm = DeviceMechanism(mech)
        mech._device = dm
    return dm


def mechanism_plan(mech) -> DevicePlan:
    dp = getattr(mech, "_plan", None)
    if dp is None:
        dp = DevicePlan(mech.n, mech.rows, mech.cols)
        mech._plan = dp
    return dp


def transposed_plan(mech) -> DevicePlan:
    """Plan of the transposed pattern (rows and columns swapped) under the
    forward plan's ordering: its entries are the same Jacobian values, its
    factorization is that of (alpha*M - J)^T, and its split matches the
    forward plan's (the no-pivot fill of A^T is the transpose of A's)."""
    dp = getattr(mech, "_plan_t", None)
    if dp is None:
        fp = mechanism_plan(mech)
        fwd = fp.plan
        kw = {k: v for k, v in fp.plan_kw.items() if k in ("rho", "m_max")}
        dp = DevicePlan(mech.n, mech.cols, mech.rows, order=fwd.perm, tile=fwd.tile, **kw)
        t = dp.plan
        if (t.c, t.m, t.tile, t.ns, t.nsch) != (fwd.c, fwd.m, fwd.tile, fwd.ns, fwd.nsch):
            raise RuntimeError("the transposed plan's split differs from the forward plan's")
        mech._plan_t = dp
    return dp


_DENSE_PLANS: dict = {}


def dense_plan(n) -> DevicePlan:
    """Plan of a fully dense n x n pattern in row-major entry order, so that a
    dense Jacobian (C, n, n) is read as nnz = n*n values per cell."""
    dp = _DENSE_PLANS.get(n)
    if dp is None:
        rows = np.repeat(np.arange(n), n)
        cols = np.tile(np.arange(n), n)
        dp = DevicePlan(n, rows, cols, order=np.arange(n))
        _DENSE_PLANS[n] = dp
    return dp
