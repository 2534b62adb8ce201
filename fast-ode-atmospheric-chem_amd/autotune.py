"""Autotuner of the batched form's launch configuration for a mechanism.

The per-cell LU engine's cost depends on how the symbolic plan splits the
ordered matrix into sparse pivots (level-scheduled LDS tasks) and the dense
register-tiled core (the core's size picks the register tile TILE), and on
whether the plan's tables are staged in LDS.  The tuner builds the distinct
plans a range of core-density thresholds and core-size caps give, drops the
ones whose workgroup does not fit in LDS, and times the batched kernel on a
sample of the caller's cells for each plan, with and without table staging,
and, for small register tiles with the tables in memory, at one or two
workgroups (cells) per CU (two need the kernel held to 256 registers per
lane); it keeps the fastest per cell-attempt.  Each candidate runs the sample's first interval
(the start-up transient, from 1e-6 of the interval) untimed and is timed on
the second, a steady one; the three fastest are then timed again over two
more steady intervals, and the fastest of that pass is kept.  Results are cached per (pattern, tableau, device,
library build) in a JSON file; a cached choice is re-checked against the
device's LDS before it is installed, and re-tuned when it no longer fits.

``tune(...)`` installs the winner as the mechanism's plan, so every later
solve (plain and batched) uses it.  Every candidate is a valid no-pivot plan
of the same ordering; which one is chosen changes roundings only.
"""
from __future__ import annotations

import hashlib
import json
import os
from dataclasses import asdict, dataclass
from pathlib import Path

import numpy as np
import torch

from . import _lib as L
from .device import DevicePlan
from .system import F64, ChemSystem

RHOS = (0.6, 0.7, 0.75, 0.8, 0.85, 0.9, 0.95)
OCC2_MAX_TILE = 5        # largest register tile with a two-waves-per-SIMD kernel (tables in memory)
M_MAXES = (160, 144, 128)
_BUILD = {}


def build_id():
    """A digest of the built library and of the host planner: any change to
    the kernels, the plan tables or the engine gives a new cache key."""
    if "id" not in _BUILD:
        h = hashlib.sha1()
        h.update(L.LIB_PATH.read_bytes())
        h.update((Path(__file__).with_name("symbolic.py")).read_bytes())
        _BUILD["id"] = h.hexdigest()[:12]
    return _BUILD["id"]


CACHE = Path(os.environ.get("FASTODE_AMD_TUNE_CACHE", Path.home() / ".cache" / "fastode_amd" / "tune.json"))


@dataclass
class Choice:
    rho: float
    m_max: int
    stage_tables: bool
    c: int = 0
    tile: int = 0
    us_per_cell_attempt: float = 0.0
    blocks_per_cu: int = 1       # workgroups (cells in flight) per CU


def pattern_key(mech, tableau, device_name):
    h = hashlib.sha1()
    h.update(np.asarray([mech.n, mech.nnz], dtype=np.int64).tobytes())
    h.update(np.asarray(mech.rows, dtype=np.int64).tobytes())
    h.update(np.asarray(mech.cols, dtype=np.int64).tobytes())
    return f"{h.hexdigest()[:16]}:{tableau.name}:{device_name}:build-{build_id()}"


def candidates(mech):
    """Distinct (rho, m_max) plans of the mechanism, by their split (c, tile)."""
    out, seen = [], set()
    for m_max in M_MAXES:
        for rho in RHOS:
            try:
                dp = DevicePlan(mech.n, mech.rows, mech.cols, rho=rho, m_max=m_max)
            except (ValueError, L.HipError):
                continue
            key = (dp.plan.c, dp.plan.tile)
            if key not in seen:
                seen.add(key)
                out.append((rho, m_max, dp))
    return out


def _load_cache():
    try:
        return json.loads(CACHE.read_text())
    except (OSError, ValueError):
        return {}


def _save_cache(d):
    try:
        CACHE.parent.mkdir(parents=True, exist_ok=True)
        CACHE.write_text(json.dumps(d, indent=1, sort_keys=True))
    except OSError:
        pass


def install(mech, choice: Choice):
    """Make the choice the mechanism's plan (clears cached derived plans)."""
    mech._plan = DevicePlan(mech.n, mech.rows, mech.cols, rho=choice.rho, m_max=choice.m_max)
    if hasattr(mech, "_plan_t"):
        del mech._plan_t
    mech._tuned = choice
    return mech._plan


def fits(mech, choice: Choice, tableau) -> bool:
    """Whether the choice's workgroup fits the device's LDS (with its tables
    staged when it asks for that)."""
    import ctypes as ct
    from .device import device_mechanism
    dp = DevicePlan(mech.n, mech.rows, mech.cols, rho=choice.rho, m_max=choice.m_max)
    b = ct.c_int64()
    L.check(L.lib.fode_batched_lds_bytes(dp.handle, device_mechanism(mech).handle, tableau.stages,
                                         ct.byref(b)), "fode_batched_lds_bytes")
    return b.value + L.CTRL_DTYPE.itemsize <= L.device_info()["lds_per_block"]


def tune(mech, k, Y, dt, atol, rtol, tableau=None, kmul=None, mass=None, use_cache=True,
         verbose=False) -> Choice:
    """Time the candidate plans on the sample (k (C, R) or (1, R), Y (C, n):
    the caller's cells, at least two groups per CU for the two-cells-per-CU
    candidates to show their worth: 16,384 cells on 256 CUs) over one
    steady interval, the three fastest again over two more, and install the
    fastest of those."""
    from .batched import BatchedAlgorithm
    from .tableaus import ROS3
    tableau = tableau or ROS3
    dev_name = torch.cuda.get_device_name(0)
    key = pattern_key(mech, tableau, dev_name)
    cache = _load_cache() if use_cache else {}
    if key in cache:
        try:
            choice = Choice(**cache[key])
            ok = fits(mech, choice, tableau)
        except (TypeError, ValueError, L.HipError):
            ok = False
        if ok:
            install(mech, choice)
            return choice
    saved = getattr(mech, "_plan", None)
    weights = np.minimum(32, len(Y) - 32 * np.arange(-(-len(Y) // 32)))

    def time_plan(dp, stage, bpc, intervals):
        """us per cell-attempt over `intervals` steady intervals after the
        untimed start-up interval, or None when the launch is refused."""
        mech._plan = dp
        balg = BatchedAlgorithm(tableau, stage_tables=stage, blocks_per_cu=bpc)
        try:
            system = ChemSystem(mech, k, mass=mass)
            run = balg.runner(system, Y, dt, atol, rtol, graph=False)
            run.step(kmul)                  # the start-up interval, untimed
            before = run.memory.counts()[:, 0].copy()
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            ev0.record()
            for _ in range(intervals):
                run.step(kmul)
            ev1.record()
            torch.cuda.synchronize()
        except L.HipError:
            return None
        att = float(((run.memory.counts()[:, 0] - before) * weights[:run.ngroups]).sum())
        return 1e3 * ev0.elapsed_time(ev1) / max(att, 1.0)

    results = []
    for rho, m_max, dp in candidates(mech):
        for stage in (True, False):
            for bpc in ((1, 2) if dp.plan.tile <= OCC2_MAX_TILE and not stage else (1,)):
                us = time_plan(dp, stage, bpc, 1)
                if us is None:
                    continue
                results.append((Choice(rho, m_max, stage, dp.plan.c, dp.plan.tile, us, bpc), dp))
                if verbose:
                    print(f"tune: rho {rho} m_max {m_max} c {dp.plan.c} tile {dp.plan.tile} "
                          f"stage {stage} blocks/CU {bpc or 'max'}: {us:.3f} us per cell-attempt")
    if not results:
        mech._plan = saved
        raise RuntimeError("no candidate plan fits this device")
    # The first pass separates close candidates by less than its own noise:
    # the three fastest are timed again over two more steady intervals.
    results.sort(key=lambda r: r[0].us_per_cell_attempt)
    final = []
    for choice, dp in results[:3]:
        us = time_plan(dp, choice.stage_tables, choice.blocks_per_cu, 2)
        if us is not None:
            choice.us_per_cell_attempt = us
            final.append(choice)
            if verbose:
                print(f"tune: again c {choice.c} tile {choice.tile} stage {choice.stage_tables} "
                      f"blocks/CU {choice.blocks_per_cu}: {us:.3f} us per cell-attempt")
    best = min(final or [r[0] for r in results], key=lambda r: r.us_per_cell_attempt)
    install(mech, best)
    if use_cache:
        cache[key] = asdict(best)
        _save_cache(cache)
    return best
