"""Throughput of the batched Rosenbrock integrator on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one host interval of the batched persistent form (one hipGraph
replay: queue reset + persistent kernel) over every group of the workload:
the seeded 200-species, 500-reaction mechanism with diurnal photolysis on
65,536 cells in groups of 32, ROS3, rtol 1e-3, atol 1e-20, 600 s intervals
(48 of them by default).  Warmup intervals run on a separate copy of the
batch; the timed run starts from the initial state.  Inputs are resident in
HBM before timing starts.

The JSON line also carries the per-group plain form (one launch sequence
per group and protocol call, the reference's call structure) measured on a
sample of groups over the same intervals, the HBM roofline of the interval
kernel, the legs of the other workloads (262,144 cells of uneven stiffness,
TS1 on 65,536 cells, the constrained RODAS3 workload on 65,536 cells) over
the same number of intervals, and the CPU oracle's rate on the box's cores.  The adjoint leg times
the gradient of sum w_s log(Y_final + 1e-17) with respect to log k and Y0
on 16,384 cells over 24 intervals against the plain batched forward of the
same intervals ("adjoint_over_forward").
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "cell-intervals/s, batched vs per-group launch, 200 spp rtol 1e-3; adjoint/fwd"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
CPU_SHARE = 16     # host processes one GPU's share of the box allows
V_PASSES = 22      # unfused stage/norm vector passes per attempt counted per species


def algorithmic_bytes(counts, cells_per_group, n, R, nnz_j, nnz_lu):
    """Compulsory bytes of a streaming implementation, summed over groups:
    8 * cells * [lu (nnz_J + nnz_LU) + solves (nnz_LU + 2n) + jac (n + R + nnz_J)
                 + rhs (2n + R) + steps V n]."""
    c = counts.astype(np.float64)
    steps, rhs, jac, lu, solves = c[:, 0], c[:, 3], c[:, 4], c[:, 5], c[:, 6]
    per_cell = (lu * (nnz_j + nnz_lu) + solves * (nnz_lu + 2 * n) + jac * (n + R + nnz_j)
                + rhs * (2 * n + R) + steps * V_PASSES * n)
    return float(8.0 * np.sum(per_cell * cells_per_group))


def algorithmic_flops(counts, cells_per_group, n, R, P, nnz_j, fact_flops, nnz_lu, jac_terms, rhs_terms):
    c = counts.astype(np.float64)
    steps, rhs, jac, lu, solves = c[:, 0], c[:, 3], c[:, 4], c[:, 5], c[:, 6]
    f_rhs = R * P + 2.0 * rhs_terms
    f_jac = R * P * (P - 1) + 2.0 * jac_terms
    f_vec = 27.0 * n
    per_cell = lu * fact_flops + solves * 2.0 * nnz_lu + jac * f_jac + rhs * f_rhs + steps * f_vec
    return float(np.sum(per_cell * cells_per_group))


def _oracle_setup(name, C, cells=None):
    """The oracle's mechanism and the workload of a worker: the first C cells
    (their draws are those of the full workload's first rows), or for the
    uneven workload the given cells only."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import fastode_oracle as O
    import __graft_entry__
    __graft_entry__.load_package()
    from fastode_amd import workloads as W
    if name == "diurnal":
        wl = W.diurnal200(C=C)
    elif name == "uneven":
        wl = W.uneven200(cells=cells)
    else:
        wl = W.constrained100(C=C)
    mech = wl.mech
    cons = None
    if len(mech.constraint_rows):
        cons = {mech.species[r]: ({mech.species[s]: float(mech.constraint_coef[i, s])
                                   for s in np.nonzero(mech.constraint_coef[i])[0]},
                                  float(mech.constraint_total[i]))
                for i, r in enumerate(mech.constraint_rows)}
    om = O.OracleMechanism(mech.species, mech.reactions_list(), cons)
    return O, wl, om


_WORKER = {}


def _oracle_warm(args):
    """Build the worker's oracle state (outside the timed region)."""
    name, C = args
    if "state" not in _WORKER:
        _WORKER["state"] = _oracle_setup(name, C, np.arange(32))
    return os.getpid()


def _oracle_job(args):
    """One group of the CPU sample through the oracle over the GPU's timed
    intervals; returns (cell-intervals, cell-attempts)."""
    name, C, group, n_int = args
    cells = np.arange(group * 32, group * 32 + 32)
    if name == "uneven":
        O, wl, om = _oracle_setup(name, C, cells)
        cells = np.arange(32)
    else:
        if "state" not in _WORKER:
            _WORKER["state"] = _oracle_setup(name, C)
        O, wl, om = _WORKER["state"]
    alg = O.Algorithm(O.ROS3 if wl.tableau == "ros3" else O.RODAS3)
    mem = alg.new_memory()
    Y = wl.Y0[cells].copy()
    c = O.Counts()
    for it in range(n_int):
        Y, _ = alg.solve(O.ChemSystem(om, wl.k(it, cells), mass=wl.mass, dense=False), Y, wl.dt, wl.atol,
                         wl.rtol, c, mem)
    return 32 * n_int, 32 * c.steps


def cpu_baseline(name, C, n_groups, n_int):
    """The oracle on the box's host cores over the same intervals the GPU
    timed (0 .. n_int-1, from the same initial state), one group per worker
    process; the pool and the workers' set-up are outside the timed region."""
    import multiprocessing as mp
    affinity = len(os.sched_getaffinity(0))
    cores = max(1, min(CPU_SHARE, affinity))
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        os.environ[k] = "1"
    ctx = mp.get_context("spawn")
    with ctx.Pool(cores) as pool:
        pool.map(_oracle_warm, [(name, C)] * cores, chunksize=1)
        t0 = time.perf_counter()
        res = pool.map(_oracle_job, [(name, C, g, n_int) for g in range(n_groups)], chunksize=1)
        wall = time.perf_counter() - t0
    done = sum(r[0] for r in res)
    attempts = sum(r[1] for r in res)
    return {"value": done / wall, "unit": "cell-intervals/s", "cores": cores, "kind": "port",
            "cell_attempts_per_s": round(attempts / wall, 1),
            "sample": f"oracle (NumPy/SciPy SuperLU restatement of the reference path), groups 0-{n_groups - 1} "
                      f"x intervals 0-{n_int - 1} (the GPU's timed intervals, same initial state) = {done} "
                      f"cell-intervals, {attempts} cell-attempts in {wall:.1f} s; {cores} worker processes "
                      f"with 1 BLAS thread each ({affinity} cores in the affinity mask, {CPU_SHARE} the "
                      f"host share of one GPU); pool start-up outside the timing"}


def adjoint_leg(fa, torch, W, C, n_int, choice=None):
    """Gradient of sum_c sum_s w_s log(Y_final + 1e-17) with respect to the
    log rate constants (C, R) and Y0, over n_int intervals: time of the plain
    batched forward (graph replays) against the taped forward (checkpoints,
    step records) plus the backward pass."""
    from fastode_amd.adjoint import BatchedAdjoint
    from fastode_amd.system import ChemSystem
    dev = torch.device("cuda", 0)
    wl = W.adjoint200(C=C, n_intervals=n_int)
    k = torch.as_tensor(np.broadcast_to(wl.k_base, (C, wl.R)).copy(), device=dev)
    kmuls = torch.as_tensor(wl.kmuls(), device=dev)
    Y0 = torch.as_tensor(wl.Y0, device=dev)
    w = torch.as_tensor(wl.weights, device=dev)
    if choice is not None:      # the plan the tuner picked for this mechanism
        from fastode_amd.autotune import install
        install(wl.mech, choice)
    adj = BatchedAdjoint(wl.mech)
    # warm: one interval of a small batch through both passes
    Yw, tw = adj.forward(k[:64], Y0[:64], kmuls[:1], wl.dt, wl.atol, wl.rtol)
    adj.backward(tw, torch.ones_like(Yw))
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
    run = fa.BatchedAlgorithm(fa.ROS3).runner(ChemSystem(wl.mech, k), Y0, wl.dt, wl.atol, wl.rtol, graph=True)
    torch.cuda.synchronize()
    ev[0].record(stream)
    for i in range(n_int):
        run.kmul.copy_(kmuls[i])
        run.step_graph_only()
    ev[1].record(stream)
    Y, tape = adj.forward(k, Y0, kmuls, wl.dt, wl.atol, wl.rtol)
    ev[2].record(stream)
    Ybar = (w[None, :] / (Y + 1e-17)).expand_as(Y).contiguous()
    ev[3].record(stream)
    kbar, y0bar = adj.backward(tape, Ybar)
    ev[4].record(stream)
    torch.cuda.synchronize()
    fwd = ev[0].elapsed_time(ev[1])
    taped = ev[1].elapsed_time(ev[2])
    bwd = ev[3].elapsed_time(ev[4])
    same = bool(torch.equal(Y, run.Y))
    nacc = torch.stack(tape.naccs).float()
    return {"cells": C, "intervals": n_int, "forward_ms": round(fwd, 1), "taped_forward_ms": round(taped, 1),
            "backward_ms": round(bwd, 1), "ratio_total": round((taped + bwd) / fwd, 3),
            "ratio_backward": round(bwd / fwd, 3), "taped_state_bitwise_equal": same,
            "gradients_finite": bool(torch.isfinite(kbar).all() and torch.isfinite(y0bar).all()),
            "accepted_steps_per_group_interval": round(float(nacc.mean()), 3),
            "note": "ratio_total = (taped forward + backward) / plain batched forward, same intervals"}


def progress(msg):
    """A progress line on stderr (stdout carries only the JSON line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def k_resident(wl, torch, dev):
    """Rate constants resident on the device: one shared row, or per cell."""
    k = wl.k_base[None, :] if wl.cell_scale is None else wl.k_base[None, :] * wl.cell_scale
    return torch.as_tensor(k, device=dev)


def read_phase_timers(reset=False):
    """Cycles per phase of the timed library's batched kernel (summed over
    workgroups since the last reset)."""
    import ctypes as ct
    from fastode_amd import _lib as L
    buf = (ct.c_uint64 * len(L.PHASES))()
    n = L.lib.fode_phase_timers(buf, len(L.PHASES), int(reset))
    if n < 0:
        raise RuntimeError(L.lib.fode_last_error().decode())
    return dict(zip(L.PHASES, (int(x) for x in buf)))


def timed_run(fa, torch, wl, tab, system, K, warmup, blocks_per_cu, longest_first, stage_tables=None,
              phase_timers=False):
    """Warm a copy of the batch, then time K intervals (graph replays) from
    the initial state; per-launch times from HIP events on the launch stream."""
    dev = torch.device("cuda", 0)
    balg = fa.BatchedAlgorithm(tab, blocks_per_cu=blocks_per_cu, longest_first=longest_first,
                               stage_tables=stage_tables)
    warm = balg.runner(system, wl.Y0, wl.dt, wl.atol, wl.rtol, graph=True)
    for it in range(warmup):
        warm.step(wl.kmul(it))
    torch.cuda.synchronize()
    del warm
    run = balg.runner(system, wl.Y0, wl.dt, wl.atol, wl.rtol, graph=True)
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    kms = [torch.as_tensor(wl.kmul(it), dtype=torch.float64, device=dev) for it in range(K)]
    counts_before = run.memory.counts()
    torch.cuda.synchronize()
    if phase_timers:
        read_phase_timers(reset=True)
    t0 = time.perf_counter()
    for it in range(K):
        run.kmul.copy_(kms[it])
        ev[it][0].record(stream)
        run.step_graph_only()
        ev[it][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = [a.elapsed_time(b) for a, b in ev]
    rec = run.memory.read()
    counts = rec["counts"] - counts_before
    cells_per_group = np.minimum(32, wl.C - 32 * np.arange(run.ngroups))
    if phase_timers:
        cyc = read_phase_timers()
        attempts = float((counts[:, 0] * cells_per_group).sum())
        total = sum(cyc.values())
        run.phase_report = {"cycles_per_cell_attempt": round(total / attempts),
                            "share": {k: round(v / total, 4) for k, v in cyc.items()},
                            "cycles_per_cell_attempt_by_phase": {k: round(v / attempts) for k, v in cyc.items()},
                            "note": "workgroup cycles (shader clock, thread 0's marks) summed over workgroups, "
                                    "over the timed intervals, per cell-attempt"}
    return run, wall, launch_ms, counts, rec, cells_per_group


def counter_profile(name):
    """The committed rocprofv3 counter summary of the kernel on a workload,
    if there is one (profiles/counters_<workload>.json)."""
    f = ROOT / "profiles" / f"counters_{name}.json"
    return json.loads(f.read_text()) if f.exists() else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cells", type=int, default=0, help="0: the workload's own size")
    ap.add_argument("--workload", choices=("diurnal", "uneven", "constrained"), default="diurnal",
                    help="diurnal: the headline 200-species workload; uneven: 262,144 cells of uneven "
                         "stiffness; constrained: RODAS3 with conservation rows (100 species)")
    ap.add_argument("--plain-groups", type=int, default=8)
    ap.add_argument("--cpu-groups", type=int, default=0, help="0: one per worker process")
    ap.add_argument("--cpu-intervals", type=int, default=0, help="0: the GPU's timed intervals")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--stage-tables", choices=("auto", "yes", "no"), default="auto",
                    help="plan tables in LDS (yes), in memory (no), or as tuned (auto)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default plan")
    ap.add_argument("--no-order", action="store_true", help="take groups in index order (not costliest first)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run a plain copy kernel over this many MiB (HBM counter calibration)")
    ap.add_argument("--adjoint-cells", type=int, default=16384, help="0 skips the adjoint leg")
    ap.add_argument("--adjoint-intervals", type=int, default=24)
    ap.add_argument("--uneven-cells", type=int, default=262144,
                    help="cells of the uneven-stiffness leg (0 skips it); it runs the timed step count")
    ap.add_argument("--phase-timers", action="store_true",
                    help="switched off: runs with libfastode_hip_timed.so (build.py --phase-timers) went "
                         "silent for 180 s on the GPU twice, with no cause found in the code")
    ap.add_argument("--ts1-cells", type=int, default=65536,
                    help="cells of the TS1 leg (the reference's real mechanism; 0 skips it)")
    ap.add_argument("--constrained-cells", type=int, default=65536,
                    help="cells of the constrained RODAS3 leg (100 species + 3 conservation rows; 0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the groups shard "
              "across GPUs with no exchange, so the figure is N times the one-GPU rate", file=sys.stderr)
        sys.exit(2)

    if args.phase_timers:
        print("bench.py --phase-timers is switched off: both runs of the timed build "
              "(libfastode_hip_timed.so) on an MI355X went silent for 180 s and were killed, and reading "
              "the code found no cause; it stays off until one is shown", file=sys.stderr)
        sys.exit(2)
    import torch
    import __graft_entry__
    fa = __graft_entry__.load_package()
    from fastode_amd import workloads as W
    from fastode_amd.system import ChemSystem
    from fastode_amd.device import mechanism_plan

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.copy_probe_mib > 0:
        # an elementwise kernel that reads and writes N MiB once each
        xs = torch.ones(args.copy_probe_mib << 17, dtype=torch.float64, device=dev)
        ys = torch.empty_like(xs)
        for _ in range(3):
            torch.mul(xs, 1.0, out=ys)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": xs.numel() * 8, "copy_probe_bytes_written": ys.numel() * 8}))
        del xs, ys
    K = args.steps
    if args.workload == "diurnal":
        wl = W.diurnal200(C=args.cells or 65536, n_intervals=K)
    elif args.workload == "uneven":
        wl = W.uneven200(C=args.cells or 262144, n_intervals=K)
    else:
        wl = W.constrained100(C=args.cells or 65536, n_intervals=K)
    tab = fa.ROS3 if wl.tableau == "ros3" else fa.RODAS3
    mech = wl.mech
    k_dev = k_resident(wl, torch, dev)

    progress(f"workload {wl.name}: {wl.C} cells, {K} timed intervals")
    tuned = None
    if not args.no_tune:
        # autotuner: the first 8,192 cells per candidate (a start-up interval,
        # then a timed steady one; the 16,384 cells that two-cells-per-CU
        # candidates need when the register tile allows them)
        from fastode_amd.autotune import tune
        n_s = min(wl.C, 16384 if wl.n <= 120 else 8192)
        t_tune = time.perf_counter()
        ch = tune(mech, k_dev[:n_s] if k_dev.shape[0] > 1 else k_dev, wl.Y0[:n_s], wl.dt, wl.atol, wl.rtol,
                  tableau=tab, kmul=wl.kmul(0), mass=wl.mass, use_cache=False)
        tuned = {"rho": ch.rho, "m_max": ch.m_max, "stage_tables": ch.stage_tables, "c": ch.c, "tile": ch.tile,
                 "blocks_per_cu": ch.blocks_per_cu or "max",
                 "us_per_cell_attempt": round(ch.us_per_cell_attempt, 3),
                 "tuning_s": round(time.perf_counter() - t_tune, 1)}
        progress(f"tuned: {tuned}")
    system = ChemSystem(mech, k_dev, mass=wl.mass)
    plan = mechanism_plan(mech).plan
    stage = {"auto": None, "yes": True, "no": False}[args.stage_tables]
    run, wall, launch_ms, counts, rec, cells_per_group = timed_run(
        fa, torch, wl, tab, system, K, args.warmup, args.blocks_per_cu, not args.no_order, stage,
        phase_timers=args.phase_timers)
    kernel_s = sum(launch_ms) / 1e3
    progress(f"timed: {wall:.2f} s wall, launches {[round(x, 1) for x in launch_ms]} ms")
    fails = int((rec["fails"] > 0).sum())
    cell_intervals = wl.C * K
    cell_attempts = float((counts[:, 0] * cells_per_group).sum())
    value = cell_intervals / wall
    fact_flops, nnz_lu = plan.flops()
    fact_flops = 2 * fact_flops
    alg_bytes = algorithmic_bytes(counts, cells_per_group, wl.n, wl.R, mech.nnz, nnz_lu)
    alg_flops = algorithmic_flops(counts, cells_per_group, wl.n, wl.R, mech.P, mech.nnz, fact_flops, nnz_lu,
                                  len(mech.jac_src), len(mech.rhs_rxn))
    achieved = alg_bytes / kernel_s / 1e9
    # HBM traffic per launch: measured bytes per cell-attempt (rocprofv3 PMC
    # FETCH_SIZE + WRITE_SIZE of this kernel, scaled by a copy kernel's
    # reading) times this run's cell-attempts per launch
    traffic = None
    counter_gbs = None
    pmc = ROOT / "profiles" / f"traffic_{wl.name}.json"
    if pmc.exists():
        t = json.loads(pmc.read_text())
        per_attempt = t.get("hbm_bytes_per_cell_attempt")
        if per_attempt:
            traffic = round(per_attempt * cell_attempts / K)
            counter_gbs = traffic / (kernel_s / K) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "achieved_is": "algorithmic bytes (the compulsory HBM traffic of a streaming, unfused "
                               "implementation of the same Counts: LU written and read per attempt) per "
                               "launch / kernel time: a work-rate index, not bytes this kernel moves",
                "alg_work_rate_gbs": round(achieved, 1),
                "alg_bytes_per_launch": round(alg_bytes / K), "avg_launch_ms": round(1e3 * kernel_s / K, 3),
                "hbm_counter_gbs": None if counter_gbs is None else round(counter_gbs, 2),
                "hbm_counter_frac": None if counter_gbs is None else round(counter_gbs / HBM_PEAK_GBS, 5),
                "fp64_tflops": round(alg_flops / kernel_s / 1e12, 3),
                "fp64_frac": round(alg_flops / kernel_s / (FP64_PEAK_TFLOPS * 1e12), 4),
                "us_per_cell_attempt": round(1e6 * kernel_s / cell_attempts, 4)}
    if args.phase_timers:
        roofline["phase_timers"] = run.phase_report
    prof = counter_profile(wl.name)
    if prof is not None:
        roofline["measured_binding"] = prof.get("binding")
        roofline["counters"] = prof.get("summary")

    out = {"metric": METRIC, "value": round(value, 1), "unit": "cell-intervals/s", "n_gpus": 1,
           "steps": K, "warmup": args.warmup, "ms_per_step": round(1e3 * wall / K, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded mechanism and states)",
           "config": {"workload": wl.name, "species": wl.n, "reactions": wl.R, "cells": wl.C,
                      "group_size": 32, "groups": run.ngroups, "intervals": K, "dt_s": wl.dt,
                      "rtol": wl.rtol, "atol": float(wl.atol[0]), "tableau": wl.tableau,
                      "form": "batched persistent kernel, one hipGraph replay per interval",
                      "lu": {"nnz_J": mech.nnz, "nnz_LU": nnz_lu, "core": plan.m, "tile": plan.tile},
                      "autotuned": tuned},
           "attempts_per_group_interval": round(float(counts[:, 0].sum()) / (run.ngroups * K), 3),
           "rejected_per_group_interval": round(float(counts[:, 2].sum()) / (run.ngroups * K), 3),
           "cell_attempts_per_s": round(cell_attempts / wall, 1),
           "failed_groups": fails,
           "roofline": roofline}

    if not args.no_plain:
        alg = fa.Algorithm(tab)
        G = min(args.plain_groups, run.ngroups)
        Yp = wl.Y0[: 32 * G].copy()
        mems = [alg.new_memory() for _ in range(G)]
        pc = fa.Counts()
        systems = [[None] * G for _ in range(K)]
        for it in range(K):
            for g in range(G):
                systems[it][g] = ChemSystem(mech, wl.k(it, np.arange(32 * g, 32 * g + 32)), mass=wl.mass)
        Yd = torch.as_tensor(Yp, device=dev)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for it in range(K):
            for g in range(G):
                sl = slice(32 * g, 32 * g + 32)
                Yd[sl], _ = alg.solve(systems[it][g], Yd[sl], wl.dt, wl.atol, wl.rtol, pc, mems[g])
        torch.cuda.synchronize()
        pwall = time.perf_counter() - t1
        plain_value = 32 * G * K / pwall
        same = bool(np.array_equal(Yd.cpu().numpy(), run.Y[: 32 * G].cpu().numpy()))
        out["plain"] = {"value": round(plain_value, 1), "unit": "cell-intervals/s",
                        "sample": f"first {G} groups x {K} intervals, one launch sequence per group and call",
                        "attempts_per_group_interval": round(pc.steps / (G * K), 3),
                        "states_bitwise_equal_to_batched": same}
        out["speedup_vs_plain"] = round(value / plain_value, 1)
    del run

    if not args.no_plain:
        progress(f"plain path: {out['plain']['value']} cell-intervals/s")
    if args.uneven_cells > 0 and args.workload == "diurnal":
        progress("uneven leg")
        out["uneven"] = uneven_leg(fa, torch, W, args.uneven_cells, K, mech, args)

    if args.ts1_cells > 0 and args.workload == "diurnal":
        progress("ts1 leg")
        out["ts1"] = ts1_leg(fa, torch, args.ts1_cells, K, args)
    if args.constrained_cells > 0 and args.workload == "diurnal":
        progress("constrained leg")
        out["constrained"] = constrained_leg(fa, torch, W, args.constrained_cells, K, args)
    if args.adjoint_cells > 0:
        progress("adjoint leg")
        out["adjoint"] = adjoint_leg(fa, torch, W, args.adjoint_cells, args.adjoint_intervals,
                                     getattr(mech, "_tuned", None))
        out["adjoint_over_forward"] = out["adjoint"]["ratio_total"]

    if not args.no_cpu:
        cores = max(1, min(CPU_SHARE, len(os.sched_getaffinity(0))))
        ng = args.cpu_groups or cores
        progress("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(args.workload, 32 * ng, ng, args.cpu_intervals or K)
    print(json.dumps(out))


def ts1_leg(fa, torch, C, K, args, scenario="diurnal-10min"):
    """TS1 (209 species, 547 reactions, the reference's real mechanism) on C
    cells: the scenario's 64 recorded cells tiled to C, rate constants on the
    device, the photolysis columns rewritten by fode_scale_columns at every
    interval start (fastode/ts1.py:205-213), one graph replay per interval;
    rtol 1e-3, atol 1e-17 (the reference's alg configuration)."""
    import fastode_amd.ts1 as T
    from fastode_amd import _lib as L
    from fastode_amd.autotune import tune
    from fastode_amd.system import ChemSystem, as_device
    dev = torch.device("cuda", 0)
    m = T.TS1()
    scn = T.Scenario.from_npz(ROOT / "tests" / "golden" / "ts1" / f"{scenario}.npz")
    base_cells = scn.make_cells(m)
    reps = -(-C // base_cells.C)
    idx = np.tile(np.arange(base_cells.C), reps)[:C]
    T_, P_ = base_cells.T[idx], base_cells.P[idx]
    params = {k: v[idx] for k, v in base_cells.params.items()}
    Y0 = base_cells.Y[idx]
    rtol, atol = 1e-3, np.full(m.mech.n, 1e-17)
    phot = [k for k in params if k.startswith("j")]
    f0 = scn.factor(0.0)
    p0 = dict(params)
    for name in phot:
        p0[name] = params[name] * f0
    k_dev = as_device(m.rate_constants(T_, P_, p0))
    cols, labels, scale = m.photolysis_table()
    cols_dev = torch.as_tensor(cols, device=dev)
    base_dev = as_device(np.stack([params[lab] for lab in labels], axis=1))
    scale_dev = as_device(scale)
    f_dev = torch.zeros(1, dtype=torch.float64, device=dev)
    tuned = None
    if not args.no_tune:
        n_s = min(C, 16384)
        ch = tune(m.mech, k_dev[:n_s], Y0[:n_s], scn.dt, atol, rtol, use_cache=False)
        tuned = {"c": ch.c, "tile": ch.tile, "stage_tables": ch.stage_tables, "blocks_per_cu": ch.blocks_per_cu or "max",
                 "us_per_cell_attempt": round(ch.us_per_cell_attempt, 3)}
        progress(f"ts1 tuned: {tuned}")
    system = ChemSystem(m.mech, k_dev)
    balg = fa.BatchedAlgorithm(fa.ROS3, blocks_per_cu=args.blocks_per_cu)
    run = balg.runner(system, Y0, scn.dt, atol, rtol, graph=True)
    run.set_kmul(None)
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    before = run.memory.counts()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for it in range(K):
        if it > 0:
            f_dev.fill_(scn.factor(it * scn.dt))
            L.check(L.lib.fode_scale_columns(C, k_dev.shape[1], L.ptr(k_dev), len(cols), L.ptr(cols_dev),
                                             L.ptr(base_dev), L.ptr(scale_dev), L.ptr(f_dev), L.stream_handle()),
                    "fode_scale_columns")
        ev[it][0].record(stream)
        run.step_graph_only()
        ev[it][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = [a.elapsed_time(b) for a, b in ev]
    counts = run.memory.counts() - before
    cpg = np.minimum(32, C - 32 * np.arange(run.ngroups))
    attempts = float((counts[:, 0] * cpg).sum())
    rec = run.memory.read()
    out = {"scenario": scenario, "cells": C, "species": m.mech.n, "reactions": m.mech.R, "intervals": K,
           "dt_s": scn.dt, "value": round(C * K / wall, 1), "unit": "cell-intervals/s",
           "ms_per_interval": round(1e3 * wall / K, 3), "avg_launch_ms": round(sum(launch_ms) / K, 3),
           "attempts_per_group_interval": round(float(counts[:, 0].sum()) / (run.ngroups * K), 3),
           "us_per_cell_attempt": round(1e3 * sum(launch_ms) / attempts, 4),
           "failed_groups": int((rec["fails"] > 0).sum()), "autotuned": tuned,
           "lu": {"nnz_J": m.mech.nnz, "core": m.mech._plan.plan.m, "tile": m.mech._plan.plan.tile},
           "note": "the scenario's 64 recorded cells tiled; atol 1e-17, rtol 1e-3"}
    del run
    return out


def constrained_leg(fa, torch, W, C, K, args):
    """The constrained workload (100 species, 250 reactions, 3 conservation
    rows replacing species rows, RODAS3 with the projection at every call)
    on C cells over the timed step count, autotuned on its first 16,384
    cells, one graph replay per interval from the initial state."""
    from fastode_amd.system import ChemSystem
    dev = torch.device("cuda", 0)
    wl = W.constrained100(C=C, n_intervals=K)
    k_dev = k_resident(wl, torch, dev)
    tuned = None
    if not args.no_tune:
        from fastode_amd.autotune import tune
        n_s = min(C, 16384)
        ch = tune(wl.mech, k_dev[:n_s] if k_dev.shape[0] > 1 else k_dev, wl.Y0[:n_s], wl.dt, wl.atol, wl.rtol,
                  tableau=fa.RODAS3, kmul=wl.kmul(0), mass=wl.mass, use_cache=False)
        tuned = {"c": ch.c, "tile": ch.tile, "stage_tables": ch.stage_tables,
                 "blocks_per_cu": ch.blocks_per_cu or "max", "us_per_cell_attempt": round(ch.us_per_cell_attempt, 3)}
        progress(f"constrained tuned: {tuned}")
    system = ChemSystem(wl.mech, k_dev, mass=wl.mass)
    run, wall, launch_ms, counts, rec, cpg = timed_run(fa, torch, wl, fa.RODAS3, system, K, 0, args.blocks_per_cu,
                                                       not args.no_order)
    attempts = float((counts[:, 0] * cpg).sum())
    out = {"cells": C, "species": wl.n, "reactions": wl.R, "constraints": int(len(wl.mech.constraint_rows)),
           "intervals": K, "dt_s": wl.dt, "tableau": wl.tableau, "value": round(C * K / wall, 1),
           "unit": "cell-intervals/s", "ms_per_interval": round(1e3 * wall / K, 3),
           "avg_launch_ms": round(sum(launch_ms) / K, 3),
           "attempts_per_group_interval": round(float(counts[:, 0].sum()) / (run.ngroups * K), 3),
           "projection_passes_per_group_interval": round(float(counts[:, 10].sum()) / (run.ngroups * K), 3),
           "us_per_cell_attempt": round(1e3 * sum(launch_ms) / attempts, 4),
           "failed_groups": int((rec["fails"] > 0).sum()), "autotuned": tuned,
           "lu": {"nnz_J": wl.mech.nnz, "core": wl.mech._plan.plan.m, "tile": wl.mech._plan.plan.tile}}
    del run
    return out


def uneven_leg(fa, torch, W, C, K, tuned_mech, args):
    """The 262,144-cell workload of uneven stiffness (per-cell temperature
    and rate multipliers) over the timed step count: throughput, and how
    unevenly the groups' work falls (attempts per group-interval) against
    what one persistent workgroup's fair share of a launch is."""
    from fastode_amd.system import ChemSystem
    dev = torch.device("cuda", 0)
    wl = W.uneven200(C=C, n_intervals=K)
    tab = fa.ROS3
    choice = getattr(tuned_mech, "_tuned", None)
    if choice is not None:        # same pattern as the headline mechanism
        from fastode_amd.autotune import install
        install(wl.mech, choice)
    system = ChemSystem(wl.mech, k_resident(wl, torch, dev), mass=wl.mass)
    run, wall, launch_ms, counts, rec, cpg = timed_run(fa, torch, wl, tab, system, K, 0, args.blocks_per_cu,
                                                       not args.no_order)
    per_group = counts[:, 0] / K                      # attempts per group-interval, each group's mean
    work = counts[:, 0] * cpg                         # cell-attempts per group over the run
    from fastode_amd import _lib as L
    cus = L.device_info()["cus"]
    fair = work.sum() / max(1, min(run.ngroups, cus))   # one workgroup per CU at this LDS size
    out = {"cells": C, "groups": run.ngroups, "intervals": K, "value": round(C * K / wall, 1),
           "unit": "cell-intervals/s", "ms_per_interval": round(1e3 * wall / K, 3),
           "launch_ms_min_median_max": [round(min(launch_ms), 1), round(float(np.median(launch_ms)), 1),
                                        round(max(launch_ms), 1)],
           "attempts_per_group_interval": {
               "mean": round(float(per_group.mean()), 3), "min": round(float(per_group.min()), 3),
               "median": round(float(np.median(per_group)), 3), "p99": round(float(np.percentile(per_group, 99)), 3),
               "max": round(float(per_group.max()), 3)},
           "largest_group_over_fair_share": round(float(work.max() / fair), 4),
           "cell_attempts_per_s": round(float(work.sum()) / wall, 1),
           "failed_groups": int((rec["fails"] > 0).sum()),
           "note": "per-group spread of the work the atomic queue balances (costliest-first order from the "
                   "previous interval); largest_group_over_fair_share << 1 means no single group can "
                   "hold up a launch"}
    del run
    return out


if __name__ == "__main__":
    main()
