"""GPU: TS1 scenarios through the device integrator against the reference's
recorded TS1 experiments.

The reference ran every configuration on the first 64 cells of each
scenario in groups of 32 and recorded the summed Counts and the digits
against the MICM reference solutions (reference-results/ts1/runs.csv,
experiments/ts1_experiments.py; reference-results/dae/ts1_families.csv,
experiments/ts1_dae.py).  Here the plain drop-in (``run_scenario`` with
``Algorithm`` / ``RestartingLoop`` / ``BackwardEuler``) and the batched form
(``run_scenario_batched``) must give those Counts exactly and the digits to
0.01 (the recorded values have 4 decimals; the device LU rounds differently
from SuperLU).
"""
import csv
import time
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden" / "ts1"
FLOOR = 1e-14
COUNT_COLS = ["steps", "accepted", "rejected", "rhs", "jac", "lu", "solves", "newton", "failures",
              "unconverged"]
FAMILY4 = ["O1D", "O", "OH", "HO2"]     # experiments/dae_experiments.py:342-347

# the configurations of experiments/ts1_experiments.py:48-69 (solver kind,
# options, atol; None is rtol times the scoring floor)
CONFIGS = {
    "alg": ("algorithm", {}, None),
    "alg-a20": ("algorithm", {}, 1e-20),
    "alg-a15": ("algorithm", {}, 1e-15),
    "alg-elementary": ("algorithm", {"controller": "elementary"}, None),
    "alg-reset": ("algorithm", {"carry": False, "quantize": False}, None),
    "alg-noquant": ("algorithm", {"quantize": False}, None),
    "alg-nomemory": ("algorithm", {"carry_memory": False}, None),
    "restart-p": ("restarting", {}, None),
    "restart": ("restarting", {}, 1e-15),
    "restart-a20": ("restarting", {}, 1e-20),
    "be": ("backward-euler", {}, 1e-3),
    "be-tight": ("backward-euler", {}, 1e-15),
    # the absolute-tolerance sweep (reference-results/ts1/atol.csv)
    **{f"alg-a{e}": ("algorithm", {}, 10.0 ** -e) for e in (12, 13, 14, 16, 17, 18)},
    "restart-a13": ("restarting", {}, 1e-13),
}
RECORDS = ("runs.csv", "atol.csv")     # reference-results/ts1/, copied verbatim
SCENARIOS = ["tier1-30s", "diurnal-30min", "diurnal-10min", "diurnal-5min", "polluted-10min", "clean-10min",
             "summer-15min", "tropics-20min", "winter60-10min", "polluted-48h-30min", "clean-48h-5min"]

_TS1 = {}


def ts1():
    import fastode_amd.ts1 as T
    if "m" not in _TS1:
        _TS1["m"] = T.TS1()
    return _TS1["m"]


def recorded_runs():
    rows = []
    for name in RECORDS:
        with open(GOLD / name, newline="") as f:
            rows += list(csv.DictReader(f))
    return rows


def runs_row(config, rtol, scenario):
    for r in recorded_runs():
        if r["config"] == config and float(r["rtol"]) == rtol and r["scenario"] == scenario:
            return r
    raise KeyError((config, rtol, scenario))


def solver_for(kind, options, batched, tableau=None, blocks_per_cu=0):
    import fastode_amd as fa
    tab = tableau or fa.ROS3
    if batched:
        if kind == "algorithm":
            return fa.BatchedAlgorithm(tab, blocks_per_cu=blocks_per_cu, **options)
        if kind == "restarting":
            return fa.BatchedRestartingLoop(tab, blocks_per_cu=blocks_per_cu)
        return fa.BatchedBackwardEuler(blocks_per_cu=blocks_per_cu)
    if kind == "algorithm":
        return fa.Algorithm(tab, **options)
    if kind == "restarting":
        return fa.RestartingLoop(tab)
    return fa.BackwardEuler()


def run(config, rtol, scenario, batched, cells=64, blocks_per_cu=0):
    import fastode_amd.ts1 as T
    kind, options, atol = CONFIGS[config]
    atol = rtol * FLOOR if atol is None else atol
    m = ts1()
    scn = T.Scenario.from_npz(GOLD / f"{scenario}.npz")
    cl = scn.make_cells(m, which=range(cells))
    solver = solver_for(kind, options, batched, blocks_per_cu=blocks_per_cu)
    t = time.time()
    if batched:
        snaps, counts, statuses, _ = T.run_scenario_batched(scn, solver, cl, m, rtol, atol)
    else:
        snaps, counts, statuses = T.run_scenario(scn, solver, cl, m, rtol, atol)
    secs = time.time() - t
    D, per, gate = T.score(snaps, range(cells), m.species, scn.reference())
    return snaps, counts, statuses, D, per, gate, secs


def compare_row(config, rtol, scenario, batched, slack=0.0):
    """Run one recorded configuration; return its snapshots, a one-line
    summary and the list of ways it differs from the record (empty: exact
    Counts, digits within 0.01, the same worst species, gate and failed
    calls).  slack > 0 lets each count differ by that fraction of the
    recorded value."""
    snaps, counts, statuses, D, per, gate, secs = run(config, rtol, scenario, batched)
    ref = runs_row(config, rtol, scenario)
    got = {c: getattr(counts, c) for c in COUNT_COLS}
    want = {c: int(ref[c]) for c in COUNT_COLS}
    worst = min(per, key=per.get)
    problems = []
    off = [c for c in COUNT_COLS if abs(got[c] - want[c]) > slack * want[c]]
    if off:
        problems.append("counts " + ", ".join(f"{c} {got[c]} (recorded {want[c]})" for c in off))
    if abs(D - float(ref["digits"])) > 0.01:
        problems.append(f"digits {D:.4f} (recorded {ref['digits']})")
    if worst != ref["worst_species"]:
        problems.append(f"worst species {worst} (recorded {ref['worst_species']})")
    if str(gate) != ref["gate"]:
        problems.append(f"gate {gate} (recorded {ref['gate']})")
    if len(statuses) != int(ref["failed_calls"]):
        problems.append(f"failed calls {len(statuses)} (recorded {ref['failed_calls']})")
    line = (f"{'batched' if batched else 'plain'} {config} rtol {rtol:g} {scenario}: steps {got['steps']} "
            f"(recorded {want['steps']}), digits {D:.4f} (recorded {ref['digits']}), {secs:.2f} s")
    return snaps, line, problems


def check_row(config, rtol, scenario, batched):
    snaps, line, problems = compare_row(config, rtol, scenario, batched)
    print(line)
    assert not problems, problems
    return snaps


PLAIN_CASES = [("alg", 1e-3, "tier1-30s"), ("alg", 1e-3, "diurnal-10min"), ("alg-a20", 1e-3, "tier1-30s"),
               ("alg-elementary", 2e-3, "tier1-30s"), ("alg-noquant", 2e-3, "tier1-30s"),
               ("alg-nomemory", 2e-3, "tier1-30s"), ("alg-reset", 1e-2, "tier1-30s"),
               ("restart-p", 1e-3, "tier1-30s"), ("restart", 1e-3, "tier1-30s"), ("be", 1e-6, "tier1-30s")]


@pytest.mark.parametrize("config,rtol,scenario", PLAIN_CASES,
                         ids=[f"{c}-{r:g}-{s}" for c, r, s in PLAIN_CASES])
def test_plain_ts1_matches_recorded_runs(config, rtol, scenario):
    check_row(config, rtol, scenario, batched=False)


BATCHED_CASES = ([("alg", 1e-3, s) for s in SCENARIOS]
                 + [("alg", r, "diurnal-10min") for r in (3e-2, 5e-4)]
                 + [(c, r, "diurnal-10min") for c in ("alg-elementary", "alg-reset", "alg-noquant",
                                                      "alg-nomemory") for r in (1e-2, 5e-4)]
                 + [(c, r, "tier1-30s") for c in ("alg-elementary", "alg-noquant", "alg-nomemory")
                    for r in (2e-3,)]
                 + [("restart-p", r, s) for s in ("tier1-30s", "diurnal-10min") for r in (1e-2, 1e-3)]
                 + [("restart", 1e-3, "tier1-30s"), ("be", 1e-6, "tier1-30s"), ("be", 1e-6, "diurnal-10min")])


@pytest.mark.parametrize("config,rtol,scenario", BATCHED_CASES,
                         ids=[f"{c}-{r:g}-{s}" for c, r, s in BATCHED_CASES])
def test_batched_ts1_matches_recorded_runs(config, rtol, scenario):
    check_row(config, rtol, scenario, batched=True)


@pytest.mark.parametrize("config,rtol", [("alg", 1e-3), ("restart-p", 1e-3), ("be", 1e-6)])
def test_batched_ts1_states_equal_the_plain_path(config, rtol):
    """The batched form fuses each cell's attempt into one workgroup; its
    states are those of the group-by-group drop-in, bit for bit."""
    a = run(config, rtol, "tier1-30s", batched=False)[0]
    b = run(config, rtol, "tier1-30s", batched=True)[0]
    for step in a:
        assert np.array_equal(a[step], b[step]), (config, step, np.abs(a[step] - b[step]).max())


@pytest.mark.parametrize("config,rtol", [("alg", 1e-3), ("restart-p", 1e-3), ("be", 1e-6)])
def test_batched_ts1_two_cells_per_cu(config, rtol):
    """The kernel held to two waves per SIMD, two workgroups (cells) per CU,
    tables in memory: the recorded Counts, and the one-wave kernel's states
    bit for bit (the same arithmetic, other registers)."""
    a = run(config, rtol, "tier1-30s", batched=True, blocks_per_cu=2)
    b = run(config, rtol, "tier1-30s", batched=True)
    ref = runs_row(config, rtol, "tier1-30s")
    assert {c: getattr(a[1], c) for c in COUNT_COLS} == {c: int(ref[c]) for c in COUNT_COLS}
    for step in a[0]:
        assert np.array_equal(a[0][step], b[0][step]), (config, step)


@pytest.mark.parametrize("config,rtol,rho,split", [("alg", 1e-3, 0.7, (177, 2)), ("be", 1e-6, 0.7, (177, 2)),
                                                    ("alg", 1e-3, 0.95, (194, 1))])
def test_batched_ts1_two_cells_per_cu_with_solve_tables_in_lds(config, rtol, rho, split):
    """The tuned TS1 plan (rho 0.7: 177 sparse pivots, TILE 2) leaves room on
    a CU for two cells with the solves' tables copied into LDS, so the
    two-waves-per-SIMD kernel reads its solve stages from LDS.  Under the
    sparsest split (194 pivots, TILE 1) every table would exceed a
    workgroup's LDS and the solves' tables two cells' share, so the kernel
    keeps its tables in memory (and the refused layout leaves no error
    behind).  Either way the one-wave kernel's Counts and states, bit for
    bit, and under the tuned plan the recorded Counts."""
    from fastode_amd.device import DevicePlan
    mech = ts1().mech
    saved = mech.__dict__.get("_plan")
    mech._plan = DevicePlan(mech.n, mech.rows, mech.cols, rho=rho, m_max=160)
    try:
        assert (mech._plan.plan.c, mech._plan.plan.tile) == split
        a = run(config, rtol, "tier1-30s", batched=True, blocks_per_cu=2)
        b = run(config, rtol, "tier1-30s", batched=True, blocks_per_cu=1)
    finally:
        if saved is None:
            del mech._plan
        else:
            mech._plan = saved
    assert {c: getattr(a[1], c) for c in COUNT_COLS} == {c: getattr(b[1], c) for c in COUNT_COLS}
    if rho == 0.7:
        ref = runs_row(config, rtol, "tier1-30s")
        assert {c: getattr(a[1], c) for c in COUNT_COLS} == {c: int(ref[c]) for c in COUNT_COLS}
    for step in a[0]:
        assert np.array_equal(a[0][step], b[0][step]), (config, step)


def family_row(rtol, form, tableau, solver):
    with open(GOLD / "ts1_families.csv", newline="") as f:
        for r in csv.DictReader(f):
            if (float(r["rtol"]), r["form"], r["tableau"], r["solver"]) == (rtol, form, tableau, solver):
                return r
    raise KeyError((rtol, form, tableau, solver))


FAMILY_CASES = [(1e-3, "dae-4", "algorithm"), (1e-3, "ode", "algorithm"), (1e-3, "dae-4", "restarting"),
                (1e-3, "ode", "restarting"), (1e-4, "dae-4", "algorithm")]
FAMILIES = {"dae-4": FAMILY4,
            "dae-9": FAMILY4 + ["CH3O2", "C2H5O2", "CH3CO3", "C3H7O2", "PO2"]}


@pytest.mark.parametrize("batched", [False, True], ids=["plain", "batched"])
@pytest.mark.parametrize("rtol,form,solver", FAMILY_CASES, ids=[f"{f}-{s}-{r:g}" for r, f, s in FAMILY_CASES])
def test_ts1_constrained_families_match_recorded_runs(rtol, form, solver, batched):
    """QSSA families held algebraic (mass 0) with RODAS3 and the projection
    at every call (experiments/ts1_dae.py:42-61)."""
    import fastode_amd as fa
    import fastode_amd.ts1 as T
    m = ts1()
    scn = T.Scenario.from_npz(GOLD / "tier1-30s.npz")
    cl = scn.make_cells(m, which=range(64))
    fam = tuple(FAMILIES.get(form, ()))
    s = solver_for(solver, {}, batched, tableau=fa.RODAS3)
    if batched:
        snaps, counts, statuses, _ = T.run_scenario_batched(scn, s, cl, m, rtol, rtol * FLOOR, constrained=fam)
    else:
        snaps, counts, statuses = T.run_scenario(scn, s, cl, m, rtol, rtol * FLOOR, constrained=fam)
    D, per, gate = T.score(snaps, range(64), m.species, scn.reference_final)
    ref = family_row(rtol, form, "rodas3", solver)
    per_interval = scn.steps * 2
    print(f"{form} {solver} rtol {rtol:g} {'batched' if batched else 'plain'}: steps {counts.steps}, "
          f"digits {D:.3f} (recorded {ref['digits']})")
    assert round(counts.steps / per_interval, 3) == float(ref["steps_per_interval"])
    assert round(counts.lu / per_interval, 3) == float(ref["lu_per_interval"])
    assert round(counts.proj_passes / per_interval, 3) == float(ref["proj_passes_per_interval"])
    assert abs(D - float(ref["digits"])) <= 0.01
    assert min(per, key=per.get) == ref["worst"]
    if fam:
        assert abs(min(per[x] for x in fam if x in per) - float(ref["constrained_min_digits"])) <= 0.01
    assert len(statuses) == int(ref["failed_calls"])


def be_substeps_row(scenario, n):
    with open(GOLD / "be_substeps.csv", newline="") as f:
        for r in csv.DictReader(f):
            if r["scenario"] == scenario and int(r["substeps"]) == n:
                return r
    raise KeyError((scenario, n))


@pytest.mark.parametrize("scenario,n", [("tier1-30s", 4), ("tier1-30s", 16), ("tier1-30s", 64),
                                        ("diurnal-10min", 4), ("diurnal-10min", 16)])
def test_batched_substepped_backward_euler_matches_recorded_runs(scenario, n):
    """Backward Euler called n times per interval on equal pieces
    (experiments/be_substeps.py): Newton passes per group-interval exact,
    digits to 0.01 of reference-results/ts1/be_substeps.csv."""
    import fastode_amd as fa
    import fastode_amd.ts1 as T
    m = ts1()
    scn = T.Scenario.from_npz(GOLD / f"{scenario}.npz")
    cl = scn.make_cells(m, which=range(64))
    snaps, counts, statuses, _ = T.run_scenario_batched(scn, fa.BatchedBackwardEuler(), cl, m, 1e-6, 1e-3,
                                                        substeps=n)
    D, per, gate = T.score(snaps, range(64), m.species, scn.reference())
    ref = be_substeps_row(scenario, n)
    per_int = 1.0 / scn.steps / 2
    print(f"be x{n} {scenario}: newton/interval {counts.newton * per_int}, digits {D:.4f} (recorded {ref['digits']})")
    assert round(counts.newton * per_int, 2) == float(ref["newton_per_interval"])
    assert (counts.failures, counts.unconverged, len(statuses)) == (int(ref["failures"]), int(ref["unconverged"]),
                                                                    int(ref["failed_calls"]))
    assert abs(D - float(ref["digits"])) <= 0.01
    assert min(per, key=per.get) == ref["worst_species"]
