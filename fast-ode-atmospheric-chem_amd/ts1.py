"""MOZART-TS1 on the device integrator: the mechanism of a MICM
configuration, its rate constants, the cells and scenarios of the
reference's data set, the scenario driver and the digits score
(fastode/ts1.py).

* ``TS1(path)`` reads the configuration (``data/ts1.json`` by default) into a
  device ``Mechanism`` (third bodies folded into k) and a rate table.
* ``TS1.rate_constants(T, P, params)`` gives k (C, R) on the host: Arrhenius
  A exp(C/T) (T/D)^B (1 + E p); Troe k0 [M]/(1 + r) Fc^(N/(N + log10(r)^2))
  with r = k0 [M]/kinf; user-defined (parameter x scaling); surface
  4 N pi r^2 / (r/D + 4/(v gamma)) with v = sqrt(8 R T/(pi Mw)); then
  [M]^power, [M] = p/(R T).  It runs once per run (the intervals only
  rescale the photolysis columns), so it stays NumPy.
* ``Cells``, ``Scenario``: the reference's CSV / scenario.txt formats, or the
  same tables in one ``.npz`` bundle (``Scenario.from_npz``).
* ``run_scenario``: the reference's driver, group by group, each group over
  every interval with its own Memory, a new ChemSystem per interval
  (fastode/ts1.py:181-223); any solver with the reference's interface runs
  under it.
* ``run_scenario_batched``: the same scenario interval by interval, every
  group in one persistent launch (``BatchedAlgorithm`` and its restarting /
  backward-Euler forms); the photolysis columns of the device rate constants
  are rewritten by ``fode_scale_columns`` at each interval start.  It gives
  the group-by-group driver's Counts and states.
* ``score``: significant digits against reference solutions, overall and per
  species, with the finite / nonnegative gate (fastode/ts1.py:226-264).
"""
from __future__ import annotations

import csv
import ctypes as ct
import json
import math
from pathlib import Path

import numpy as np

from .counts import COUNT_FIELDS, Counts
from .mechanism import Mechanism

DATA = Path(__file__).with_name("data") / "ts1.json"
BOLTZMANN = 1.380649e-23
AVOGADRO = 6.02214076e23
R_GAS = BOLTZMANN * AVOGADRO


def _arrhenius(r, T, P, M):
    A, B, C = r.get("A", 1.0), r.get("B", 0.0), r.get("C", 0.0)
    D, E = r.get("D", 300.0), r.get("E", 0.0)
    return A * np.exp(C / T) * (T / D) ** B * (1.0 + E * P)


def _troe(r, T, P, M):
    def limit(prefix):
        return (r.get(prefix + "_A", 1.0) * np.exp(r.get(prefix + "_C", 0.0) / T)
                * (T / 300.0) ** r.get(prefix + "_B", 0.0))
    k0, kinf = limit("k0"), limit("kinf")
    Fc, N = r.get("Fc", 0.6), r.get("N", 1.0)
    ratio = k0 * M / kinf
    return k0 * M / (1.0 + ratio) * Fc ** (N / (N + np.log10(ratio) ** 2))


class TS1:
    """The mechanism of a MICM configuration (fastode/ts1.py:35-63)."""

    def __init__(self, path: Path = DATA):
        conf = json.loads(Path(path).read_text())
        entries = conf["camp-data"]
        specs = [e for e in entries if e.get("type") == "CHEM_SPEC"]
        mechanism = next(e for e in entries if e.get("type") == "MECHANISM")
        self.third = {s["name"] for s in specs if s.get("third body")}
        self.species = [s["name"] for s in specs if not s.get("third body")]
        self.props = {s["name"]: s for s in specs}
        self.init_conc = np.array([self.props[s].get("initial concentration [mol m-3]", 0.0)
                                   for s in self.species])
        init = conf.get("initial conditions", {})
        self.T0 = float(init.get("temperature", 298.15))
        self.P0 = float(init.get("pressure", 101325.0))
        self.params0 = {k: float(v) for k, v in conf.get("initial rate parameters", {}).items()}
        self.reactions = mechanism["reactions"]

        def equation(r):
            if r["type"] == "SURFACE":
                return [r["gas-phase reactant"]], r["gas-phase products"]
            return list(r["reactants"]), r.get("products", [])
        pairs = []
        for r in self.reactions:
            reac, prods = equation(r)
            pairs.append((reac, [(q["species"], float(q.get("yield", 1.0))) for q in prods]))
        self.mech = Mechanism(self.species, pairs, third_bodies=self.third)
        self.index = {s: i for i, s in enumerate(self.species)}
        self.photolysis_reactions = [j for j, r in enumerate(self.reactions)
                                     if r["type"] == "USER_DEFINED" and r["MUSICA name"].startswith("j")]

    def _surface(self, r, T, params):
        label = r["MUSICA name"]
        radius = params[label + ".effective radius [m]"]
        number = params[label + ".particle number concentration [# m-3]"]
        gas = self.props[r["gas-phase reactant"]]
        speed = np.sqrt(8.0 * R_GAS / (math.pi * gas["molecular weight [kg mol-1]"]) * T)
        gamma = r.get("reaction probability", 1.0)
        return (4.0 * number * math.pi * radius * radius
                / (radius / gas["diffusion coefficient [m2 s-1]"] + 4.0 / (speed * gamma)))

    def rate_constants(self, T, P, params):
        """k (C, R) at temperatures T, pressures P (C,) and rate parameters
        params: name -> (C,) (fastode/ts1.py:66-104)."""
        T = np.asarray(T, dtype=float)
        P = np.asarray(P, dtype=float)
        M = P / (R_GAS * T)
        k = np.empty((len(T), len(self.reactions)))
        for j, r in enumerate(self.reactions):
            kind = r["type"]
            if kind == "ARRHENIUS":
                k[:, j] = _arrhenius(r, T, P, M)
            elif kind == "TROE":
                k[:, j] = _troe(r, T, P, M)
            elif kind == "USER_DEFINED":
                k[:, j] = params[r["MUSICA name"]] * r.get("scaling factor", 1.0)
            elif kind == "SURFACE":
                k[:, j] = self._surface(r, T, params)
            else:
                raise ValueError(f"rate type {kind} is not one of TS1's")
        power = self.mech.third_body_power
        if np.any(power):
            k *= M[:, None] ** power[None, :]
        return k

    def photolysis_table(self):
        """(columns, parameter labels, scaling factors) of the photolysis
        reactions, whose k an interval start rewrites."""
        cols = np.asarray(self.photolysis_reactions, dtype=np.int32)
        labels = [self.reactions[j]["MUSICA name"] for j in self.photolysis_reactions]
        scale = np.array([self.reactions[j].get("scaling factor", 1.0) for j in self.photolysis_reactions])
        return cols, labels, scale


def read_csv_table(path):
    """(header, float rows) of a CSV (fastode/ts1.py:106-111)."""
    with open(path, newline="") as f:
        reader = csv.reader(f)
        header = next(reader)
        rows = [[float(v) for v in row] for row in reader if row]
    return header, np.array(rows)


class Cells:
    """The cells of a cells table: TS1's initial state, then each row's
    values (fastode/ts1.py:114-144)."""

    def __init__(self, ts1: TS1, path=None, which=None, table=None):
        header, rows = table if table is not None else read_csv_table(path)
        rows = np.asarray(rows, dtype=float)
        if which is not None:
            rows = rows[np.asarray(which)]
        C = rows.shape[0]
        self.C = C
        self.T = np.full(C, ts1.T0)
        self.P = np.full(C, ts1.P0)
        self.Y = np.tile(ts1.init_conc, (C, 1))
        self.params = {k: np.full(C, v) for k, v in ts1.params0.items()}
        for col, name in enumerate(header):
            name = str(name)
            if name == "temperature":
                self.T = rows[:, col].copy()
            elif name == "pressure":
                self.P = rows[:, col].copy()
            elif name in ts1.index:
                self.Y[:, ts1.index[name]] = rows[:, col]
            else:
                self.params[name] = rows[:, col].copy()

    def subset(self, sl):
        other = object.__new__(Cells)
        other.C = len(range(*sl.indices(self.C)))
        other.T = self.T[sl].copy()
        other.P = self.P[sl].copy()
        other.Y = self.Y[sl].copy()
        other.params = {k: v[sl].copy() for k, v in self.params.items()}
        return other


class Scenario:
    """A scenario: interval length, number of intervals, snapshot spacing, the
    photolysis factor table and the cells (fastode/ts1.py:147-178)."""

    def __init__(self, path=None):
        if path is None:
            return
        self.path = Path(path)
        values = {}
        for line in self.path.read_text().splitlines():
            line = line.split("#", 1)[0].strip()
            if line:
                key, value = (s.strip() for s in line.split("=", 1))
                values[key] = value
        base = self.path.parent
        self.name = base.name
        self.cells = base / values["cells"]
        self.dt = float(values["dt"])
        self.steps = int(values["steps"])
        self.snapshot_every = int(values.get("snapshot_every", 0))
        _, table = read_csv_table(base / values["photolysis"])
        self._ft, self._fv = table[:, 0], table[:, 1]
        self.reference_snapshots = base / "reference_snapshots.csv"
        self.reference_final = base / "reference_final.csv"
        self._tables = None

    @classmethod
    def from_npz(cls, path):
        """A scenario bundled as one .npz: dt, steps, snapshot_every,
        photolysis (time_s, factor), cells_header/cells, final_header/final
        and optionally snap_header/snaps (the reference's CSV tables)."""
        self = cls()
        with np.load(path, allow_pickle=False) as z:
            t = {k: z[k] for k in z.files}
        self.path = Path(path)
        self.name = self.path.stem
        self.dt = float(t["dt"])
        self.steps = int(t["steps"])
        self.snapshot_every = int(t["snapshot_every"])
        self._ft, self._fv = t["photolysis"][:, 0], t["photolysis"][:, 1]
        self.cells = ([str(h) for h in t["cells_header"]], t["cells"])
        self._tables = t
        self.reference_final = ([str(h) for h in t["final_header"]], t["final"])
        self.reference_snapshots = (([str(h) for h in t["snap_header"]], t["snaps"])
                                    if "snaps" in t else None)
        return self

    def make_cells(self, ts1, which=None):
        if isinstance(self.cells, tuple):
            return Cells(ts1, which=which, table=self.cells)
        return Cells(ts1, self.cells, which=which)

    def reference(self):
        """The reference solutions the experiments score against: the
        snapshots when the scenario keeps them, else the final states."""
        return self.reference_snapshots if self.snapshot_every else self.reference_final

    def factor(self, t):
        return float(np.interp(t, self._ft, self._fv))

    def snapshot_steps(self):
        if self.snapshot_every <= 0:
            return [self.steps]
        steps = list(range(self.snapshot_every, self.steps + 1, self.snapshot_every))
        if not steps or steps[-1] != self.steps:
            steps.append(self.steps)
        return steps


def _plan_run(scenario, ts1, atol, constrained, steps):
    n = ts1.mech.n
    atol_v = np.full(n, atol) if np.isscalar(atol) else np.asarray(atol, dtype=float)
    mass = np.ones(n)
    for s in constrained:
        mass[ts1.index[s]] = 0.0
    nsteps = scenario.steps if steps is None else steps
    snap_steps = [s for s in scenario.snapshot_steps() if s <= nsteps] if steps is None else [nsteps]
    return atol_v, mass, nsteps, snap_steps


def run_scenario(scenario: Scenario, solver, cells: Cells, ts1: TS1, rtol, atol, group_size=32,
                 constrained=(), warmup_steps=0, steps=None, on_step=None):
    """Run a solver on a scenario's cells in groups that share a step
    (fastode/ts1.py:181-223).  Returns (snapshots: step -> (C, n) NumPy,
    counts, statuses: (first cell, interval, status) of every call that did
    not converge)."""
    from .system import ChemSystem
    n = ts1.mech.n
    atol_v, mass, nsteps, snap_steps = _plan_run(scenario, ts1, atol, constrained, steps)
    snapshots = {s: np.empty((cells.C, n)) for s in snap_steps}
    counts = Counts()
    statuses = []
    photolysis = [k for k in cells.params if k.startswith("j")]
    for g0 in range(0, cells.C, group_size):
        sl = slice(g0, min(g0 + group_size, cells.C))
        group = cells.subset(sl)
        base = {k: group.params[k].copy() for k in photolysis}
        Y = group.Y.copy()
        memory = solver.new_memory()
        k_static = None
        for step in range(-warmup_steps, nsteps):
            f = scenario.factor(max(step, 0) * scenario.dt)
            for name in photolysis:
                group.params[name] = base[name] * f
            if k_static is None:
                k_static = ts1.rate_constants(group.T, group.P, group.params)
                k = k_static
            else:
                k = k_static.copy()
                for j in ts1.photolysis_reactions:
                    r = ts1.reactions[j]
                    k[:, j] = group.params[r["MUSICA name"]] * r.get("scaling factor", 1.0)
            system = ChemSystem(ts1.mech, k, mass=mass, dense=False)
            Y, status = solver.solve(system, Y, scenario.dt, atol_v, rtol, counts, memory)
            if status != "converged":
                statuses.append((g0, step, status))
            if on_step is not None:
                on_step(g0, step, Y)
            if step + 1 in snapshots:
                snapshots[step + 1][sl] = Y
    return snapshots, counts, statuses


def run_scenario_batched(scenario: Scenario, solver, cells: Cells, ts1: TS1, rtol, atol, group_size=32,
                         constrained=(), warmup_steps=0, steps=None, graph=True, on_step=None, substeps=1):
    """The scenario with every group in one persistent launch per interval.
    solver: ``BatchedAlgorithm`` / ``BatchedRestartingLoop`` /
    ``BatchedBackwardEuler`` (its group_size is used).  Rate constants are
    computed once for all cells (the same per-cell arithmetic as the group
    loop's), moved to the device, and the photolysis columns rewritten on the
    device from the second interval on.  substeps > 1 calls the solver that
    many times per interval on equal pieces of it, the photolysis held over
    the interval (the sub-stepped backward Euler of
    experiments/be_substeps.py:33-52).  Returns what ``run_scenario``
    returns, plus the runner (device state, per-group Memory and Counts)."""
    import torch

    from . import _lib as L
    from .system import ChemSystem, as_device
    group_size = solver.group_size
    n = ts1.mech.n
    atol_v, mass, nsteps, snap_steps = _plan_run(scenario, ts1, atol, constrained, steps)
    photolysis = [k for k in cells.params if k.startswith("j")]
    f0 = scenario.factor(0.0)
    params = dict(cells.params)
    for name in photolysis:
        params[name] = cells.params[name] * f0
    k_dev = as_device(ts1.rate_constants(cells.T, cells.P, params))
    cols, labels, scale = ts1.photolysis_table()
    cols_dev = torch.as_tensor(cols, device=k_dev.device)
    base_dev = as_device(np.stack([cells.params[lab] for lab in labels], axis=1) if labels
                         else np.zeros((cells.C, 0)))
    scale_dev = as_device(scale)
    f_dev = torch.zeros(1, dtype=torch.float64, device=k_dev.device)
    system = ChemSystem(ts1.mech, k_dev, mass=mass, dense=False)
    runner = solver.runner(system, cells.Y, scenario.dt / substeps, atol_v, rtol, graph=graph)
    runner.set_kmul(None)
    snapshots = {}
    statuses = []
    first = True
    for step in range(-warmup_steps, nsteps):
        if not first:
            f_dev.fill_(scenario.factor(max(step, 0) * scenario.dt))
            L.check(L.lib.fode_scale_columns(cells.C, k_dev.shape[1], L.ptr(k_dev), len(cols), L.ptr(cols_dev),
                                             L.ptr(base_dev), L.ptr(scale_dev), L.ptr(f_dev),
                                             L.stream_handle()), "fode_scale_columns")
        first = False
        failed = {}
        for _ in range(substeps):
            runner.step_graph_only()
            st = runner.memory.read()["status"]
            for g in np.flatnonzero(st != L.CONVERGED):
                failed[int(g)] = L.STATUS[int(st[g])]
        for g, status in sorted(failed.items()):
            statuses.append((g * group_size, step, status))
        if on_step is not None:
            Yh = runner.Y.cpu().numpy()
            for g0 in range(0, cells.C, group_size):
                on_step(g0, step, Yh[g0:g0 + group_size])
        if step + 1 in snap_steps:
            snapshots[step + 1] = runner.Y.cpu().numpy()
    counts = Counts()
    for name, v in zip(COUNT_FIELDS, runner.memory.counts().sum(axis=0)):
        setattr(counts, name, int(v))
    return snapshots, counts, statuses, runner


def score(result: dict, cells_index, species, reference, threshold=1e-14, steps=None):
    """Digits overall and per species (Sandu et al.) and the gate over the
    given cells: result maps step -> (C, n) in the order of cells_index;
    reference is a CSV path or a (header, rows) table, keyed by (step, cell)
    or one row per cell (fastode/ts1.py:226-264).  Returns (overall,
    per-species dict, gate)."""
    header, rows = reference if isinstance(reference, tuple) else read_csv_table(reference)
    header = [str(h) for h in header]
    keyed = header[:2] == ["step", "cell"]
    start = 4 if keyed else 2
    ref_species = header[start:]
    column = {s: start + i for i, s in enumerate(ref_species)}
    where = {s: i for i, s in enumerate(species)}
    want = list(cells_index)
    position = {c: i for i, c in enumerate(want)}
    squares = {s: [] for s in ref_species}
    gate = True
    for step, Y in result.items():
        if not np.all(np.isfinite(Y)) or np.any(Y < -1e-20):
            gate = False
        if keyed:
            chosen = rows[(rows[:, 0] == step) & np.isin(rows[:, 1], want)]
            pairs = [(Y[position[int(row[1])]], row) for row in chosen]
        else:
            pairs = [(Y[i], rows[c]) for i, c in enumerate(want)]
        for y_cell, row in pairs:
            for s in ref_species:
                ref = row[column[s]]
                if ref < threshold or s not in where:
                    continue
                squares[s].append(((y_cell[where[s]] - ref) / ref) ** 2)

    def digits(values):
        if not values:
            return float("nan")
        ms = float(np.mean(values))
        return 16.0 if ms <= 0.0 else min(16.0, -math.log10(math.sqrt(ms)))

    per = {s: digits(v) for s, v in squares.items() if v}
    return digits([v for vs in squares.values() for v in vs]), per, gate
