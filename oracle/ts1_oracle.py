"""ORACLE — test infrastructure only (see fastode_oracle.py's header).

A CPU restatement of the reference's TS1 driver (fastode/ts1.py): the MICM
configuration parse, the rate constants, the cells, the scenarios, the
group loop of run_scenario and the digits score.  It drives the oracle's
solvers (fastode_oracle.Algorithm / RestartingLoop / BackwardEuler) with the
oracle's SuperLU ChemSystem, so that its Counts can be checked against the
reference's recorded runs (reference-results/ts1/runs.csv).

Scenario data come from the fixtures tests/golden/ts1/<name>.npz (the
reference's CSVs parsed as fastode/ts1.py:106-111 parses them).
"""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np

import fastode_oracle as O

BOLTZMANN = 1.380649e-23              # fastode/ts1.py:28-30
AVOGADRO = 6.02214076e23
R_GAS = BOLTZMANN * AVOGADRO


class TS1:
    """The mechanism of a MICM configuration (fastode/ts1.py:35-63)."""

    def __init__(self, path):
        d = json.loads(Path(path).read_text())
        cd = d["camp-data"]
        specs = [e for e in cd if e.get("type") == "CHEM_SPEC"]
        mech = next(e for e in cd if e.get("type") == "MECHANISM")
        self.third = {s["name"] for s in specs if s.get("third body")}
        self.species = [s["name"] for s in specs if not s.get("third body")]
        self.props = {s["name"]: s for s in specs}
        self.init_conc = np.array([self.props[s].get("initial concentration [mol m-3]", 0.0)
                                   for s in self.species])
        ic = d.get("initial conditions", {})
        self.T0 = float(ic.get("temperature", 298.15))
        self.P0 = float(ic.get("pressure", 101325.0))
        self.params0 = {k: float(v) for k, v in d.get("initial rate parameters", {}).items()}
        self.reactions = mech["reactions"]
        pairs = []
        for r in self.reactions:
            if r["type"] == "SURFACE":                            # ts1.py:52-54
                reac = [r["gas-phase reactant"]]
                prods = [(p["species"], float(p.get("yield", 1.0))) for p in r["gas-phase products"]]
            else:
                reac = list(r["reactants"])
                prods = [(p["species"], float(p.get("yield", 1.0))) for p in r.get("products", [])]
            pairs.append((reac, prods))
        self.pairs = pairs
        self.mech = O.OracleMechanism(self.species, pairs, third_bodies=self.third)
        self.index = {s: i for i, s in enumerate(self.species)}
        self.photolysis_reactions = [j for j, r in enumerate(self.reactions)
                                     if r["type"] == "USER_DEFINED" and r["MUSICA name"].startswith("j")]

    def rate_constants(self, T, P, params):
        """k (C, R) (fastode/ts1.py:66-104), reaction by reaction."""
        T = np.asarray(T, dtype=float)
        P = np.asarray(P, dtype=float)
        M = P / (R_GAS * T)
        k = np.empty((len(T), len(self.reactions)))
        for j, r in enumerate(self.reactions):
            t = r["type"]
            if t == "ARRHENIUS":
                A, B, C = r.get("A", 1.0), r.get("B", 0.0), r.get("C", 0.0)
                D, E = r.get("D", 300.0), r.get("E", 0.0)
                k[:, j] = A * np.exp(C / T) * (T / D) ** B * (1.0 + E * P)
            elif t == "TROE":
                k0 = r.get("k0_A", 1.0) * np.exp(r.get("k0_C", 0.0) / T) * (T / 300.0) ** r.get("k0_B", 0.0)
                ki = r.get("kinf_A", 1.0) * np.exp(r.get("kinf_C", 0.0) / T) * (T / 300.0) ** r.get("kinf_B", 0.0)
                Fc, N = r.get("Fc", 0.6), r.get("N", 1.0)
                ratio = k0 * M / ki
                k[:, j] = k0 * M / (1.0 + ratio) * Fc ** (N / (N + np.log10(ratio) ** 2))
            elif t == "USER_DEFINED":
                k[:, j] = params[r["MUSICA name"]] * r.get("scaling factor", 1.0)
            elif t == "SURFACE":
                lab = r["MUSICA name"]
                rad = params[lab + ".effective radius [m]"]
                num = params[lab + ".particle number concentration [# m-3]"]
                sp = self.props[r["gas-phase reactant"]]
                Dc = sp["diffusion coefficient [m2 s-1]"]
                mw = sp["molecular weight [kg mol-1]"]
                v = np.sqrt(8.0 * R_GAS / (math.pi * mw) * T)
                gam = r.get("reaction probability", 1.0)
                k[:, j] = 4.0 * num * math.pi * rad * rad / (rad / Dc + 4.0 / (v * gam))
            else:
                raise ValueError(t)
        pw = self.mech.third_body_power
        if np.any(pw):
            k *= M[:, None] ** pw[None, :]
        return k


class Cells:
    """Cells of a cells table (fastode/ts1.py:114-144)."""

    def __init__(self, ts1, header, rows):
        C = rows.shape[0]
        self.C = C
        self.T = np.full(C, ts1.T0)
        self.P = np.full(C, ts1.P0)
        self.Y = np.tile(ts1.init_conc, (C, 1))
        self.params = {k: np.full(C, v) for k, v in ts1.params0.items()}
        for col, name in enumerate(header):
            if name == "temperature":
                self.T = rows[:, col].copy()
            elif name == "pressure":
                self.P = rows[:, col].copy()
            elif name in ts1.index:
                self.Y[:, ts1.index[name]] = rows[:, col]
            else:
                self.params[name] = rows[:, col].copy()

    def subset(self, sl):
        o = object.__new__(Cells)
        o.C = len(range(*sl.indices(self.C)))
        o.T, o.P, o.Y = self.T[sl].copy(), self.P[sl].copy(), self.Y[sl].copy()
        o.params = {k: v[sl].copy() for k, v in self.params.items()}
        return o


class Scenario:
    """A scenario fixture (fastode/ts1.py:147-178)."""

    def __init__(self, npz):
        with np.load(npz, allow_pickle=False) as z:
            self.z = {k: z[k] for k in z.files}
        self.name = Path(npz).stem
        self.dt = float(self.z["dt"])
        self.steps = int(self.z["steps"])
        self.snapshot_every = int(self.z["snapshot_every"])
        self._ft, self._fv = self.z["photolysis"][:, 0], self.z["photolysis"][:, 1]

    def cells(self, ts1, which=None):
        rows = self.z["cells"] if which is None else self.z["cells"][np.asarray(which)]
        return Cells(ts1, [str(h) for h in self.z["cells_header"]], rows)

    def factor(self, t):
        return float(np.interp(t, self._ft, self._fv))

    def snapshot_steps(self):
        if self.snapshot_every <= 0:
            return [self.steps]
        s = list(range(self.snapshot_every, self.steps + 1, self.snapshot_every))
        if not s or s[-1] != self.steps:
            s.append(self.steps)
        return s

    def reference(self):
        """(header, rows) of the snapshots when the scenario has them, else of
        the final states (the choice of experiments/ts1_experiments.py:98)."""
        if self.snapshot_every:
            return [str(h) for h in self.z["snap_header"]], self.z["snaps"]
        return [str(h) for h in self.z["final_header"]], self.z["final"]


def run_scenario(scn, solver, cells, ts1, rtol, atol, group_size=32, constrained=(), steps=None):
    """Groups of group_size cells, each over all intervals with its own
    Memory; photolysis parameters rescaled at every interval start, the
    static rate constants computed once per group (fastode/ts1.py:181-223).
    Returns (snapshots, counts, statuses)."""
    n = ts1.mech.n
    atol_v = np.full(n, atol) if np.isscalar(atol) else np.asarray(atol)
    mass = np.ones(n)
    for s in constrained:
        mass[ts1.index[s]] = 0.0
    nsteps = scn.steps if steps is None else steps
    snap_steps = [s for s in scn.snapshot_steps() if s <= nsteps] if steps is None else [nsteps]
    snaps = {s: np.empty((cells.C, n)) for s in snap_steps}
    counts = O.Counts()
    statuses = []
    phot = [k for k in cells.params if k.startswith("j")]
    for g0 in range(0, cells.C, group_size):
        sl = slice(g0, min(g0 + group_size, cells.C))
        grp = cells.subset(sl)
        base = {k: grp.params[k].copy() for k in phot}
        Y = grp.Y.copy()
        mem = solver.new_memory()
        k_static = None
        for step in range(nsteps):
            f = scn.factor(step * scn.dt)
            for name in phot:
                grp.params[name] = base[name] * f
            if k_static is None:
                k_static = ts1.rate_constants(grp.T, grp.P, grp.params)
                k = k_static
            else:
                k = k_static.copy()
                for j in ts1.photolysis_reactions:
                    r = ts1.reactions[j]
                    k[:, j] = grp.params[r["MUSICA name"]] * r.get("scaling factor", 1.0)
            system = O.ChemSystem(ts1.mech, k, mass=mass, dense=False)
            Y, st = solver.solve(system, Y, scn.dt, atol_v, rtol, counts, mem)
            if st != "converged":
                statuses.append((g0, step, st))
            if step + 1 in snaps:
                snaps[step + 1][sl] = Y
    return snaps, counts, statuses


def score(result, cells_index, species, header, rows, threshold=1e-14):
    """Significant digits overall and per species, and the finite /
    nonnegative gate (fastode/ts1.py:226-264)."""
    keyed = header[:2] == ["step", "cell"]
    start = 4 if keyed else 2
    ref_species = header[start:]
    col = {s: start + i for i, s in enumerate(ref_species)}
    sp_index = {s: i for i, s in enumerate(species)}
    want = list(cells_index)
    pos = {c: i for i, c in enumerate(want)}
    sq = {s: [] for s in ref_species}
    gate = True
    for step, Y in result.items():
        if not np.all(np.isfinite(Y)) or np.any(Y < -1e-20):
            gate = False
        if keyed:
            sel = rows[(rows[:, 0] == step) & np.isin(rows[:, 1], want)]
            pairs = [(Y[pos[int(r[1])]], r) for r in sel]
        else:
            pairs = [(Y[i], rows[c]) for i, c in enumerate(want)]
        for y, r in pairs:
            for s in ref_species:
                ref = r[col[s]]
                if ref < threshold or s not in sp_index:
                    continue
                sq[s].append(((y[sp_index[s]] - ref) / ref) ** 2)
    everything = [v for vs in sq.values() for v in vs]

    def digits(v):
        if not v:
            return float("nan")
        ms = float(np.mean(v))
        return 16.0 if ms <= 0.0 else min(16.0, -math.log10(math.sqrt(ms)))

    return digits(everything), {s: digits(v) for s, v in sq.items() if v}, gate


def run_job(spec):
    """One recorded experiment, for a process pool: spec = dict(scenario
    (npz path), mechanism (json path), solver ("algorithm" | "restarting" |
    "backward-euler"), options, tableau ("ros3" | "rodas3"), rtol, atol,
    constrained (species names), cells).  Returns the Counts, the digits
    and the statuses."""
    import os
    import time
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    t0 = time.time()
    ts1 = TS1(spec["mechanism"])
    scn = Scenario(spec["scenario"])
    cells = scn.cells(ts1, range(spec.get("cells", 64)))
    tab = O.ROS3 if spec.get("tableau", "ros3") == "ros3" else O.RODAS3
    kind = spec["solver"]
    if kind == "algorithm":
        solver = O.Algorithm(tab, **spec.get("options", {}))
    elif kind == "restarting":
        solver = O.RestartingLoop(tab)
    else:
        solver = O.BackwardEuler()
    fam = tuple(spec.get("constrained", ()))
    snaps, counts, statuses = run_scenario(scn, solver, cells, ts1, spec["rtol"], spec["atol"],
                                           constrained=fam)
    header, rows = scn.reference()
    D, per, gate = score(snaps, range(cells.C), ts1.species, header, rows)
    worst = min(per, key=per.get)
    fam_digits = min(per[s] for s in fam if s in per) if fam else float("nan")
    return {"counts": counts.as_dict(), "digits": D, "worst": worst, "worst_digits": per[worst],
            "family_digits": fam_digits, "gate": gate, "statuses": statuses,
            "seconds": time.time() - t0}
