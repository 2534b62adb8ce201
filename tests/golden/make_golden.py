"""Write golden vectors for the seeded workloads by running the REFERENCE
implementation (davidfillmore/Fast-ODE-Atmospheric-Chem, imported from
/root/reference in the build container only).

Only the reference itself can pin Algorithm.solve's Counts, per-attempt
(H, err) sequence and final states on these seeded workloads: nothing in the
reference's own files holds them.  The spy follows the reference's own
rejection study (experiments/ts1_rejections.py:26-47): error_norm is wrapped
at module level (solve looks it up at call time), and system.factor records
alpha = 1/(H gamma) for each attempt.

Run:  python tests/golden/make_golden.py [name ...]
Writes tests/golden/<name>.npz (allow_pickle=False data only).
"""
from __future__ import annotations

import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]
sys.path.insert(0, "/root/reference")
sys.path.insert(0, str(REPO))

import fastode.solvers as ref_solvers            # noqa: E402  (the reference)
from fastode import Algorithm, Counts, ROS3, RODAS3  # noqa: E402
from fastode.chem import ChemSystem, Mechanism    # noqa: E402
from fastode.system import DenseSystem            # noqa: E402

import __graft_entry__                            # noqa: E402

fa = __graft_entry__.load_package()
from fastode_amd import workloads as W           # noqa: E402

COUNT_NAMES = ["steps", "accepted", "rejected", "rhs", "jac", "lu", "solves", "newton",
               "failures", "unconverged", "proj_passes", "proj_trials", "calls"]


class Spy:
    def __init__(self):
        self.errs = []
        self.alphas = []
        self._orig = ref_solvers.error_norm

    def __enter__(self):
        def wrapped(*a, **k):
            v = self._orig(*a, **k)
            self.errs.append(v)
            return v
        ref_solvers.error_norm = wrapped
        return self

    def __exit__(self, *exc):
        ref_solvers.error_norm = self._orig


def recording(system, spy):
    orig = system.factor

    def factor(alpha, J, algebraic_rows_only=False):
        if not algebraic_rows_only:
            spy.alphas.append(alpha)
        return orig(alpha, J, algebraic_rows_only)
    system.factor = factor
    return system


def make_system(wl, k):
    mech = wl.mech
    if wl.mech.has_constraints:
        # The reference has no constraint rows in its Mechanism: build the
        # constrained system as a DenseSystem (fastode/system.py:33-50) whose
        # rhs/jac replace the constraint rows, like the conservation form of
        # fastode/problems.py:33-49.
        rmech = Mechanism(mech.species, mech.reactions_list())
        rows = mech.constraint_rows
        coef = mech.constraint_coef
        tot = mech.constraint_total

        def rhs(Y):
            F = rmech.rhs(k, Y)
            F[:, rows] = Y @ coef.T - tot[None, :]
            return F

        def jac(Y):
            J = rmech.dense_jacobian(rmech.jac_values(k, Y))
            J[:, rows, :] = coef[None, :, :]
            return J
        return DenseSystem(mech.n, rhs, jac, mass=wl.mass)
    rmech = Mechanism(mech.species, mech.reactions_list())
    if wl.dense:
        return DenseSystem(mech.n, lambda Y: rmech.rhs(k, Y),
                           lambda Y: rmech.dense_jacobian(rmech.jac_values(k, Y)))
    return ChemSystem(rmech, k, dense=False)


def run(wl, groups, n_intervals, snap_every):
    tab = {"ros3": ROS3, "rodas3": RODAS3}[wl.tableau]
    alg = Algorithm(tab)
    G = 32
    out = dict(counts=[], status=[], att_group=[], att_interval=[], att_alpha=[], att_err=[],
               snap_interval=[], snaps=[])
    cells = np.concatenate([np.arange(g * G, (g + 1) * G) for g in groups])
    Yall = wl.Y0[cells].copy()
    mems = [alg.new_memory() for _ in groups]
    for it in range(n_intervals):
        for gi, g in enumerate(groups):
            sl = slice(gi * G, (gi + 1) * G)
            k = wl.k(it, cells[sl])
            system = make_system(wl, k)
            spy = Spy()
            recording(system, spy)
            c = Counts()
            with spy:
                Y, st = alg.solve(system, Yall[sl], wl.dt, wl.atol, wl.rtol, c, mems[gi])
            Yall[sl] = Y
            assert len(spy.errs) == len(spy.alphas), (len(spy.errs), len(spy.alphas))
            out["counts"].append([getattr(c, f) for f in COUNT_NAMES])
            out["status"].append(st)
            out["att_group"] += [gi] * len(spy.errs)
            out["att_interval"] += [it] * len(spy.errs)
            out["att_alpha"] += spy.alphas
            out["att_err"] += spy.errs
        if (it + 1) % snap_every == 0 or it + 1 == n_intervals:
            out["snap_interval"].append(it + 1)
            out["snaps"].append(Yall.copy())
    mem = np.array([[m.h_next, m.err_prev, m.h_prev] for m in mems])
    return cells, out, mem


SPECS = {
    # name: (workload factory, groups, intervals, snapshot every)
    "mech20_dense": (lambda: W.small(), [0, 1], 10, 1),
    "mech200_diurnal": (lambda: W.diurnal200(C=32 * 64), [0, 37], 48, 12),
    "mech200_uneven": (lambda: W.uneven200(cells=np.r_[0:64, 32 * 777:32 * 778]), [0, 2], 24, 6),
    "mech100_constrained": (lambda: W.constrained100(C=64), [0, 1], 24, 6),
}


def main(names):
    for name in names:
        make, groups, nint, every = SPECS[name]
        wl = make()
        t0 = time.time()
        cells, out, mem = run(wl, groups, nint, every)
        dt = time.time() - t0
        status = np.array(out["status"])
        np.savez_compressed(
            HERE / f"{name}.npz",
            cells=cells, groups=np.array(groups), n_intervals=np.array(nint),
            Y0=wl.Y0[cells], k_base=wl.k_base,
            react=wl.mech.react, S=wl.mech.S,
            counts=np.array(out["counts"], dtype=np.int64).reshape(nint, len(groups), 13),
            status=status.reshape(nint, len(groups)),
            att_group=np.array(out["att_group"], dtype=np.int64),
            att_interval=np.array(out["att_interval"], dtype=np.int64),
            att_alpha=np.array(out["att_alpha"]), att_err=np.array(out["att_err"]),
            snap_interval=np.array(out["snap_interval"]), snaps=np.array(out["snaps"]),
            memory=mem, seconds=np.array(dt))
        c = np.array(out["counts"]).sum(axis=0)
        print(f"{name}: {dt:.1f} s, statuses {set(out['status'])}, counts {dict(zip(COUNT_NAMES, c))}")


if __name__ == "__main__":
    main(sys.argv[1:] or list(SPECS))
