"""Write the tangent-linear fixture of the adjoint workload with the ORACLE
(oracle/fastode_oracle.py, the NumPy/SciPy restatement of the reference's
path; the reference itself is not imported).

Workload: the 200-species diurnal mechanism, its first 64 cells (2 groups of
32), 24 intervals of 600 s of ROS3, loss sum_c sum_s w_s log(Y_final + 1e-17).
For each direction (dlogk, dY0) below, the oracle's tangent-linear run gives
the directional derivative of the loss; the fixture keeps those, the loss,
the final state and the accepted step sizes per group and interval.

Run:  python tests/golden/make_adjoint_golden.py      (about 5 minutes, 4 processes)
Writes tests/golden/adjoint200_tlm.npz (allow_pickle=False data only).
"""
from __future__ import annotations

import sys
import time
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]
sys.path.insert(0, str(REPO))
sys.path.insert(0, str(REPO / "oracle"))

CELLS = 64
INTERVALS = 24
FLOOR = 1e-17


def directions(C, R, n, Y0):
    """The fixture's directions (dlogk (C, R), dY0 (C, n)), regenerated the
    same way by the tests."""
    rng = np.random.default_rng(777)
    d = []
    d.append((rng.normal(size=(C, R)), np.zeros((C, n))))               # log k, all cells
    d.append((np.zeros((C, R)), Y0 * rng.normal(size=(C, n))))          # relative initial state
    unit = np.zeros((C, R))
    unit[5, 17] = 1.0
    d.append((unit, np.zeros((C, n))))                                  # one rate constant of one cell
    d.append((rng.normal(size=(C, R)), 0.3 * Y0 * rng.normal(size=(C, n))))
    return d


def setup():
    import __graft_entry__
    __graft_entry__.load_package()
    import fastode_oracle as O
    from fastode_amd import workloads as W
    wl = W.adjoint200(C=CELLS, n_intervals=INTERVALS)
    om = O.OracleMechanism(wl.mech.species, wl.mech.reactions_list())
    k = np.broadcast_to(wl.k_base, (CELLS, wl.R)).copy()
    return O, wl, om, k


def run(i):
    O, wl, om, k = setup()
    dlogk, dY0 = directions(CELLS, wl.R, wl.n, wl.Y0)[i]
    t0 = time.time()
    Y, dY, steps = O.tangent_linear(om, k, wl.kmuls(), O.ROS3, wl.Y0, dY0, dlogk, wl.dt, wl.atol, wl.rtol)
    dl = float((wl.weights[None, :] * dY / (Y + FLOOR)).sum())
    print(f"direction {i}: dL = {dl:.12e} ({time.time() - t0:.0f} s)", flush=True)
    return dl, Y, steps


def main():
    with ProcessPoolExecutor(4) as ex:
        res = list(ex.map(run, range(4)))
    O, wl, om, k = setup()
    Y, steps = res[0][1], res[0][2]
    for _, Yi, si in res[1:]:
        assert np.array_equal(Yi, Y) and si == steps
    loss = float((wl.weights[None, :] * np.log(Y + FLOOR)).sum())
    nacc = np.array([[len(h) for h in row] for row in steps], dtype=np.int32)      # (I, groups)
    hs = np.concatenate([np.asarray(h, dtype=float) for row in steps for h in row])
    np.savez_compressed(HERE / "adjoint200_tlm.npz", dL=np.array([r[0] for r in res]), loss=np.array(loss),
                        Y=Y, nacc=nacc, H=hs, weights=wl.weights)
    print("loss", loss, "accepted steps per group-interval", nacc.mean())


if __name__ == "__main__":
    main()
