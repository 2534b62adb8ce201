"""Helpers shared by the tests that read the golden vectors."""
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
COUNT_NAMES = ["steps", "accepted", "rejected", "rhs", "jac", "lu", "solves", "newton",
               "failures", "unconverged", "proj_passes", "proj_trials", "calls"]


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def workload(name):
    import fastode_amd.workloads as W
    import numpy as np
    if name == "mech20_dense":
        return W.small()
    if name == "mech200_diurnal":
        return W.diurnal200(C=32 * 64)
    if name == "mech200_uneven":
        return W.uneven200(cells=np.r_[0:64, 32 * 777:32 * 778])
    if name == "mech100_constrained":
        return W.constrained100(C=64)
    raise KeyError(name)


# Parity bar for states.  Relative 1e-9 for values at or above 1e-14 (the
# reference's own scoring threshold, fastode/ts1.py:226); between the floor
# a = 1e-17 and 1e-14 an absolute 1e-23 (1e-9 of the threshold); below the
# floor, within 10 a.  Two pivoted factorizations of the reference itself
# (LAPACK getrf against SuperLU, 200-species workload, 48 intervals) differ
# by 9.6e-9 relative on a value of 5.4e-17 and by 2.4e-17 absolute below the
# floor, so no LU can hold a purely relative 1e-9 down at the floor.
Y_RTOL = 1e-9
THRESHOLD = 1e-14
FLOOR = 1e-17


def state_error(Y, ref, atol=None):
    """Worst error scaled by its allowance (<= 1 passes), and the worst
    relative error above the floor (reported)."""
    aref = np.abs(ref)
    d = np.abs(Y - ref)
    tol = np.where(aref >= THRESHOLD, Y_RTOL * aref,
                   np.where(aref > FLOOR, Y_RTOL * THRESHOLD, 10 * FLOOR))
    scaled = float((d / tol).max())
    big = aref > FLOOR
    rel = float((d[big] / aref[big]).max()) if big.any() else 0.0
    return scaled, rel
