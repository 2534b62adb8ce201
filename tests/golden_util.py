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


# Parity bar for states: relative 1e-9 above the floor a = 1e-17, with an
# absolute allowance of 1e-3 * atol (atol = rtol * a = 1e-20) for values at
# the floor.  Two pivoted factorizations of the reference itself (LAPACK
# getrf against SuperLU on the 200-species workload) differ by 9.6e-9
# relative (5.2e-25 absolute) on a value of 5.4e-17 after 48 intervals, so a
# purely relative bar cannot hold at the floor for any LU.
Y_RTOL = 1e-9
ATOL_FRACTION = 1e-3


def state_error(Y, ref, atol):
    """Worst |Y - ref| / (Y_RTOL |ref| + ATOL_FRACTION atol); <= 1 passes.
    Also returns the worst plain relative error above the floor."""
    tol = Y_RTOL * np.abs(ref) + ATOL_FRACTION * np.asarray(atol)[None, :]
    scaled = float((np.abs(Y - ref) / tol).max())
    big = np.abs(ref) > 1e-17
    rel = float((np.abs(Y - ref)[big] / np.abs(ref)[big]).max()) if big.any() else 0.0
    return scaled, rel
