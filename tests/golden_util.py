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
