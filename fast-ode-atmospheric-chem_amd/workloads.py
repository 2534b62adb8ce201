"""The seeded synthetic workloads the integrator is measured and checked on.

Each workload fixes a mechanism, per-cell rate constants per host interval, an
initial state, tolerances and the interval length.  Host intervals follow the
reference's scenario runner (fastode/ts1.py:181-223): rate constants are set
at the *start* of each interval (photolysis times the diurnal factor there),
every group of 32 cells keeps its own controller memory across intervals.

Choices the workload text leaves open, fixed here:
* second-order reactions have two distinct reactants; products are distinct;
  every yield is 1 (see ``mechanism.synthetic``);
* atol = rtol * 1e-17 for every species;
* the 262,144-cell workload scales *thermal* reactions by exp(-Ea_j/(R T_c))
  with Ea_j ~ U(0, 20) kJ/mol per reaction and T_c ~ U(200, 310) K per cell,
  and *all* reactions of a cell by one multiplier 10^U(-3, 1) per cell (four
  decades; centred at 1 the stiffest cells of almost every group exhaust the
  1000-attempt budget every interval); its photolysis also follows the
  diurnal factor;
* the constrained workload holds 3 linear conservation rows with total
  1e-2 mol m-3 each, and every cell starts on them.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .mechanism import R_GAS, Mechanism, diurnal_factor, log_uniform_state, synthetic

FLOOR = 1.0e-17
GROUP = 32


@dataclass
class Workload:
    name: str
    mech: Mechanism
    C: int                      # cells
    n_intervals: int
    dt: float
    rtol: float
    atol: np.ndarray            # (n,)
    tableau: str                # "ros3" | "rodas3"
    seed: int
    dense: bool = False         # solved as a DenseSystem by the reference
    k_base: np.ndarray = None   # (R,)
    photolysis: np.ndarray = None
    cell_scale: np.ndarray = None   # (C, R) multiplicative factor, or None
    diurnal: bool = False
    mass: np.ndarray = None
    Y0: np.ndarray = None       # (C, n)
    weights: np.ndarray = None  # (n,) loss weights of the adjoint workload

    @property
    def n(self):
        return self.mech.n

    @property
    def R(self):
        return self.mech.R

    def kmul(self, interval):
        """Per-reaction multiplier (R,) of an interval: the diurnal factor at
        its start on photolysis reactions, 1 elsewhere."""
        m = np.ones(self.R)
        if self.diurnal:
            m[self.photolysis] = diurnal_factor(interval * self.dt)
        return m

    def kmuls(self):
        return np.stack([self.kmul(i) for i in range(self.n_intervals)])

    def k(self, interval, cells=slice(None)):
        """Rate constants (C_sel, R) for the given interval."""
        k = np.broadcast_to(self.k_base, (self.C, self.R))[cells].copy()
        if self.cell_scale is not None:
            k *= self.cell_scale[cells]
        if self.diurnal:
            f = diurnal_factor(interval * self.dt)
            k[:, self.photolysis] *= f
        return k


def _base(n, R, seed):
    return synthetic(n, R, seed)


def small(C=64, seed=1):
    """20 species, 40 reactions, 64 cells = 2 groups, 10 intervals of 30 s."""
    syn = _base(20, 40, seed)
    rng = np.random.default_rng(seed + 1000)
    rtol = 1.0e-3
    return Workload("mech20-dense", syn.mech, C, 10, 30.0, rtol, np.full(20, rtol * FLOOR), "ros3",
                    seed, dense=True, k_base=syn.k, photolysis=syn.photolysis,
                    mass=np.ones(20), Y0=log_uniform_state(C, 20, rng))


def diurnal200(C=65536, seed=2, n_intervals=48):
    """200 species, 500 reactions, diurnal photolysis, 48 intervals of 600 s."""
    syn = _base(200, 500, seed)
    rng = np.random.default_rng(seed + 1000)
    rtol = 1.0e-3
    return Workload("mech200-diurnal", syn.mech, C, n_intervals, 600.0, rtol, np.full(200, rtol * FLOOR),
                    "ros3", seed, k_base=syn.k, photolysis=syn.photolysis, diurnal=True,
                    mass=np.ones(200), Y0=log_uniform_state(C, 200, rng))


def uneven200(C=262144, seed=2, n_intervals=24, cells=None):
    """The same mechanism with per-cell temperature and rate multipliers.
    ``cells``: build only these cell indices (the draws are those of the full
    262,144-cell workload)."""
    syn = _base(200, 500, seed)
    rng = np.random.default_rng(seed + 3000)
    T = rng.uniform(200.0, 310.0, size=C)
    mult = 10.0 ** rng.uniform(-3.0, 1.0, size=C)
    rng_y = np.random.default_rng(seed + 2000)
    Y0 = log_uniform_state(C, 200, rng_y)
    sel = np.arange(C) if cells is None else np.asarray(cells)
    scale = np.exp(-syn.ea[None, :] / (R_GAS * T[sel, None])) * mult[sel, None]
    rtol = 1.0e-3
    return Workload("mech200-uneven", syn.mech, len(sel), n_intervals, 600.0, rtol,
                    np.full(200, rtol * FLOOR), "ros3", seed, k_base=syn.k,
                    photolysis=syn.photolysis, diurnal=True, cell_scale=scale,
                    mass=np.ones(200), Y0=Y0[sel])


def constrained100(C=65536, seed=4, n_intervals=24):
    """100 species, 250 reactions + 3 conservation rows, RODAS3, 24 x 300 s.
    Each algebraic species starts at its row's total minus the rest of its
    subset, so the initial state satisfies the constraints."""
    syn = synthetic(100, 250, seed, constraints=3)
    mech = syn.mech
    rng = np.random.default_rng(seed + 1000)
    Y0 = log_uniform_state(C, 100, rng)
    for r, row in enumerate(mech.constraint_rows):
        others = mech.constraint_coef[r].copy()
        others[row] = 0.0
        Y0[:, row] = mech.constraint_total[r] - Y0 @ others
    mass = np.ones(100)
    mass[mech.constraint_rows] = 0.0
    rtol = 1.0e-3
    return Workload("mech100-constrained", mech, C, n_intervals, 300.0, rtol, np.full(100, rtol * FLOOR),
                    "rodas3", seed, k_base=syn.k, photolysis=syn.photolysis, mass=mass, Y0=Y0)


def adjoint200(C=16384, seed=2, n_intervals=24):
    """The diurnal 200-species workload on 16,384 cells over 24 intervals, with
    the loss sum_c sum_s w_s log(Y_final[c, s] + 1e-17), w_s ~ U(0, 1)."""
    w = diurnal200(C=C, seed=seed, n_intervals=n_intervals)
    w.name = "mech200-adjoint"
    w.weights = np.random.default_rng(seed + 5000).uniform(0.0, 1.0, size=w.n)
    return w
