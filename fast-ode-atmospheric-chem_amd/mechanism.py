"""Mass-action mechanisms compiled to the gather tables the HIP kernels read.

Each reaction j fires at r_j = k_j * (product of its reactant concentrations)
and changes species i by S[i, j] per event (reactants -1 each, products +
their yield).  Third-body species (``M``) are not state variables: their
count per reaction is recorded in ``third_body_power`` and the caller folds
[M]^power into k.  The arrays are those the reference's ``Mechanism`` exposes
(fastode/chem.py:19-79): ``react`` (R, P) with the pad value n standing for a
factor 1, ``S`` (n, R), and the Jacobian pattern ``rows``/``cols`` with the
diagonal first, then the entries (i, k) met in the order reaction, reactant
slot, row.

The device never forms the dense products the reference evaluates (F =
rates @ S^T, J = GT @ d, chem.py:86-102).  It reads *gather* tables, CSR over
the outputs, so that the thread owning one output sums its terms without
atomics:

* ``rhs_ptr/rhs_rxn/rhs_coef``: for species i, the reactions j with
  S[i, j] != 0 in increasing j, and S[i, j];
* ``jac_ptr/jac_src/jac_coef``: for pattern entry e = (i, k), the flattened
  slots j*P + p with react[j, p] == k and S[i, j] != 0, increasing, and
  S[i, j].

Linear constraint rows (``constraints``) replace the chemistry of a species'
row by g_i(Y) = sum_s c_s Y_s - b_i, the conservation form the reference
writes by hand for Robertson (fastode/problems.py:33-49); their Jacobian row
is the constant coefficient vector (``jac_const``).

``Mechanism(species, reactions, ...)`` compiles reaction lists;
``Mechanism.adopt(m)`` compiles any object carrying the reference
Mechanism's arrays (``species``, ``react``, ``S``, ``pattern.rows/cols``,
optionally ``third_body_power``), keeping its pattern order, so that a
mechanism the reference built drives the device systems unchanged.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def _csr(keys, n_keys):
    """Row pointer of sorted integer keys over n_keys rows."""
    return np.concatenate([[0], np.cumsum(np.bincount(keys, minlength=n_keys))]).astype(np.int64)


class Mechanism:
    def __init__(self, species, reactions, third_bodies=("M",), constraints=None):
        """species: names of the state species; reactions: (reactants,
        products) pairs, reactants a list of names (repeats allowed),
        products a list of (name, yield); constraints: {row species name:
        ({species name: coefficient}, total)}."""
        species = list(species)
        index = {name: i for i, name in enumerate(species)}
        third = frozenset(third_bodies)
        n, R = len(species), len(reactions)
        slots = [[index[x] for x in reac if x not in third] for reac, _ in reactions]
        power = np.array([float(sum(x in third for x in reac)) for reac, _ in reactions])
        P = max([1] + [len(x) for x in slots])
        react = np.full((R, P), n, dtype=np.int64)
        S = np.zeros((n, R))
        for j, ((_, products), x) in enumerate(zip(reactions, slots)):
            react[j, :len(x)] = x
            for i in x:
                S[i, j] -= 1.0
            for name, y in products:
                if name not in third:
                    S[index[name], j] += y
        self._compile(species, react, S, power, None, constraints)

    @classmethod
    def adopt(cls, m, constraints=None):
        """A device mechanism from any object exposing the reference
        Mechanism's arrays: ``species``, ``react`` (R, P; pad n), ``S`` (n, R),
        ``pattern.rows``/``pattern.cols`` (whose order the Jacobian values
        keep) and, if present, ``third_body_power``."""
        if isinstance(m, cls):
            return m
        self = cls.__new__(cls)
        species = list(m.species)
        react = np.asarray(m.react, dtype=np.int64)
        S = np.asarray(m.S, dtype=np.float64)
        power = np.asarray(getattr(m, "third_body_power", np.zeros(react.shape[0])), dtype=np.float64)
        pattern = None
        if not constraints:
            pattern = (np.asarray(m.pattern.rows, dtype=np.int64), np.asarray(m.pattern.cols, dtype=np.int64))
        self._compile(species, react, S, power, pattern, constraints)
        return self

    # ------------------------------------------------------------------
    def _compile(self, species, react, S, power, pattern, constraints):
        n = len(species)
        R, P = react.shape
        if S.shape != (n, R):
            raise ValueError(f"S has shape {S.shape}, expected {(n, R)}")
        if react.min(initial=0) < 0 or react.max(initial=n) > n:
            raise ValueError("react holds an index outside 0..n")
        self.species, self.n, self.R, self.P = species, n, R, P
        self.react, self.S, self.third_body_power = react, S, power
        index = {name: i for i, name in enumerate(species)}

        rows_c, coefs, totals = [], [], []
        for name, (coef, total) in (constraints or {}).items():
            c = np.zeros(n)
            for sp_name, v in coef.items():
                c[index[sp_name]] = v
            rows_c.append(index[name])
            coefs.append(c)
            totals.append(float(total))
        order = np.argsort(rows_c, kind="stable")
        self.constraint_rows = np.asarray(rows_c, dtype=np.int64)[order]
        self.constraint_coef = np.asarray(coefs, dtype=np.float64).reshape(-1, n)[order]
        self.constraint_total = np.asarray(totals, dtype=np.float64)[order]
        is_con = np.zeros(n, dtype=bool)
        is_con[self.constraint_rows] = True
        self.is_constraint = is_con

        # chemistry terms of the Jacobian: (row i, column k, slot j*P + p, S[i, j])
        # in the order reaction j, slot p, row i (the pattern's order)
        jj, ii = np.nonzero(S.T)
        t_row, t_col, t_slot = [], [], []
        by_j = np.split(ii, np.flatnonzero(np.diff(jj)) + 1) if len(jj) else []
        js = np.unique(jj)
        for j, rows_j in zip(js, by_j):
            rows_j = rows_j[~is_con[rows_j]]
            for p in range(P):
                k = int(react[j, p])
                if k == n or not len(rows_j):
                    continue
                t_row.append(rows_j)
                t_col.append(np.full(len(rows_j), k))
                t_slot.append(np.full(len(rows_j), j * P + p))
        t_row = np.concatenate(t_row) if t_row else np.zeros(0, dtype=np.int64)
        t_col = np.concatenate(t_col) if t_col else np.zeros(0, dtype=np.int64)
        t_slot = np.concatenate(t_slot) if t_slot else np.zeros(0, dtype=np.int64)
        t_coef = S[t_row, t_slot // P]

        con_r, con_c = np.nonzero(self.constraint_coef)
        con_row = self.constraint_rows[con_r]
        if pattern is None:
            # first appearance of each (i, k): diagonal, chemistry terms, constraints
            cand_r = np.concatenate([np.arange(n), t_row, con_row])
            cand_c = np.concatenate([np.arange(n), t_col, con_c])
            key = cand_r * n + cand_c
            _, first = np.unique(key, return_index=True)
            first.sort()
            rows, cols = cand_r[first], cand_c[first]
        else:
            rows, cols = pattern
        self.rows, self.cols = rows.astype(np.int64), cols.astype(np.int64)
        self.nnz = len(self.rows)
        self.diag = np.arange(n, dtype=np.int64)
        entry = np.full(n * n, -1, dtype=np.int64)
        entry[self.rows * n + self.cols] = np.arange(self.nnz)
        if np.any(entry[np.arange(n) * (n + 1)] < 0):
            raise ValueError("the pattern lacks a diagonal entry")

        # species gather: reactions of each row in increasing j
        si, sj = np.nonzero(S)
        keep = ~is_con[si]
        self.rhs_ptr = _csr(si[keep], n)
        self.rhs_rxn = sj[keep].astype(np.int64)
        self.rhs_coef = S[si[keep], sj[keep]]

        # Jacobian gather: terms of each entry in increasing slot
        t_entry = entry[t_row * n + t_col]
        if np.any(t_entry < 0):
            raise ValueError("the pattern lacks entries the mechanism's Jacobian has")
        o = np.lexsort((t_slot, t_entry))
        self.jac_ptr = _csr(t_entry[o], self.nnz)
        self.jac_src = t_slot[o].astype(np.int64)
        self.jac_coef = t_coef[o]
        self.jac_const = np.zeros(self.nnz)
        c_entry = entry[con_row * n + con_c]
        if np.any(c_entry < 0):
            raise ValueError("the pattern lacks constraint entries")
        np.add.at(self.jac_const, c_entry, self.constraint_coef[con_r, con_c])

    @property
    def has_constraints(self) -> bool:
        return len(self.constraint_rows) > 0

    def reactions_list(self):
        """The (reactants, products) list this mechanism was compiled from, with
        species names, recovered from react and S (used to hand the same
        mechanism to other implementations)."""
        out = []
        for j in range(self.R):
            reac = [self.species[k] for k in self.react[j] if k < self.n]
            col = self.S[:, j].copy()
            for k in self.react[j]:
                if k < self.n:
                    col[k] += 1.0
            out.append((reac, [(self.species[i], float(col[i])) for i in np.nonzero(col)[0]]))
        return out


# ---------------------------------------------------------------------------
# Seeded synthetic mechanisms and workloads.
# ---------------------------------------------------------------------------

R_GAS = 1.380649e-23 * 6.02214076e23      # J mol-1 K-1, as fastode/ts1.py:28-30
CONSTRAINT_TOTAL = 1.0e-2                 # mol m-3, every synthetic conservation row


@dataclass
class Synthetic:
    """A seeded synthetic mass-action mechanism.

    Reaction j is first-order photolysis (one reactant, k log-uniform in
    [1e-6, 1e-2] s-1) for the round(0.2 R) reactions picked by a permutation,
    and second-order otherwise (two distinct reactants, k log-uniform in
    [1e-3, 1e3] m3 mol-1 s-1).  Every reaction has 1, 2 or 3 products
    (count uniform) drawn uniformly without replacement from all species,
    each with yield 1.  Reactant species are drawn uniformly."""
    mech: Mechanism
    k: np.ndarray                 # (R,) base rate constants
    photolysis: np.ndarray        # (R,) bool
    seed: int
    ea: np.ndarray = field(default=None)   # (R,) J mol-1, thermal reactions only


def synthetic_reactions(n, R, rng):
    n_phot = int(round(0.2 * R))
    phot = np.zeros(R, dtype=bool)
    phot[rng.permutation(R)[:n_phot]] = True
    names = [f"S{i:03d}" for i in range(n)]
    reactions = []
    k = np.empty(R)
    for j in range(R):
        if phot[j]:
            reac = [int(rng.integers(n))]
            k[j] = 10.0 ** rng.uniform(-6.0, -2.0)
        else:
            reac = [int(x) for x in rng.choice(n, size=2, replace=False)]
            k[j] = 10.0 ** rng.uniform(-3.0, 3.0)
        npro = int(rng.integers(1, 4))
        prod = [int(x) for x in rng.choice(n, size=npro, replace=False)]
        reactions.append(([names[i] for i in reac], [(names[i], 1.0) for i in prod]))
    return names, reactions, k, phot


def synthetic(n, R, seed, constraints=0):
    """The seeded mechanism; with constraints > 0, that many linear
    conservation rows over disjoint random species subsets of 3 to 6 species
    with 0/1 coefficients, each replacing the row of the first species of its
    subset, each with the total CONSTRAINT_TOTAL."""
    rng = np.random.default_rng(seed)
    names, reactions, k, phot = synthetic_reactions(n, R, rng)
    cons = None
    if constraints:
        perm = rng.permutation(n)
        cons = {}
        at = 0
        for _ in range(constraints):
            size = int(rng.integers(3, 7))
            subset = sorted(int(x) for x in perm[at: at + size])
            at += size
            cons[names[subset[0]]] = ({names[s]: 1.0 for s in subset}, CONSTRAINT_TOTAL)
    ea = rng.uniform(0.0, 20.0e3, size=R)
    ea[phot] = 0.0
    return Synthetic(Mechanism(names, reactions, constraints=cons), k, phot, seed, ea)


def log_uniform_state(C, n, rng, lo=1e-15, hi=1e-3):
    return 10.0 ** rng.uniform(np.log10(lo), np.log10(hi), size=(C, n))


def diurnal_factor(t):
    """Photolysis scaling max(0, cos(2 pi t / 86400)) at a host time t (s)."""
    return max(0.0, float(np.cos(2.0 * np.pi * t / 86400.0)))
