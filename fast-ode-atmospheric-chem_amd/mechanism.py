"""Mass-action mechanisms compiled to the gather tables the HIP kernels read.

A reaction with rate constant k and reactants A_1 ... A_p proceeds at the rate
r = k x_1 ... x_p; it removes one of each reactant per event and adds the yield
of each product.  Third bodies (``M``) are not state variables and are folded
into k by the caller.  This follows the reference's ``Mechanism``
(fastode/chem.py:19-79): the reactant table ``react`` is (R, P) with the pad
value n pointing at a column of ones (chem.py:39-41), S is (n, R)
(chem.py:42-51), and the Jacobian pattern lists the diagonal first, then the
entries (i, k) in reaction order, reactant order, row order (chem.py:53-77).

The reference evaluates F = rates @ S^T as a dense GEMM (chem.py:89-90) and the
Jacobian values as a sparse product GT @ d (chem.py:78-79, 92-102).  Here both
are compiled to *gather* tables (CSR over the output), so that a device thread
that owns one output value sums its contributions without atomics:

* ``rhs_ptr/rhs_rxn/rhs_coef``: for species i, the reactions j with
  S[i, j] != 0 in increasing j, and S[i, j].
* ``jac_ptr/jac_src/jac_coef``: for Jacobian entry e = (i, k), the flattened
  (reaction, reactant slot) indices j*P + p with react[j, p] == k and
  S[i, j] != 0, in increasing j*P + p, and S[i, j].

Rows listed in ``constraints`` are algebraic linear conservation rows
g_i(Y) = sum_s c_s Y_s - b_i (as the conservation form of Robertson,
fastode/problems.py:33-49, replaces one species row by A + B + C - 1); their
RHS is g_i and their Jacobian row is the constant coefficient vector.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


class Mechanism:
    def __init__(self, species, reactions, third_bodies=("M",), constraints=None):
        """species: names; reactions: (reactants, products), reactants a list of
        names (repeats allowed), products a list of (name, yield);
        constraints: {row species name: ({species name: coefficient}, total)}."""
        self.species = list(species)
        n = self.n = len(self.species)
        index = {s: i for i, s in enumerate(self.species)}
        R = self.R = len(reactions)
        lists = []
        self.third_body_power = np.zeros(R)
        for j, (reactants, _) in enumerate(reactions):
            lst = []
            for s in reactants:
                if s in third_bodies:
                    self.third_body_power[j] += 1
                else:
                    lst.append(index[s])
            lists.append(lst)
        P = self.P = max(1, max((len(lst) for lst in lists), default=1))
        self.react = np.full((R, P), n, dtype=np.int64)
        for j, lst in enumerate(lists):
            self.react[j, : len(lst)] = lst
        S = np.zeros((n, R))
        for j, (_, products) in enumerate(reactions):
            for i in lists[j]:
                S[i, j] -= 1.0
            for s, y in products:
                if s in third_bodies:
                    continue
                S[index[s], j] += y
        self.S = S

        # Linear constraint rows: they replace the chemistry of their row.
        self.constraint_rows = np.zeros(0, dtype=np.int64)
        self.constraint_coef = np.zeros((0, n))
        self.constraint_total = np.zeros(0)
        if constraints:
            rows, coefs, totals = [], [], []
            for name, (coef, total) in constraints.items():
                rows.append(index[name])
                c = np.zeros(n)
                for s, v in coef.items():
                    c[index[s]] = v
                coefs.append(c)
                totals.append(float(total))
            order = np.argsort(rows)
            self.constraint_rows = np.asarray(rows, dtype=np.int64)[order]
            self.constraint_coef = np.asarray(coefs)[order]
            self.constraint_total = np.asarray(totals)[order]
        is_con = np.zeros(n, dtype=bool)
        is_con[self.constraint_rows] = True
        self.is_constraint = is_con

        # Jacobian pattern: diagonal first, then reaction order (chem.py:64-77);
        # constraint rows contribute their coefficient columns afterwards.
        entries: dict[tuple[int, int], int] = {}
        rows_l, cols_l = [], []

        def entry(i, k):
            key = (i, k)
            if key not in entries:
                entries[key] = len(rows_l)
                rows_l.append(i)
                cols_l.append(k)
            return entries[key]

        for i in range(n):
            entry(i, i)
        contrib: dict[int, list[tuple[int, float]]] = {}
        for j in range(R):
            nz = np.nonzero(S[:, j])[0]
            for p in range(P):
                kk = int(self.react[j, p])
                if kk == n:
                    continue
                for i in nz:
                    if is_con[i]:
                        continue
                    e = entry(int(i), kk)
                    contrib.setdefault(e, []).append((j * P + p, float(S[i, j])))
        con_entries = []
        for r, row in enumerate(self.constraint_rows):
            for s in np.nonzero(self.constraint_coef[r])[0]:
                e = entry(int(row), int(s))
                con_entries.append((e, float(self.constraint_coef[r, s])))
        self.rows = np.asarray(rows_l, dtype=np.int64)
        self.cols = np.asarray(cols_l, dtype=np.int64)
        self.nnz = len(rows_l)
        self.diag = np.arange(n, dtype=np.int64)

        # Gather tables.
        ptr = [0]
        rxn, cf = [], []
        for i in range(n):
            if not is_con[i]:
                for j in np.nonzero(S[i])[0]:
                    rxn.append(int(j))
                    cf.append(float(S[i, j]))
            ptr.append(len(rxn))
        self.rhs_ptr = np.asarray(ptr, dtype=np.int64)
        self.rhs_rxn = np.asarray(rxn, dtype=np.int64)
        self.rhs_coef = np.asarray(cf, dtype=np.float64)

        ptr = [0]
        src, cf = [], []
        for e in range(self.nnz):
            for jp, c in sorted(contrib.get(e, [])):
                src.append(jp)
                cf.append(c)
            ptr.append(len(src))
        self.jac_ptr = np.asarray(ptr, dtype=np.int64)
        self.jac_src = np.asarray(src, dtype=np.int64)
        self.jac_coef = np.asarray(cf, dtype=np.float64)
        # Constant Jacobian values (constraint rows), added after the gather.
        self.jac_const = np.zeros(self.nnz)
        for e, c in con_entries:
            self.jac_const[e] += c

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
            prod = []
            col = self.S[:, j].copy()
            for k in self.react[j]:
                if k < self.n:
                    col[k] += 1.0
            for i in np.nonzero(col)[0]:
                prod.append((self.species[i], float(col[i])))
            out.append((reac, prod))
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
