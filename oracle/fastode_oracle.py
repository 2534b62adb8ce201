"""ORACLE — test infrastructure only.

A CPU restatement (NumPy + SciPy) of the reference's hot path,
``Algorithm(ROS3|RODAS3).solve`` with its ``System`` protocol.  It is the
checker for the HIP path: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and the product never
routes through it.

Each function cites the reference file:line whose behaviour it restates.
How each part is pinned (golden vectors written by the reference itself,
closed forms, invariants) is listed in ``PINS`` at the bottom; a part pinned
by nothing would say "parity unpinned" there.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, fields

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

EPS = float(np.finfo(float).eps)      # fastode/solvers.py:31
H_MIN_FRACTION = 1.0e-15              # fastode/solvers.py:32
H_START_FRACTION = 1.0e-6             # fastode/solvers.py:33
ERROR_MIN = 1.0e-10                   # fastode/solvers.py:34


# ----------------------------------------------------------------- tableaus
@dataclass(frozen=True)
class Tableau:
    """KPP-form Rosenbrock coefficients (fastode/tableaus.py:14-42)."""
    name: str
    gamma: float
    a: np.ndarray
    c: np.ndarray
    m: np.ndarray
    e: np.ndarray
    order: int
    embedded: int

    @property
    def stages(self):
        return len(self.m)

    @property
    def k(self):
        return float(self.embedded + 1)                       # tableaus.py:30-33

    @property
    def new_forcing(self):
        # stage r needs F anew unless its point equals stage r-1's (tableaus.py:35-42)
        out = [True]
        for r in range(1, self.stages):
            same = np.array_equal(self.a[r, :r], np.append(self.a[r - 1, : r - 1], 0.0))
            out.append(not (same and self.a[r, r - 1] == 0.0))
        return out

    def standard_form(self):
        """(A, Gamma, b, bhat) of the standard form (tableaus.py:44-51)."""
        S = self.stages
        Gam = np.linalg.inv(np.eye(S) / self.gamma - self.c)
        return self.a @ Gam, Gam, Gam.T @ self.m, Gam.T @ self.m - Gam.T @ self.e

    def index1(self, q):
        A, Gam, b, _ = self.standard_form()
        return float(b @ np.linalg.solve(A + Gam, A.sum(axis=1) ** q))   # tableaus.py:53-57


def _strict_lower(S, values):
    M = np.zeros((S, S))
    M[np.tril_indices(S, -1)] = values       # row-major over (i > j), tableaus.py:60-67
    return M


ROS3 = Tableau("ros3", 0.43586652150845899941601945119356,           # tableaus.py:70-80
               _strict_lower(3, [1.0, 1.0, 0.0]),
               _strict_lower(3, [-0.10156171083877702091975600115545e01,
                                 0.40759956452537699824805835358067e01,
                                 0.92076794298330791242156818474003e01]),
               np.array([1.0, 0.61697947043828245592553615689730e01,
                         -0.42772256543218573326238373806514]),
               np.array([0.5, -0.29079558716805469821718236208017e01,
                         0.22354069897811569627360909276199]), 3, 2)
RODAS3 = Tableau("rodas3", 0.5,                                        # tableaus.py:82-90
                 _strict_lower(4, [0.0, 2.0, 0.0, 2.0, 0.0, 1.0]),
                 _strict_lower(4, [4.0, 1.0, -1.0, 1.0, -1.0, -8.0 / 3.0]),
                 np.array([2.0, 0.0, 1.0, 1.0]), np.array([0.0, 0.0, 0.0, 1.0]), 3, 2)


# ------------------------------------------------------------------ counts
@dataclass
class Counts:
    """The 13 per-group operation counters (fastode/counts.py:9-30)."""
    steps: int = 0
    accepted: int = 0
    rejected: int = 0
    rhs: int = 0
    jac: int = 0
    lu: int = 0
    solves: int = 0
    newton: int = 0
    failures: int = 0
    unconverged: int = 0
    proj_passes: int = 0
    proj_trials: int = 0
    calls: int = 0

    def add(self, other):
        for f in fields(self):
            setattr(self, f.name, getattr(self, f.name) + getattr(other, f.name))

    def as_dict(self):
        return {f.name: getattr(self, f.name) for f in fields(self)}


@dataclass
class Memory:
    """Carried controller state (fastode/solvers.py:181-186)."""
    h_next: float = 0.0
    err_prev: float = 0.0
    h_prev: float = 0.0


# ---------------------------------------------------------------- systems
class LapackLU:
    """Per-cell LAPACK getrf/getrs (fastode/linalg.py:15-21)."""

    def __init__(self, A):
        self.f = [sla.lu_factor(A[c], check_finite=False) for c in range(A.shape[0])]

    def solve(self, B):
        return np.stack([sla.lu_solve(f, B[c], check_finite=False) for c, f in enumerate(self.f)])


def stage_matrix_dense(alpha, mass, J, algebraic_rows_only=False):
    """alpha*diag(mass) - J; the projection keeps only algebraic rows of J
    (fastode/linalg.py:24-32)."""
    if algebraic_rows_only:
        J = J * (mass == 0.0)[None, :, None]
    A = -J
    d = np.arange(J.shape[1])
    A[:, d, d] += alpha * mass
    return A


class DenseSystem:
    """rhs/jac from callables, LAPACK factorization (fastode/system.py:33-50)."""

    def __init__(self, n, rhs, jac, mass=None):
        self.n = n
        self._rhs, self._jac = rhs, jac
        self.mass = np.ones(n) if mass is None else np.asarray(mass, dtype=float)

    @property
    def has_constraints(self):
        return bool(np.any(self.mass == 0.0))               # system.py:27-30

    def rhs(self, Y):
        return self._rhs(Y)

    def jac(self, Y):
        return self._jac(Y)

    def factor(self, alpha, J, algebraic_rows_only=False):
        return LapackLU(stage_matrix_dense(alpha, self.mass, J, algebraic_rows_only))


class OracleMechanism:
    """Mass-action rates, forcing and Jacobian values (fastode/chem.py:19-108),
    restated: react (R, P) with pad n, dense S^T for the forcing GEMM, the
    pattern with the diagonal first, and GT mapping d(rate)/dy to J values.
    Optional linear constraint rows replace a species' row by
    sum_s c_s y_s - total (the conservation form of fastode/problems.py:33-49)."""

    def __init__(self, species, reactions, constraints=None, third_bodies=("M",)):
        self.species = list(species)
        n = self.n = len(species)
        idx = {s: i for i, s in enumerate(species)}
        R = self.R = len(reactions)
        third = set(third_bodies)
        # third bodies are not state variables; their count per reaction is
        # the power of [M] folded into k by the caller (chem.py:28-36)
        lists = [[idx[s] for s in reac if s not in third] for reac, _ in reactions]
        self.third_body_power = np.array([float(sum(s in third for s in reac)) for reac, _ in reactions])
        P = self.P = max(1, max((len(x) for x in lists), default=1))
        self.react = np.full((R, P), n, dtype=np.int64)
        for j, x in enumerate(lists):
            self.react[j, : len(x)] = x
        S = np.zeros((n, R))
        for j, (_, prods) in enumerate(reactions):
            for i in lists[j]:
                S[i, j] -= 1.0
            for s, y in prods:
                if s not in third:
                    S[idx[s], j] += y
        self.con_rows, self.con_coef, self.con_total = [], [], []
        if constraints:
            for name, (coef, total) in sorted(constraints.items(), key=lambda kv: idx[kv[0]]):
                c = np.zeros(n)
                for s, v in coef.items():
                    c[idx[s]] = v
                self.con_rows.append(idx[name])
                self.con_coef.append(c)
                self.con_total.append(float(total))
        self.con_rows = np.asarray(self.con_rows, dtype=np.int64)
        self.con_coef = np.asarray(self.con_coef).reshape(-1, n)
        self.con_total = np.asarray(self.con_total)
        Sm = S.copy()
        Sm[self.con_rows, :] = 0.0                           # constraint rows carry no chemistry
        self.ST = np.ascontiguousarray(Sm.T)
        pos = {}
        rows, cols = [], []

        def at(i, k):
            if (i, k) not in pos:
                pos[(i, k)] = len(rows)
                rows.append(i)
                cols.append(k)
            return pos[(i, k)]

        for i in range(n):
            at(i, i)
        jp, ent, coef = [], [], []
        for j in range(R):
            for p in range(P):
                kk = int(self.react[j, p])
                if kk == n:
                    continue
                for i in np.nonzero(Sm[:, j])[0]:
                    jp.append(j * P + p)
                    ent.append(at(int(i), kk))
                    coef.append(Sm[i, j])
        self.jconst = {}
        for r, row in enumerate(self.con_rows):
            for s in np.nonzero(self.con_coef[r])[0]:
                self.jconst[at(int(row), int(s))] = self.con_coef[r, s]
        self.rows, self.cols = np.asarray(rows), np.asarray(cols)
        self.nnz = len(rows)
        self.GT = sp.csr_matrix((coef, (ent, jp)), shape=(self.nnz, R * P))
        self.jc_idx = np.asarray(list(self.jconst.keys()), dtype=np.int64)
        self.jc_val = np.asarray(list(self.jconst.values()), dtype=float)

    def _gather(self, Y):
        return np.concatenate([Y, np.ones((Y.shape[0], 1))], axis=1)[:, self.react]

    def rhs(self, k, Y):
        F = (k * self._gather(Y).prod(axis=2)) @ self.ST        # chem.py:86-90
        if len(self.con_rows):
            F[:, self.con_rows] = Y @ self.con_coef.T - self.con_total[None, :]
        return F

    def jac_values(self, k, Y):
        X = self._gather(Y)                                    # chem.py:92-102
        d = np.empty_like(X)
        for p in range(self.P):
            prod = np.broadcast_to(k, X.shape[:2]).copy()
            for q in range(self.P):
                if q != p:
                    prod *= X[:, :, q]
            d[:, :, p] = prod
        V = np.asarray((self.GT @ d.reshape(X.shape[0], -1).T).T)
        if len(self.jc_idx):
            V[:, self.jc_idx] = self.jc_val[None, :]
        return V

    def dense(self, vals):
        J = np.zeros((vals.shape[0], self.n, self.n))           # chem.py:104-108
        J[:, self.rows, self.cols] = vals
        return J


class SuperLU:
    """SuperLU on the block-diagonal CSC of the group (fastode/linalg.py:70-81)."""

    def __init__(self, mech, vals):
        C, n = vals.shape[0], mech.n
        rows = (mech.rows[None, :] + n * np.arange(C)[:, None]).ravel()
        cols = (mech.cols[None, :] + n * np.arange(C)[:, None]).ravel()
        A = sp.csc_matrix((vals.ravel(), (rows, cols)), shape=(C * n, C * n))
        self.lu = spla.splu(A, permc_spec="MMD_AT_PLUS_A")
        self.shape = (C, n)

    def solve(self, B):
        return self.lu.solve(np.ascontiguousarray(B).ravel()).reshape(self.shape)


class ChemSystem:
    """A mechanism with rate constants k (C, R) (fastode/chem.py:111-134):
    dense LAPACK when n <= 60, SuperLU otherwise."""

    def __init__(self, mech, k, mass=None, dense=None):
        self.mech = mech
        self.k = np.atleast_2d(k)
        self.n = mech.n
        self.mass = np.ones(self.n) if mass is None else np.asarray(mass, dtype=float)
        self.dense = (self.n <= 60) if dense is None else dense

    @property
    def has_constraints(self):
        return bool(np.any(self.mass == 0.0))

    def rhs(self, Y):
        return self.mech.rhs(self.k, Y)

    def jac(self, Y):
        v = self.mech.jac_values(self.k, Y)
        return self.mech.dense(v) if self.dense else v

    def factor(self, alpha, J, algebraic_rows_only=False):
        if self.dense:
            return LapackLU(stage_matrix_dense(alpha, self.mass, J, algebraic_rows_only))
        vals = -J                                                # linalg.py:60-67
        if algebraic_rows_only:
            vals = vals * (self.mass[self.mech.rows] == 0.0)[None, :]
        else:
            vals = vals.copy()
        vals[:, np.arange(self.n)] += alpha * self.mass[None, :]   # diagonal entries come first
        return SuperLU(self.mech, vals)


# ------------------------------------------------------------- the solver
def error_norm(Y, Ynew, E, atol, rtol):
    """RMS over the whole group of E / (atol + rtol max(|Y|, |Ynew|)), floored
    at 1e-10 unless NaN (fastode/solvers.py:37-43)."""
    w = atol[None, :] + rtol * np.maximum(np.abs(Y), np.abs(Ynew))
    s = E / w
    v = math.sqrt(float(np.sum(s * s)) / s.size)
    return v if math.isnan(v) else max(v, ERROR_MIN)


def quantize(h, interval):
    """The largest divisor interval/n not above h (fastode/solvers.py:46-51)."""
    if h >= interval:
        return interval
    return interval / max(math.ceil(interval / h - 1.0e-9), 1)


def stages(system, tab, Y, F0, LU, H, counts):
    """One KPP-form Rosenbrock attempt (fastode/solvers.py:54-86)."""
    dae = system.has_constraints
    mass = system.mass
    K = []
    forcing = F0
    flags = tab.new_forcing
    for r in range(tab.stages):
        if r > 0 and flags[r]:
            Yr = Y.copy()
            for s in range(r):
                if tab.a[r, s] != 0.0:
                    Yr += tab.a[r, s] * K[s]
            forcing = system.rhs(Yr)
            counts.rhs += 1
        rhs = forcing.copy()
        for s in range(r):
            coef = tab.c[r, s] / H
            if coef != 0.0:
                rhs += coef * (K[s] * mass[None, :] if dae else K[s])
        K.append(LU.solve(rhs))
        counts.solves += 1
    Ynew = Y.copy()
    E = np.zeros_like(Y)
    for r in range(tab.stages):
        if tab.m[r] != 0.0:
            Ynew += tab.m[r] * K[r]
        if tab.e[r] != 0.0:
            E += tab.e[r] * K[r]
    return Ynew, E


def clamp(system, Y):
    """max(Y, 0), on the differential columns only with constraints
    (fastode/solvers.py:89-97)."""
    if system.has_constraints:
        d = system.mass == 1.0
        Y[:, d] = np.maximum(Y[:, d], 0.0)
    else:
        np.maximum(Y, 0.0, out=Y)
    return Y


def project(system, Y, atol, rtol, counts, max_updates=10, tol=0.1, max_backtracks=24,
            backtrack_factor=0.5, sigma=1.0e-4):
    """Damped Newton on the algebraic rows for a consistent state
    (fastode/solvers.py:100-178)."""
    alg = system.mass == 0.0
    if not np.any(alg):
        return Y, "converged"
    atol_a = atol[alg]

    def residual(Z):
        R = system.rhs(Z)
        counts.rhs += 1
        R[:, ~alg] = 0.0
        return R

    def finite(A):
        return bool(np.all(np.isfinite(A[:, alg])))

    def weighted(Z, D):
        z, d = Z[:, alg], D[:, alg]
        sc = atol_a[None, :] + rtol * np.maximum(np.abs(z), np.abs(z + d))
        with np.errstate(divide="ignore", invalid="ignore"):
            q = np.where(sc > 0.0, np.abs(d) / sc, np.where(d == 0.0, 0.0, np.inf))
        return float(q.max())

    saved = None
    for update in range(max_updates + 1):
        Gv = residual(Y)
        if not finite(Gv):
            break
        counts.proj_passes += 1
        if float(np.abs(Gv[:, alg]).max()) == 0.0:
            return Y, "converged"
        J = system.jac(Y)
        counts.jac += 1
        LU = system.factor(1.0, J, algebraic_rows_only=True)
        counts.lu += 1
        delta = LU.solve(Gv)
        counts.solves += 1
        if not finite(delta):
            break
        q = weighted(Y, delta)
        if q <= tol:
            return Y, "converged"
        if update == max_updates:
            break
        if saved is None:
            saved = Y.copy()
        step = 1.0
        taken = None
        for _ in range(max_backtracks + 1):
            cand = Y.copy()
            cand[:, alg] += step * delta[:, alg]
            used = step
            step *= backtrack_factor
            if not finite(cand):
                continue
            Gc = residual(cand)
            counts.proj_trials += 1
            if not finite(Gc):
                continue
            if float(np.abs(Gc[:, alg]).max()) == 0.0:
                taken = cand
                break
            trial = LU.solve(Gc)
            counts.solves += 1
            if not finite(trial):
                continue
            qt = weighted(cand, trial)
            if qt <= tol or qt < (1.0 - sigma * used) * q:
                taken = cand
                break
        if taken is None:
            break
        Y = taken
    if saved is not None:
        Y = saved
    return Y, "projection failed"


class Algorithm:
    """The carried/quantized H211b Rosenbrock integrator
    (fastode/solvers.py:270-415).  ``trace``, when a list, receives one
    (H, err, accepted) tuple per attempt."""

    name = "algorithm"

    def __init__(self, tableau=ROS3, controller="h211b", b=4.0, safety=0.9, fac_min=0.2,
                 fac_max=6.0, carry=True, quantize=True, max_steps=1000, carry_memory=True):
        self.tableau = tableau
        self.controller = controller
        self.b, self.safety, self.fac_min, self.fac_max = b, safety, fac_min, fac_max
        self.carry, self.quantize, self.max_steps, self.carry_memory = carry, quantize, max_steps, carry_memory
        self.trace = None

    def new_memory(self):
        return Memory()

    def factor(self, err, H, mem):
        """H211b with the old memory, else elementary; clamped (solvers.py:297-303)."""
        k = self.tableau.k
        fac = self.safety / err ** (1.0 / k)
        if self.controller == "h211b" and mem.err_prev > 0.0 and mem.h_prev > 0.0:
            fac = (self.safety * (err * mem.err_prev) ** (-1.0 / (self.b * k))
                   * (H / mem.h_prev) ** (-1.0 / self.b))
        return min(self.fac_max, max(self.fac_min, fac))

    def solve(self, system, Y, dt, atol, rtol, counts, memory=None):
        tab = self.tableau
        mem = memory if memory is not None else Memory()
        if not (self.carry and self.carry_memory):
            mem.err_prev = mem.h_prev = 0.0
        counts.calls += 1
        h_min, h_max = H_MIN_FRACTION * dt, dt
        H = mem.h_next if (self.carry and mem.h_next > 0.0) else H_START_FRACTION * dt
        H = min(max(h_min, H), h_max)
        if self.quantize:
            H = quantize(H, dt)
        proposed = H
        Y = Y.copy()
        if system.has_constraints:
            Y, st = project(system, Y, atol, rtol, counts)
            if st != "converged":
                return Y, st
        t = 0.0
        rej_last = rej_more = False
        nsteps = 0
        status = "running"
        while (t - dt + EPS) <= 0.0 and status == "running":
            if nsteps >= self.max_steps:
                status = "too many steps"
                break
            if (t + 0.1 * H == t) or H < h_min:
                status = "step too small"
                break
            last = dt - t <= H * (1.0 + 1.0e-8)
            short = last and dt - t < H * (1.0 - 1.0e-8)
            H_ctrl = H
            if last:
                H = dt - t
            F0 = system.rhs(Y)
            counts.rhs += 1
            J = system.jac(Y)
            counts.jac += 1
            done = False
            while not done:
                if nsteps >= self.max_steps:
                    status = "too many steps"
                    break
                if H < h_min or t + 0.1 * H == t:
                    status = "step too small"
                    break
                LU = system.factor(1.0 / (H * tab.gamma), J)
                counts.lu += 1
                Ynew, E = stages(system, tab, Y, F0, LU, H, counts)
                err = error_norm(Y, Ynew, E, atol, rtol)
                nsteps += 1
                counts.steps += 1
                if self.trace is not None:
                    self.trace.append((H, err, err < 1.0))
                if math.isnan(err) or math.isinf(err):
                    status = "nan or inf"
                    break
                if err < 1.0 and short:
                    counts.accepted += 1
                    t = dt
                    Y = Ynew
                    proposed = H_ctrl
                    done = True
                elif err < 1.0:
                    fac = self.factor(err, H, mem)
                    counts.accepted += 1
                    t = dt if last else t + H
                    Y = Ynew
                    mem.err_prev, mem.h_prev = err, H
                    Hn = H * fac
                    if rej_last:
                        Hn = min(Hn, H)
                    Hn = max(h_min, min(Hn, h_max))
                    if self.quantize:
                        Hn = quantize(Hn, dt)
                    rej_last = rej_more = False
                    proposed = H = Hn
                    done = True
                else:
                    Hn = H * max(self.fac_min, self.safety / err ** (1.0 / tab.k))
                    if rej_more:
                        Hn = H * 0.1
                    rej_more = rej_last
                    rej_last = True
                    if self.quantize:
                        Hn = quantize(Hn, dt)
                    H = Hn
                    last = dt - t <= H * (1.0 + 1.0e-8)
                    short = last and dt - t < H * (1.0 - 1.0e-8)
                    H_ctrl = H
                    if last:
                        H = dt - t
                    counts.rejected += 1
        mem.h_next = proposed
        if status == "running":
            status = "converged"
        return clamp(system, Y), status


class RestartingLoop:
    """The baseline Rosenbrock loop (fastode/solvers.py:189-267): every call
    restarts from 1e-6 of the interval with the elementary controller and
    fixed constants (safety 0.9, factors in [0.2, 6]); the attempt budget is
    checked as ``nsteps > max_steps`` at the loop top only; a step is accepted
    when err < 1 or H < h_min; with constraints the algebraic rows of E are
    replaced by the step change (step_change_rows)."""

    name = "restarting"

    def __init__(self, tableau=ROS3, max_steps=1000, step_change_rows=True):
        self.tableau = tableau
        self.max_steps = max_steps
        self.step_change_rows = step_change_rows
        self.trace = None

    def new_memory(self):
        return None

    def solve(self, system, Y, dt, atol, rtol, counts, memory=None):
        tab = self.tableau
        counts.calls += 1
        h_min, h_max = H_MIN_FRACTION * dt, dt
        H = min(max(h_min, H_START_FRACTION * dt), h_max)
        Y = Y.copy()
        if system.has_constraints:
            Y, st = project(system, Y, atol, rtol, counts)
            if st != "converged":
                return Y, st
        alg = system.mass == 0.0
        t = 0.0
        rej_last = rej_more = False
        nsteps = 0
        status = "running"
        while (t - dt + EPS) <= 0.0 and status == "running":
            if nsteps > self.max_steps:
                status = "too many steps"
                break
            if (t + 0.1 * H == t) or H <= min(EPS, h_min):
                status = "step too small"
                break
            H = min(H, abs(dt - t))
            F0 = system.rhs(Y)
            counts.rhs += 1
            J = system.jac(Y)
            counts.jac += 1
            done = False
            while not done:
                LU = system.factor(1.0 / (H * tab.gamma), J)
                counts.lu += 1
                Ynew, E = stages(system, tab, Y, F0, LU, H, counts)
                if system.has_constraints and self.step_change_rows:
                    E[:, alg] = Ynew[:, alg] - Y[:, alg]
                err = error_norm(Y, Ynew, E, atol, rtol)
                nsteps += 1
                counts.steps += 1
                if self.trace is not None:
                    self.trace.append((H, err, err < 1.0 or H < h_min))
                if math.isnan(err) or math.isinf(err):
                    status = "nan or inf"
                    break
                fac = min(6.0, max(0.2, 0.9 / err ** (1.0 / tab.k)))
                Hn = H * fac
                if err < 1.0 or H < h_min:
                    counts.accepted += 1
                    t += H
                    Y = Ynew
                    Hn = max(h_min, min(Hn, h_max))
                    if rej_last:
                        Hn = min(Hn, H)
                    rej_last = rej_more = False
                    H = Hn
                    done = True
                else:
                    if rej_more:
                        Hn = H * 0.1
                    rej_more = rej_last
                    rej_last = True
                    H = Hn
                    counts.rejected += 1
        if status == "running":
            status = "converged"
        return clamp(system, Y), status


class BackwardEuler:
    """Backward Euler with Newton's iteration (fastode/solvers.py:418-491):
    each Newton pass evaluates F and J at the iterate, factors I/H - J
    (mass-weighted) and solves for the correction, then clamps the iterate at
    0; convergence is tested from the second pass on, over the whole group;
    a failed step restarts from the step's start with H cut by the next
    reduction, and after the last reduction the step is taken unconverged;
    two successes in a row double H."""

    name = "backward-euler"

    def __init__(self, max_iterations=11, reductions=(0.5, 0.5, 0.5, 0.5, 0.1), small=1.0e-40):
        self.max_iterations = max_iterations
        self.reductions = tuple(reductions)
        self.small = small

    def new_memory(self):
        return None

    def converged(self, delta, Y, atol, rtol):
        if not np.all(np.isfinite(delta)):
            return False
        bound = np.maximum(np.maximum(self.small, atol[None, :]), rtol * np.abs(Y))
        return bool(np.all(np.abs(delta) <= bound))

    def solve(self, system, Y, dt, atol, rtol, counts, memory=None):
        counts.calls += 1
        H = dt
        t = 0.0
        n_success = n_fail = 0
        Yn = Y.copy()
        Yn1 = Y.copy()
        status = "converged"
        while t < dt:
            ok = False
            iterations = 0
            while True:
                counts.newton += 1
                F = system.rhs(Yn1)
                counts.rhs += 1
                J = system.jac(Yn1)
                counts.jac += 1
                LU = system.factor(1.0 / H, J)
                counts.lu += 1
                delta = LU.solve(F - (Yn1 - Yn) / H)
                counts.solves += 1
                Yn1 = np.maximum(Yn1 + delta, 0.0)
                iterations += 1
                if iterations > 1:
                    ok = self.converged(delta, Yn1, atol, rtol)
                if ok or iterations >= self.max_iterations:
                    break
            counts.steps += 1
            if not ok:
                counts.rejected += 1
                counts.failures += 1
                n_success = 0
                if n_fail >= len(self.reductions):
                    t += H
                    counts.unconverged += 1
                    status = "accepted unconverged"
                    if not np.all(np.isfinite(Yn1)):
                        status = "nan or inf"
                    break
                Yn1 = Yn.copy()
                H *= self.reductions[n_fail]
                n_fail += 1
            else:
                counts.accepted += 1
                t += H
                Yn = Yn1.copy()
                n_success += 1
                if n_success >= 2:
                    n_success = 0
                    H *= 2.0
            H = min(H, dt - t)
        return clamp(system, Yn1), status


def fixed_steps(system, tab, Y, t_end, n_steps, counts=None):
    """n uncontrolled equal steps (fastode/solvers.py:494-504)."""
    counts = counts or Counts()
    H = t_end / n_steps
    for _ in range(n_steps):
        F0 = system.rhs(Y)
        J = system.jac(Y)
        LU = system.factor(1.0 / (H * tab.gamma), J)
        Y, _ = stages(system, tab, Y, F0, LU, H, counts)
    return Y


# ----------------------------------------------------- tangent-linear model
# Forward-mode derivative of the accepted steps of Algorithm.solve with the
# step sequence held fixed (quantized steps are locally constant), for a
# direction (dY, dlogk) in initial state and log rate constants.  It
# differentiates stages (fastode/solvers.py:54-86) with the mass-action
# rates and Jacobian of fastode/chem.py:81-102 and the clamp of
# solvers.py:89-97; unconstrained mechanisms only.

def _slots(mech, Y, pad):
    """Y (C, n) gathered at the reactant slots (C, R, P), pad for no species."""
    return np.concatenate([Y, np.full((Y.shape[0], 1), pad)], axis=1)[:, mech.react]


def _partials(mech, k, X):
    """d rate_j / d y_{react[j,p]} (C, R, P) from the slot values X."""
    d = np.empty_like(X)
    for p in range(mech.P):
        prod = np.broadcast_to(k, X.shape[:2]).copy()
        for q in range(mech.P):
            if q != p:
                prod *= X[:, :, q]
        d[:, :, p] = prod
    return d


def forcing_tangent(mech, k, Y, dY, dlogk):
    """d F(Y) along (dY, dlogk): J(Y) dY + S (r(Y) * dlogk)."""
    X = _slots(mech, Y, 1.0)
    dX = _slots(mech, dY, 0.0)
    r = k * X.prod(axis=2)
    v = (_partials(mech, k, X) * dX).sum(axis=2) + r * dlogk
    return v @ mech.ST


def jacobian_tangent(mech, k, Y, dY, dlogk, V):
    """(d J(Y) along (dY, dlogk)) V for mass action: the Hessian of the
    rates contracted with dY and V, plus the k-linear part of J."""
    X = _slots(mech, Y, 1.0)
    dX = _slots(mech, dY, 0.0)
    VX = _slots(mech, V, 0.0)
    P = mech.P
    d = _partials(mech, k, X)
    v = (d * VX).sum(axis=2) * dlogk
    for p in range(P):
        for q in range(P):
            if p == q:
                continue
            prod = np.broadcast_to(k, X.shape[:2]).copy()
            for s in range(P):
                if s != p and s != q:
                    prod *= X[:, :, s]
            v = v + prod * VX[:, :, p] * dX[:, :, q]
    return v @ mech.ST


def tangent_step(mech, k, tab, Y, dY, dlogk, H):
    """One step of size H from Y (the arithmetic of ``stages``) and its
    tangent: returns (Ynew, dYnew)."""
    system = ChemSystem(mech, k)
    F0 = system.rhs(Y)
    J = system.jac(Y)
    LU = system.factor(1.0 / (H * tab.gamma), J)
    K, dK = [], []
    forcing, dforcing = F0, forcing_tangent(mech, k, Y, dY, dlogk)
    for r in range(tab.stages):
        if r > 0 and tab.new_forcing[r]:
            Yr, dYr = Y.copy(), dY.copy()
            for s in range(r):
                if tab.a[r, s] != 0.0:
                    Yr += tab.a[r, s] * K[s]
                    dYr += tab.a[r, s] * dK[s]
            forcing, dforcing = system.rhs(Yr), forcing_tangent(mech, k, Yr, dYr, dlogk)
        rhs, drhs = forcing.copy(), dforcing.copy()
        for s in range(r):
            coef = tab.c[r, s] / H
            if coef != 0.0:
                rhs += coef * K[s]
                drhs += coef * dK[s]
        Kr = LU.solve(rhs)
        drhs = drhs + jacobian_tangent(mech, k, Y, dY, dlogk, Kr)
        K.append(Kr)
        dK.append(LU.solve(drhs))
    Ynew, dYnew = Y.copy(), dY.copy()
    for r in range(tab.stages):
        if tab.m[r] != 0.0:
            Ynew += tab.m[r] * K[r]
            dYnew += tab.m[r] * dK[r]
    return Ynew, dYnew


def tangent_linear(mech, k, kmuls, tab, Y0, dY0, dlogk, dt, atol, rtol, group_size=32, **options):
    """Groups of ``group_size`` cells, each with its own carried Memory, over
    len(kmuls) intervals (rate constants k * kmuls[i] in interval i), the
    loop of fastode/ts1.py:196-216.  Returns (Y_final, dY_final, steps), steps
    [interval][group] = the accepted step sizes."""
    alg = Algorithm(tab, **options)
    C = Y0.shape[0]
    groups = [slice(g, min(g + group_size, C)) for g in range(0, C, group_size)]
    mems = [alg.new_memory() for _ in groups]
    Y, dY = Y0.copy(), dY0.copy()
    steps = []
    for m in kmuls:
        kk = k * m[None, :]
        row = []
        for g, sl in enumerate(groups):
            alg.trace = []
            Yg, _ = alg.solve(ChemSystem(mech, kk[sl]), Y[sl], dt, atol, rtol, Counts(), mems[g])
            hs = [h for (h, _, acc) in alg.trace if acc]
            y, dy = Y[sl].copy(), dY[sl].copy()
            for h in hs:
                y, dy = tangent_step(mech, kk[sl], tab, y, dy, dlogk[sl], h)
            dY[sl] = np.where(y > 0.0, dy, 0.0)
            Y[sl] = Yg
            row.append(hs)
        steps.append(row)
    alg.trace = None
    return Y, dY, steps


PINS = {
    "Tableau/ROS3/RODAS3":"order conditions to 3, index-1 values, new_forcing flags (tests/test_oracle.py)",
    "error_norm": "closed form on hand vectors; golden per-attempt err from the reference",
    "quantize": "divisor property h/2 < q <= h, n q = interval",
    "Algorithm.factor": "closed-form H211b/elementary values; golden step sequences",
    "stages": "order > 2.8 on y' = -y^3; stability function R(z) of the standard form; Robertson invariant",
    "Algorithm.solve": "reference tests' exact-count cases; golden Counts/trace/Y of the seeded configs",
    "clamp": "golden final states (clamped) and hand vectors",
    "project": "affine covariance; Robertson manifold; golden constrained-config Counts",
    "OracleMechanism": "hand RHS of Robertson; Jacobian vs central differences; golden Y",
    "SuperLU/LapackLU": "residual against numpy.linalg.solve",
    "RestartingLoop": "exact TS1 Counts of the restart and restart-p rows of the reference's "
                      "reference-results/ts1/runs.csv; the reference test that a restart costs >= 9 "
                      "steps where the carried loop takes 1 (tests/test_fastode.py:96-110)",
    "BackwardEuler": "exact TS1 Counts of the be rows of reference-results/ts1/runs.csv; implicit "
                     "Euler's closed form on y' = -k y (one Newton pass is exact for a linear system)",
    "tangent_linear": "central differences of Algorithm.solve itself (fixed-step and controlled), "
                      "closed-form dy/dy0 and dy/dlogk of y' = -k y; forcing/jacobian tangents vs "
                      "complex-step derivatives of rhs/jac_values",
}
