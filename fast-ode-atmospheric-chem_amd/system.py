"""Device systems: the reference's System protocol on torch device tensors.

The protocol is that of fastode/system.py:14-30: ``n``, ``mass``, ``names``,
``rhs(Y) -> (C, n)``, ``jac(Y) -> J`` (opaque, consumed only by ``factor``),
``factor(alpha, J, algebraic_rows_only=False) -> obj`` with
``obj.solve(B) -> (C, n)``, and ``has_constraints``.  Every array is a
float64 tensor on the HIP device; ``alpha`` may be a Python float or a
0-dim device tensor (the device controller hands the latter).

* ``ChemSystem(mech, k, mass)`` mirrors fastode/chem.py:111-134: the forcing,
  Jacobian values and the factorization run in the HIP kernels.  ``mech`` is
  a ``fastode_amd.mechanism.Mechanism`` or any object with the reference
  Mechanism's arrays (compiled once by ``Mechanism.adopt`` and kept on it).
* ``DenseSystem(n, rhs, jac, mass)`` mirrors fastode/system.py:33-50: rhs and
  jac are the caller's callables, on torch tensors or on NumPy arrays (found
  out at the first call), the factorization is the HIP LU on a dense plan.

NumPy arguments get NumPy results: ``rhs(Y)`` and ``LU.solve(B)`` copy a
host array to the device and the result back, and ``jac(Y)`` returns the
device values (opaque to the solver, as the protocol has it), so the
reference's own solvers (fastode/solvers.py) can drive these systems.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from .device import dense_plan, device_mechanism, mechanism_plan
from .mechanism import Mechanism

F64 = torch.float64


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("fastode_amd needs a HIP device (MI355X); none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def as_device(x, shape=None):
    t = torch.as_tensor(x, dtype=F64, device=device())
    if shape is not None:
        t = t.reshape(shape)
    return t.contiguous()


def alpha_tensor(alpha):
    if isinstance(alpha, torch.Tensor):
        return alpha
    return torch.tensor([float(alpha)], dtype=F64, device=device())


class HipLU:
    """A stored factorization of a group's stage matrices (C, lu_len)."""

    def __init__(self, plan, LU, n):
        self.plan = plan
        self.LU = LU
        self.n = n

    def solve(self, B, out=None):
        host = not isinstance(B, torch.Tensor)
        B = as_device(B) if host else B.contiguous()
        C = B.shape[0]
        X = torch.empty_like(B) if out is None else out
        L.check(L.lib.fode_solve(self.plan.handle, C, L.ptr(self.LU), L.ptr(B), L.ptr(X),
                                 L.stream_handle()), "fode_solve")
        return X.cpu().numpy() if host else X


def adopted(mech) -> Mechanism:
    """The device Mechanism of mech: itself, or the compiled form of a
    reference-built mechanism, made once and kept on the object."""
    if isinstance(mech, Mechanism):
        return mech
    m = getattr(mech, "_fastode_amd_mech", None)
    if m is None:
        m = Mechanism.adopt(mech)
        try:
            mech._fastode_amd_mech = m
        except AttributeError:
            pass
    return m


class _Base:
    n: int
    mass_np: np.ndarray

    @property
    def has_constraints(self) -> bool:
        return bool(np.any(self.mass_np == 0.0))

    @property
    def mass(self):
        return self.mass_np

    def _factor(self, plan, alpha, J, algebraic_rows_only):
        J = J.contiguous()
        C = J.shape[0]
        LU = torch.empty((C, plan.lu_len), dtype=F64, device=J.device)
        a = alpha_tensor(alpha)
        L.check(L.lib.fode_factor(plan.handle, C, L.ptr(a), L.ptr(self.mass_dev), L.ptr(J),
                                  int(bool(algebraic_rows_only)), L.ptr(LU), L.stream_handle()),
                "fode_factor")
        return HipLU(plan, LU, self.n)


class ChemSystem(_Base):
    """A mechanism with rate constants k (C, R); mass[i] = 0 makes row i
    algebraic (fastode/chem.py:111-134).  ``dense`` is accepted for the
    reference's signature; the device always factors on the symbolic plan."""

    def __init__(self, mech, k, mass=None, dense=None):
        mech = adopted(mech)
        self.mech = mech
        self.n = mech.n
        self.names = mech.species
        self.k = as_device(np.atleast_2d(k) if not isinstance(k, torch.Tensor) else k)
        if self.k.dim() == 1:
            self.k = self.k[None, :]
        self.mass_np = np.ones(self.n) if mass is None else np.asarray(mass, dtype=float)
        self.mass_dev = as_device(self.mass_np)
        self.dense = dense
        self._dm = device_mechanism(mech)
        self._plan = mechanism_plan(mech)

    def _k_for(self, C):
        """k with one row per cell; a shared row is expanded once and kept, so
        the kernels never read a freed temporary."""
        if self.k.shape[0] == C:
            return self.k
        if self.k.shape[0] == 1:
            kc = getattr(self, "_k_expanded", None)
            if kc is None or kc.shape[0] != C:
                kc = self.k.expand(C, -1).contiguous()
                self._k_expanded = kc
            return kc
        raise ValueError(f"k has {self.k.shape[0]} rows for {C} cells")

    def rhs(self, Y, out=None):
        host = not isinstance(Y, torch.Tensor)
        Y = as_device(Y) if host else Y.contiguous()
        C = Y.shape[0]
        F = torch.empty_like(Y) if out is None else out
        L.check(L.lib.fode_rhs(self._dm.handle, C, L.ptr(Y), L.ptr(self._k_for(C)), L.ptr(F),
                               L.stream_handle()), "fode_rhs")
        return F.cpu().numpy() if host else F

    def jac(self, Y):
        Y = as_device(Y) if not isinstance(Y, torch.Tensor) else Y.contiguous()
        C = Y.shape[0]
        J = torch.empty((C, self.mech.nnz), dtype=F64, device=Y.device)
        L.check(L.lib.fode_jac(self._dm.handle, C, L.ptr(Y), L.ptr(self._k_for(C)), L.ptr(J),
                               L.stream_handle()), "fode_jac")
        return J

    def factor(self, alpha, J, algebraic_rows_only=False):
        return self._factor(self._plan, alpha, J, algebraic_rows_only)

    def dense_jacobian(self, J):
        """(C, nnz) values as dense (C, n, n) matrices (fastode/chem.py:104-108)."""
        D = torch.zeros((J.shape[0], self.n, self.n), dtype=F64, device=J.device)
        r = torch.as_tensor(self.mech.rows, device=J.device)
        c = torch.as_tensor(self.mech.cols, device=J.device)
        D[:, r, c] = J
        return D


class DenseSystem(_Base):
    """A system given by callables rhs(Y) -> (C, n) and jac(Y) -> (C, n, n)
    (fastode/system.py:33-50), written for torch tensors or for NumPy
    arrays: a callable that fails on a device tensor is called from then on
    with host copies, and results come back in the caller's kind of array."""

    def __init__(self, n, rhs, jac, mass=None, names=None):
        self.n = n
        self._rhs = rhs
        self._jac = jac
        self._host = {}
        self.mass_np = np.ones(n) if mass is None else np.asarray(mass, dtype=float)
        self.mass_dev = as_device(self.mass_np)
        self.names = names or [f"y{i}" for i in range(n)]
        self._plan = dense_plan(n)

    def _call(self, which, fn, Y):
        tensor_in = isinstance(Y, torch.Tensor)
        if tensor_in and not self._host.get(which, False):
            try:
                out = fn(Y)
            except (TypeError, RuntimeError, AttributeError, ValueError):
                self._host[which] = True
                out = fn(Y.cpu().numpy())
        elif tensor_in:
            out = fn(Y.cpu().numpy())
        else:
            out = fn(Y)
        if tensor_in and not isinstance(out, torch.Tensor):
            return as_device(out)
        if not tensor_in and isinstance(out, torch.Tensor):
            return out.cpu().numpy()
        return out

    def rhs(self, Y):
        return self._call("rhs", self._rhs, Y)

    def jac(self, Y):
        return self._call("jac", self._jac, Y)

    def factor(self, alpha, J, algebraic_rows_only=False):
        J = torch.as_tensor(J, dtype=F64, device=device())
        return self._factor(self._plan, alpha, J.reshape(J.shape[0], self.n * self.n),
                            algebraic_rows_only)
