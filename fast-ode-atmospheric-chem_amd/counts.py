"""Operation counters and carried controller memory.

``Counts`` mirrors fastode/counts.py:9-30: 13 counters, each counting calls
on a whole group (so per group and interval they are per cell and interval).
On the device a group's counters are 13 consecutive int64 in this order.
``Memory`` mirrors fastode/solvers.py:181-186 (3 float64 per group on device).
"""
from __future__ import annotations

from dataclasses import dataclass, fields

COUNT_FIELDS = ("steps", "accepted", "rejected", "rhs", "jac", "lu", "solves", "newton",
                "failures", "unconverged", "proj_passes", "proj_trials", "calls")


@dataclass
class Counts:
    steps: int = 0          # attempted steps (accepted + rejected)
    accepted: int = 0
    rejected: int = 0
    rhs: int = 0            # right-hand-side evaluations
    jac: int = 0            # Jacobian evaluations
    lu: int = 0             # factorizations
    solves: int = 0         # triangular-solve pairs
    newton: int = 0         # Newton iterations (backward Euler; kept for the layout)
    failures: int = 0
    unconverged: int = 0
    proj_passes: int = 0    # projection passes
    proj_trials: int = 0    # projection line-search trials
    calls: int = 0          # solver calls (intervals)

    def add(self, other: "Counts") -> None:
        for f in fields(self):
            setattr(self, f.name, getattr(self, f.name) + getattr(other, f.name))

    def as_dict(self) -> dict:
        return {f.name: getattr(self, f.name) for f in fields(self)}

    def add_array(self, arr) -> None:
        """Add a device counter block (13 integers in COUNT_FIELDS order)."""
        for name, v in zip(COUNT_FIELDS, arr):
            setattr(self, name, getattr(self, name) + int(v))


@dataclass
class Memory:
    h_next: float = 0.0
    err_prev: float = 0.0
    h_prev: float = 0.0
