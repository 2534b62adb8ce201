"""fastode_amd: the batched Rosenbrock integrator of
davidfillmore/Fast-ODE-Atmospheric-Chem (``Algorithm(ROS3|RODAS3).solve``),
rebuilt for AMD MI355X (gfx950) with hand-written FP64 HIP kernels behind a
C ABI (``include/fastode_hip.h``) and PyTorch-ROCm for device memory.

The public surface mirrors the reference's ``fastode/__init__.py:15-20``:
``ROS3, RODAS3, Tableau, Counts, Algorithm, project`` plus the device
systems (``ChemSystem``, ``DenseSystem``), the batched persistent form
(``BatchedAlgorithm``) and the discrete adjoint (``adjoint``).

This directory is loaded as the package ``fastode_amd`` (its own name holds
hyphens); ``__graft_entry__.load_package()`` registers it.
"""
from .tableaus import ROS3, RODAS3, Tableau
from .counts import Counts, Memory

__all__ = ["ROS3", "RODAS3", "Tableau", "Counts", "Memory"]


def __getattr__(name):
    # Device-facing pieces load the HIP library on first use, so that the
    # host-only modules (mechanism, symbolic, workloads) import without it.
    if name in ("Algorithm", "RestartingLoop", "BackwardEuler", "project", "clamp", "fixed_steps"):
        from . import solvers
        return getattr(solvers, name)
    if name in ("ChemSystem", "DenseSystem"):
        from . import system
        return getattr(system, name)
    if name in ("BatchedAlgorithm", "BatchedRestartingLoop", "BatchedBackwardEuler"):
        from . import batched
        return getattr(batched, name)
    if name in ("TS1", "Cells", "Scenario", "run_scenario", "run_scenario_batched", "score"):
        from . import ts1
        return getattr(ts1, name)
    if name in ("BatchedAdjoint", "integrate"):
        from . import adjoint
        return getattr(adjoint, name)
    if name in ("tune",):
        from . import autotune
        return getattr(autotune, name)
    raise AttributeError(name)
