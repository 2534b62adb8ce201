"""GPU: fode_batched_interval_noalloc (the entry point that allocates
nothing) against fode_batched_interval on the same inputs: states and
per-group controller blocks equal bit for bit, on two calls in a row, the
second with a smaller batch; and a batch whose group cannot fit the
library's scratch is refused with an error, not launched."""
import ctypes as ct

import numpy as np
import pytest
import torch

from golden_util import workload

pytestmark = pytest.mark.gpu


def pair(name, C, solver, kmul_interval=0):
    """Two runners on the same inputs: one for fode_batched_interval, one
    whose buffers (state, controller blocks) the no-alloc call uses."""
    from fastode_amd.system import ChemSystem
    wl = workload(name)
    cells = np.arange(C)
    k = wl.k(kmul_interval, cells)
    system = ChemSystem(wl.mech, k, mass=wl.mass)
    a = solver.runner(system, wl.Y0[:C], wl.dt, wl.atol, wl.rtol, graph=False)
    b = solver.runner(system, wl.Y0[:C], wl.dt, wl.atol, wl.rtol, graph=False)
    return wl, a, b


def noalloc_step(L, runner):
    desc = L.BatchDesc.from_buffer_copy(runner.desc)
    desc.work = None
    desc.queue = None
    rc = L.lib.fode_batched_interval_noalloc(runner.plan.handle, runner.dmech.handle, ct.byref(runner.params),
                                             ct.byref(desc), int(runner.balg.blocks_per_cu), L.stream_handle())
    L.check(rc, "fode_batched_interval_noalloc")


def compare(a, b):
    torch.cuda.synchronize()
    assert torch.equal(a.Y, b.Y), (a.Y - b.Y).abs().max().item()
    assert torch.equal(a.memory.dev, b.memory.dev)


@pytest.mark.parametrize("kind", ["algorithm", "restarting", "backward-euler", "constrained"])
def test_noalloc_equals_the_allocating_entry_point(kind):
    import fastode_amd as fa
    from fastode_amd import _lib as L
    name = "mech100_constrained" if kind == "constrained" else "mech200_diurnal"
    make = {"algorithm": lambda: fa.BatchedAlgorithm(fa.ROS3),
            "restarting": lambda: fa.BatchedRestartingLoop(fa.ROS3),
            "backward-euler": lambda: fa.BatchedBackwardEuler(),
            "constrained": lambda: fa.BatchedAlgorithm(fa.RODAS3)}[kind]
    sizes = (32 + 17, 20) if kind == "constrained" else (5 * 32 + 7, 2 * 32 + 3)   # the constrained workload has 64 cells
    for C in sizes:                              # a second call with a smaller batch
        print(f"no-alloc {kind}: {C} cells", flush=True)
        wl, a, b = pair(name, C, make())
        for it in range(2):
            kmul = wl.kmul(it)
            a.step(kmul)
            b.set_kmul(kmul)
            noalloc_step(L, b)
            compare(a, b)
        rec = b.memory.read()
        assert (rec["counts"][:, 12] == 2).all()      # two calls per group


def test_noalloc_refuses_a_group_beyond_the_scratch():
    import fastode_amd as fa
    from fastode_amd import _lib as L
    from fastode_amd.system import ChemSystem
    wl = workload("mech200_diurnal")
    gs = (32 << 20) // (8 * wl.n) + 1               # one group's rows exceed the 32 MiB scratch
    solver = fa.BatchedAlgorithm(fa.ROS3, group_size=gs)
    Y = np.tile(wl.Y0[:1], (gs, 1))
    runner = solver.runner(ChemSystem(wl.mech, wl.k(0, np.zeros(1, dtype=int))), Y, wl.dt, wl.atol, wl.rtol,
                           graph=False)
    before = runner.Y.clone()
    desc = L.BatchDesc.from_buffer_copy(runner.desc)
    desc.work = None
    desc.queue = None
    rc = L.lib.fode_batched_interval_noalloc(runner.plan.handle, runner.dmech.handle, ct.byref(runner.params),
                                             ct.byref(desc), 0, L.stream_handle())
    assert rc != 0
    assert b"scratch" in L.lib.fode_last_error()
    torch.cuda.synchronize()
    assert torch.equal(runner.Y, before)
    # work and queue given: refused too (the scratch takes their place)
    rc = L.lib.fode_batched_interval_noalloc(runner.plan.handle, runner.dmech.handle, ct.byref(runner.params),
                                             ct.byref(runner.desc), 0, L.stream_handle())
    assert rc != 0
