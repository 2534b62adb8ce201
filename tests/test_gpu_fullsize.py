"""GPU: the batched form at the workloads' full sizes, with the golden
groups inside them.

The reference-written golden vectors cover two groups of each workload; the
full workloads draw the same rows for them (the 200-species diurnal workload:
groups 0 and 37 of 2,048; the uneven one: groups 0 and 777 of 8,192).  So a
run over every group must give those two groups the golden Counts, statuses
and states, exactly as the two-group runs do, while the other groups share
the launches and the queue.  Every group must also converge.
"""
import numpy as np
import pytest

from golden_util import load, state_error

pytestmark = pytest.mark.gpu


def full_run(name, wl, golden_groups):
    import fastode_amd as fa
    import fastode_amd._lib as L
    import torch
    from fastode_amd.system import ChemSystem
    g = load(name)
    k = wl.k_base[None, :] if wl.cell_scale is None else wl.k_base[None, :] * wl.cell_scale
    system = ChemSystem(wl.mech, torch.as_tensor(k, device="cuda"), mass=wl.mass)
    run = fa.BatchedAlgorithm(fa.ROS3).runner(system, wl.Y0, wl.dt, wl.atol, wl.rtol, graph=True)
    prev = run.memory.counts()
    worst = 0.0
    cells = np.concatenate([np.arange(32 * gg, 32 * gg + 32) for gg in golden_groups])
    for it in range(int(g["n_intervals"])):
        run.step(wl.kmul(it))
        rec = run.memory.read()
        cnt = rec["counts"] - prev
        prev = rec["counts"].copy()
        for i, gg in enumerate(golden_groups):
            assert list(cnt[gg]) == list(g["counts"][it, i]), (it, gg)
            assert L.STATUS[int(rec["status"][gg])] == g["status"][it, i]
        assert (rec["status"] == L.CONVERGED).all(), (it, np.flatnonzero(rec["status"] != L.CONVERGED)[:5])
        if (it + 1) in list(g["snap_interval"]):
            s = list(g["snap_interval"]).index(it + 1)
            Y = run.Y[torch.as_tensor(cells, device="cuda")].cpu().numpy()
            scaled, rel = state_error(Y, g["snaps"][s])
            print(f"{name} full size, interval {it + 1}: scaled error {scaled:.3e}, rel {rel:.3e}", flush=True)
            worst = max(worst, scaled)
    assert np.allclose(rec["h_next"][list(golden_groups)], g["memory"][:, 0], rtol=1e-12)
    att = rec["counts"][:, 0] / int(g["n_intervals"])
    print(f"{name}: {run.ngroups} groups, attempts per group-interval min {att.min():.2f} "
          f"median {np.median(att):.2f} max {att.max():.2f}")
    return worst


def test_full_65536_cell_diurnal_workload_keeps_the_golden_groups():
    from fastode_amd import workloads as W
    wl = W.diurnal200(C=65536, n_intervals=48)
    assert full_run("mech200_diurnal", wl, (0, 37)) <= 1.0


def test_full_262144_cell_uneven_workload_keeps_the_golden_groups():
    from fastode_amd import workloads as W
    wl = W.uneven200(C=262144, n_intervals=24)
    assert full_run("mech200_uneven", wl, (0, 777)) <= 1.0
