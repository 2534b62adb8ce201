"""GPU: the autotuner's candidate plans all keep parity with the golden
vectors (the choice changes roundings only), and tuning installs and caches
a plan."""
import numpy as np
import pytest
import torch

from golden_util import load, state_error, workload

pytestmark = pytest.mark.gpu


def kmul_for(wl, it):
    return wl.kmul(it)


@pytest.fixture(scope="module")
def mech200():
    return workload("mech200_diurnal")


def test_every_candidate_plan_matches_golden_counts(mech200):
    import fastode_amd as fa
    from fastode_amd.autotune import candidates
    from fastode_amd.system import ChemSystem
    g = load("mech200_diurnal")
    wl = mech200
    cells = g["cells"]
    k = np.broadcast_to(wl.k_base, (len(cells), wl.R)).copy()
    cands = candidates(wl.mech)
    assert len(cands) >= 4
    default = (cands[0][0], cands[0][1])
    for rho, m_max, dp in cands:
        wl.mech._plan = dp
        for stage in ((True, False) if (rho, m_max) == default else (True,)):
            run = fa.BatchedAlgorithm(fa.ROS3, stage_tables=stage).runner(
                ChemSystem(wl.mech, k), g["Y0"], wl.dt, wl.atol, wl.rtol, graph=False)
            prev = run.memory.counts()
            n_int = 24
            for it in range(n_int):
                run.step(kmul_for(wl, it))
                cnt = run.memory.counts()
                assert np.array_equal(cnt - prev, g["counts"][it]), (rho, m_max, stage, it)
                prev = cnt
            s = list(g["snap_interval"]).index(n_int)
            scaled, rel = state_error(run.Y.cpu().numpy(), g["snaps"][s])
            print(f"plan rho {rho} m_max {m_max} c {dp.plan.c} tile {dp.plan.tile} stage {stage}: "
                  f"scaled state error {scaled:.2e}")
            assert scaled <= 1.0
    del wl.mech._plan


def test_tune_installs_and_caches(tmp_path, monkeypatch):
    import fastode_amd.autotune as A
    from fastode_amd import workloads as W
    monkeypatch.setattr(A, "CACHE", tmp_path / "tune.json")
    wl = W.diurnal200(C=1024)
    choice = A.tune(wl.mech, wl.k_base[None, :], wl.Y0, wl.dt, wl.atol, wl.rtol, kmul=wl.kmul(0),
                    verbose=True)
    assert wl.mech._tuned == choice and wl.mech._plan.plan.c == choice.c
    assert (tmp_path / "tune.json").exists()
    again = A.tune(wl.mech, wl.k_base[None, :], wl.Y0, wl.dt, wl.atol, wl.rtol, kmul=wl.kmul(0))
    assert again == choice
    torch.cuda.synchronize()
