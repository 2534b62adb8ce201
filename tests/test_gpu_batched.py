"""GPU: the batched persistent form against the golden vectors and against
the plain drop-in (identical Counts and bitwise-identical states), eager and
replayed from its hipGraph."""
import numpy as np
import pytest
import torch

from golden_util import GOLDEN, load, state_error, workload

pytestmark = pytest.mark.gpu


def kmul_for(wl, it):
    from fastode_amd.mechanism import diurnal_factor
    m = np.ones(wl.R)
    if wl.diurnal:
        m[wl.photolysis] = diurnal_factor(it * wl.dt)
    return m


def static_k(wl, cells):
    k = np.broadcast_to(wl.k_base, (len(cells), wl.R)).copy()
    if wl.cell_scale is not None:
        k *= wl.cell_scale[cells]
    return k


def replay_batched(name, graph):
    import fastode_amd as fa
    from fastode_amd.system import ChemSystem
    g = load(name)
    wl = workload(name)
    tab = fa.ROS3 if wl.tableau == "ros3" else fa.RODAS3
    cells = g["cells"]
    system = ChemSystem(wl.mech, static_k(wl, cells), mass=wl.mass)
    balg = fa.BatchedAlgorithm(tab)
    run = balg.runner(system, g["Y0"], wl.dt, wl.atol, wl.rtol, graph=graph)
    nint = int(g["n_intervals"])
    prev = run.memory.counts()
    worst = 0.0
    for it in range(nint):
        run.step(kmul_for(wl, it))
        rec = run.memory.read()
        cnt = rec["counts"] - prev
        prev = rec["counts"].copy()
        for gi in range(len(g["groups"])):
            assert list(cnt[gi]) == list(g["counts"][it, gi]), (it, gi, list(cnt[gi]), list(g["counts"][it, gi]))
            assert fa_status(rec["status"][gi]) == g["status"][it, gi]
        if (it + 1) in list(g["snap_interval"]):
            s = list(g["snap_interval"]).index(it + 1)
            ref = g["snaps"][s]
            Y = run.Y.cpu().numpy()
            scaled, rel = state_error(Y, ref, wl.atol)
            print(f"{name} interval {it + 1}: scaled error {scaled:.3e}, relative above the floor {rel:.3e}")
            worst = max(worst, scaled)
    mem = rec
    assert np.allclose(mem["h_next"], g["memory"][:, 0], rtol=1e-12)
    return worst, run.Y.cpu().numpy()


def fa_status(code):
    import fastode_amd._lib as L
    return L.STATUS[int(code)]


@pytest.mark.parametrize("graph", [False, True])
def test_batched_matches_golden_mech20(graph):
    worst, _ = replay_batched("mech20_dense", graph)
    assert worst <= 1.0


def test_batched_matches_golden_mech200_diurnal():
    worst, _ = replay_batched("mech200_diurnal", True)
    print("batched mech200 worst rel", worst)
    assert worst <= 1.0


def test_batched_matches_golden_mech200_uneven():
    if not (GOLDEN / "mech200_uneven.npz").exists():
        pytest.skip("golden vectors not generated")
    worst, _ = replay_batched("mech200_uneven", True)
    assert worst <= 1.0


@pytest.mark.parametrize("graph", [False, True])
def test_batched_matches_golden_mech100_constrained(graph):
    """RODAS3 with three conservation rows: the per-group projection runs in
    the persistent kernel at the start of every call."""
    worst, _ = replay_batched("mech100_constrained", graph)
    print("batched mech100 constrained worst", worst)
    assert worst <= 1.0


@pytest.mark.parametrize("stage_tables", [True, False])
def test_batched_constrained_equals_plain_bitwise(stage_tables):
    """The constrained workload on a ragged batch (40 cells: 32 + 8), the
    initial state pushed off its constraints so that the projection takes a
    Newton update: batched and plain agree bit for bit, Counts included, with
    the plan's tables staged in LDS and read from memory (two kernel
    instantiations)."""
    import fastode_amd as fa
    from fastode_amd import workloads as W
    from fastode_amd.system import ChemSystem
    wl = W.constrained100(C=40)
    Y0 = wl.Y0.copy()
    Y0[:, wl.mech.constraint_rows] *= 1.0 + 1e-3 * np.arange(1, 41)[:, None]
    k = static_k(wl, np.arange(40))
    system = ChemSystem(wl.mech, k, mass=wl.mass)
    run = fa.BatchedAlgorithm(fa.RODAS3, stage_tables=stage_tables).runner(system, Y0, wl.dt, wl.atol, wl.rtol,
                                                                          graph=False)
    alg = fa.Algorithm(fa.RODAS3)
    mems = [alg.new_memory() for _ in range(2)]
    cnt = [fa.Counts() for _ in range(2)]
    Yp = Y0.copy()
    for it in range(2):
        run.step(None)
        for gi, sl in enumerate([slice(0, 32), slice(32, 40)]):
            Yp[sl], st = alg.solve(ChemSystem(wl.mech, k[sl], mass=wl.mass), Yp[sl], wl.dt, wl.atol,
                                   wl.rtol, cnt[gi], mems[gi])
            assert st == "converged"
        assert np.array_equal(run.Y.cpu().numpy(), Yp)
    rec = run.memory.read()
    for gi in range(2):
        assert list(rec["counts"][gi]) == [getattr(cnt[gi], f) for f in
                                           ("steps", "accepted", "rejected", "rhs", "jac", "lu", "solves",
                                            "newton", "failures", "unconverged", "proj_passes",
                                            "proj_trials", "calls")]
        assert rec["h_next"][gi] == mems[gi].h_next
    assert rec["counts"][:, 10].min() >= 2          # the projection did a Newton pass


@pytest.mark.parametrize("rho,cut", [(0.75, 22), (0.85, 29)], ids=["all-tables-in-lds", "solve-tables-in-lds"])
def test_two_cells_per_cu_constrained_equals_plain_bitwise(rho, cut):
    """Two cells per CU on the constrained workload.  Under the tuned plan
    (22 sparse pivots, TILE 5) every table fits in LDS beside two cells;
    under the denser split (29 pivots) only the solves' tables do.  Either
    way the states, Counts and step proposals are the plain path's, bit for
    bit, on a ragged batch whose projection takes a Newton update."""
    import fastode_amd as fa
    from fastode_amd import workloads as W
    from fastode_amd.device import DevicePlan
    from fastode_amd.system import ChemSystem
    wl = W.constrained100(C=40)
    mech = wl.mech
    saved = mech.__dict__.get("_plan")
    mech._plan = DevicePlan(mech.n, mech.rows, mech.cols, rho=rho, m_max=160)
    try:
        assert (mech._plan.plan.c, mech._plan.plan.tile) == (cut, 5)
        Y0 = wl.Y0.copy()
        Y0[:, mech.constraint_rows] *= 1.0 + 1e-3 * np.arange(1, 41)[:, None]
        k = static_k(wl, np.arange(40))
        system = ChemSystem(mech, k, mass=wl.mass)
        run = fa.BatchedAlgorithm(fa.RODAS3, blocks_per_cu=2).runner(system, Y0, wl.dt, wl.atol, wl.rtol,
                                                                      graph=False)
        alg = fa.Algorithm(fa.RODAS3)
        mems = [alg.new_memory() for _ in range(2)]
        cnt = [fa.Counts() for _ in range(2)]
        Yp = Y0.copy()
        for it in range(2):
            run.step(None)
            for gi, sl in enumerate([slice(0, 32), slice(32, 40)]):
                Yp[sl], st = alg.solve(ChemSystem(mech, k[sl], mass=wl.mass), Yp[sl], wl.dt, wl.atol,
                                       wl.rtol, cnt[gi], mems[gi])
                assert st == "converged"
            assert np.array_equal(run.Y.cpu().numpy(), Yp)
        rec = run.memory.read()
    finally:
        if saved is None:
            del mech._plan
        else:
            mech._plan = saved
    for gi in range(2):
        assert list(rec["counts"][gi]) == [getattr(cnt[gi], f) for f in
                                           ("steps", "accepted", "rejected", "rhs", "jac", "lu", "solves",
                                            "newton", "failures", "unconverged", "proj_passes",
                                            "proj_trials", "calls")]
        assert rec["h_next"][gi] == mems[gi].h_next


def test_batched_equals_plain_bitwise_and_graph_replays_match_eager():
    """Groups of uneven size (the last one short), 3 intervals: the batched
    states equal the plain path's bit for bit, with eager work between graph
    replays."""
    import fastode_amd as fa
    from fastode_amd import workloads as W
    from fastode_amd.system import ChemSystem
    wl = W.small(C=80)          # 3 groups: 32, 32, 16
    k = static_k(wl, np.arange(80))
    system = ChemSystem(wl.mech, k)
    balg = fa.BatchedAlgorithm(fa.ROS3)
    eager = balg.runner(system, wl.Y0, wl.dt, wl.atol, wl.rtol, graph=False)
    graphed = balg.runner(system, wl.Y0, wl.dt, wl.atol, wl.rtol, graph=True)
    alg = fa.Algorithm(fa.ROS3)
    mems = [alg.new_memory() for _ in range(3)]
    Yp = wl.Y0.copy()
    for it in range(3):
        km = kmul_for(wl, it)
        eager.step(km)
        junk = torch.randn(1000, device="cuda").sum()      # eager work between replays
        graphed.step(km)
        torch.cuda.synchronize()
        _ = float(junk)
        for gi, sl in enumerate([slice(0, 32), slice(32, 64), slice(64, 80)]):
            c = fa.Counts()
            Yp[sl], st = alg.solve(ChemSystem(wl.mech, wl.k(it)[sl]), Yp[sl], wl.dt, wl.atol, wl.rtol, c,
                                   mems[gi])
        Ye = eager.Y.cpu().numpy()
        Yg = graphed.Y.cpu().numpy()
        assert np.array_equal(Ye, Yg)
        assert np.array_equal(Ye, Yp)
    re, rg = eager.memory.read(), graphed.memory.read()
    assert np.array_equal(re["counts"], rg["counts"])
    for gi in range(3):
        assert re["h_next"][gi] == mems[gi].h_next


@pytest.mark.parametrize("graph", [False, True])
def test_shared_rate_row_equals_per_cell_rows(graph):
    """One shared rate-constant row (k_stride 0, scaled once per launch) and
    the same row repeated per cell give the same bits, over intervals with
    different multipliers (the per-launch scaling follows each interval's)."""
    import fastode_amd as fa
    from fastode_amd import workloads as W
    from fastode_amd.system import ChemSystem
    wl = W.diurnal200(C=40)     # photolysis multipliers change every interval
    shared = ChemSystem(wl.mech, wl.k_base[None, :])
    per_cell = ChemSystem(wl.mech, np.broadcast_to(wl.k_base, (40, wl.R)).copy())
    balg = fa.BatchedAlgorithm(fa.ROS3)
    a = balg.runner(shared, wl.Y0, wl.dt, wl.atol, wl.rtol, graph=graph)
    b = balg.runner(per_cell, wl.Y0, wl.dt, wl.atol, wl.rtol, graph=graph)
    for it in range(3):
        km = kmul_for(wl, it + 3)
        if it:
            assert not np.array_equal(km, kmul_for(wl, it + 2))
        a.step(km)
        b.step(km)
        assert np.array_equal(a.Y.cpu().numpy(), b.Y.cpu().numpy()), it
    assert np.array_equal(a.memory.counts(), b.memory.counts())
