"""GPU: RestartingLoop and BackwardEuler on the device (plain and batched),
and the drop-in boundary with the reference's own kinds of arguments.

* RestartingLoop / BackwardEuler against the oracle (fastode/solvers.py:
  189-267, 418-491 restated) on seeded synthetic mechanisms: exact Counts and
  statuses, states to the golden bar; the batched forms equal the plain ones
  bit for bit.
* The reference's known answers: a restart costs >= 9 steps where the
  carried loop takes 1 (tests/test_fastode.py:96-110); backward Euler's
  closed form on a linear decay; its failure ladder ending "accepted
  unconverged".
* The boundary: a mechanism object with exactly the reference Mechanism's
  attributes (fastode/chem.py:19-79) drives ``ChemSystem``; NumPy states in,
  NumPy states out; a NumPy solver written like the reference's (the
  oracle's restatement of fastode/solvers.py) drives the device systems
  through rhs / jac / factor / solve; DenseSystem takes NumPy callables.
"""
import types

import numpy as np
import pytest
import torch

from golden_util import COUNT_NAMES, load, state_error, workload

pytestmark = pytest.mark.gpu


def dev(x):
    return torch.as_tensor(x, dtype=torch.float64, device="cuda")


def counts_list(c):
    return [getattr(c, f) for f in COUNT_NAMES]


# ------------------------------------------------- restart / backward Euler
def oracle_and_device(kind, wl, groups, n_intervals, solver_kw=None):
    """Run the oracle and the plain device solver side by side group by
    group; return the per-call Counts of both and the final states."""
    import fastode_amd as fa
    import fastode_oracle as O
    from fastode_amd.system import ChemSystem
    solver_kw = solver_kw or {}
    om = O.OracleMechanism(wl.mech.species, wl.mech.reactions_list())
    if kind == "restarting":
        osol, dsol = O.RestartingLoop(O.ROS3, **solver_kw), fa.RestartingLoop(fa.ROS3, **solver_kw)
    else:
        osol, dsol = O.BackwardEuler(**solver_kw), fa.BackwardEuler(**solver_kw)
    Yo = wl.Y0.copy()
    Yd = wl.Y0.copy()
    rows = []
    for it in range(n_intervals):
        for g in groups:
            sl = slice(g * 32, g * 32 + 32)
            k = wl.k(it, np.arange(sl.start, sl.stop))
            oc, dc = O.Counts(), fa.Counts()
            Yo[sl], ost = osol.solve(O.ChemSystem(om, k, dense=wl.dense), Yo[sl], wl.dt, wl.atol, wl.rtol, oc)
            Yd[sl], dst = dsol.solve(ChemSystem(wl.mech, k), Yd[sl], wl.dt, wl.atol, wl.rtol, dc)
            rows.append((it, g, ost, dst, counts_list(oc), counts_list(dc)))
    return rows, Yo, Yd


@pytest.mark.parametrize("kind", ["restarting", "backward-euler"])
@pytest.mark.parametrize("name,groups,n_int", [("mech20_dense", [0, 1], 10), ("mech200_diurnal", [0, 37], 3)])
def test_plain_solver_matches_the_oracle(kind, name, groups, n_int):
    wl = workload(name)
    rows, Yo, Yd = oracle_and_device(kind, wl, groups, n_int)
    for it, g, ost, dst, oc, dc in rows:
        assert ost == dst, (it, g, ost, dst)
        assert oc == dc, (it, g, oc, dc)
    cells = np.concatenate([np.arange(g * 32, g * 32 + 32) for g in groups])
    scaled, rel = state_error(Yd[cells], Yo[cells])
    print(f"{kind} {name}: {sum(r[4][0] for r in rows)} steps, scaled state error {scaled:.2e}, rel {rel:.2e}")
    assert scaled <= 1.0


@pytest.mark.parametrize("kind", ["restarting", "backward-euler"])
def test_batched_solver_equals_plain(kind):
    """Batched restart / backward Euler over 4 groups (one ragged) take the
    plain path's steps and land on its states bit for bit."""
    import fastode_amd as fa
    from fastode_amd.system import ChemSystem
    wl = workload("mech200_diurnal")
    C = 3 * 32 + 17
    k = wl.k(0, np.arange(C))
    system = ChemSystem(wl.mech, k)
    if kind == "restarting":
        plain, batched = fa.RestartingLoop(fa.ROS3), fa.BatchedRestartingLoop(fa.ROS3)
    else:
        plain, batched = fa.BackwardEuler(), fa.BatchedBackwardEuler()
    Y0 = wl.Y0[:C].copy()
    runner = batched.runner(system, Y0, wl.dt, wl.atol, wl.rtol, graph=True)
    from fastode_amd import _lib as L
    Yp = Y0.copy()
    for it in range(2):
        runner.step(None)
        rec = runner.memory.read()
        for gi, g0 in enumerate(range(0, C, 32)):
            sl = slice(g0, min(g0 + 32, C))
            c = fa.Counts()
            Yp[sl], st = plain.solve(ChemSystem(wl.mech, k[sl]), Yp[sl], wl.dt, wl.atol, wl.rtol, c)
            assert L.STATUS[int(rec["status"][gi])] == st
    Yb = runner.Y.cpu().numpy()
    assert np.array_equal(Yb, Yp), np.abs(Yb - Yp).max()
    assert (runner.memory.read()["counts"][:, COUNT_NAMES.index("calls")] == 2).all()


def decay_np(rate):
    import fastode_amd as fa
    return fa.DenseSystem(1, lambda Y: -rate * Y, lambda Y: np.full((Y.shape[0], 1, 1), -rate))


def test_restart_costs_nine_steps():
    """tests/test_fastode.py:96-110 with NumPy callables, as the reference
    writes them."""
    import fastode_amd as fa
    sys_ = decay_np(1e-6)
    c = fa.Counts()
    fa.RestartingLoop(fa.ROS3).solve(sys_, np.array([[1.0]]), 600.0, np.array([1e-10]), 1e-3, c)
    assert c.accepted >= 9
    alg = fa.Algorithm(fa.ROS3)
    mem = alg.new_memory()
    Y = np.array([[1.0]])
    c2 = fa.Counts()
    for _ in range(5):
        Y, _ = alg.solve(sys_, Y, 600.0, np.array([1e-10]), 1e-3, c2, mem)
    c3 = fa.Counts()
    alg.solve(sys_, Y, 600.0, np.array([1e-10]), 1e-3, c3, mem)
    assert c3.accepted == 1


def test_backward_euler_closed_form():
    """On y' = -k y the first Newton pass solves (1 + kH) y1 = y0 exactly;
    the second meets the bound: two passes, one step, y0/(1 + k dt)."""
    import fastode_amd as fa
    k, dt = 0.3, 2.0
    c = fa.Counts()
    Y, st = fa.BackwardEuler().solve(decay_np(k), np.array([[1.0]]), dt, np.array([1e-12]), 1e-6, c)
    assert st == "converged"
    assert (c.steps, c.accepted, c.newton, c.rhs, c.jac, c.lu, c.solves) == (1, 1, 2, 2, 2, 2, 2)
    assert abs(Y[0, 0] - 1.0 / (1.0 + k * dt)) < 1e-15


@pytest.mark.parametrize("batched", [False, True])
def test_backward_euler_failure_ladder(batched):
    """With one Newton pass allowed no step converges (the test needs a
    second pass): five reductions of H, then the step is taken unconverged
    (fastode/solvers.py:462-474).  Counts and state as the oracle's."""
    import fastode_amd as fa
    import fastode_oracle as O
    from fastode_amd.system import ChemSystem
    wl = workload("mech20_dense")
    k = wl.k(0, np.arange(32))
    om = O.OracleMechanism(wl.mech.species, wl.mech.reactions_list())
    oc = O.Counts()
    Yo, ost = O.BackwardEuler(max_iterations=1).solve(O.ChemSystem(om, k, dense=True), wl.Y0[:32], wl.dt,
                                                      wl.atol, wl.rtol, oc)
    assert ost == "accepted unconverged" and oc.unconverged == 1 and oc.failures == 6
    c = fa.Counts()
    if batched:
        r = fa.BatchedBackwardEuler(max_iterations=1).runner(ChemSystem(wl.mech, k), wl.Y0[:32], wl.dt, wl.atol,
                                                            wl.rtol, graph=False)
        r.step(None)
        rec = r.memory.read()
        c.add_array(rec["counts"][0])
        from fastode_amd import _lib as L
        st = L.STATUS[int(rec["status"][0])]
        Y = r.Y.cpu().numpy()
    else:
        Y, st = fa.BackwardEuler(max_iterations=1).solve(ChemSystem(wl.mech, k), wl.Y0[:32], wl.dt, wl.atol,
                                                         wl.rtol, c)
    assert st == ost
    assert counts_list(c) == counts_list(oc)
    assert state_error(Y, Yo)[0] <= 1.0


# ------------------------------------------------------------- the boundary
def reference_shaped(mech):
    """An object with exactly the attributes the reference's Mechanism
    builds (fastode/chem.py:19-79): species, n, R, third_body_power, P,
    react, S, ST, pattern (n, rows, cols, nnz), GT (here a placeholder:
    nothing on the device side reads it)."""
    pattern = types.SimpleNamespace(n=mech.n, rows=np.asarray(mech.rows), cols=np.asarray(mech.cols),
                                    nnz=len(mech.rows))
    return types.SimpleNamespace(species=list(mech.species), n=mech.n, R=mech.R,
                                 third_body_power=np.zeros(mech.R), P=mech.P, react=mech.react.copy(),
                                 S=mech.S.copy(), ST=np.ascontiguousarray(mech.S.T), pattern=pattern, GT=None)


def test_reference_shaped_mechanism_runs_the_mech20_golden_with_numpy_arrays():
    """The mech20 golden vectors (written by the reference) through
    ChemSystem(reference-shaped mechanism) with NumPy Y, k, atol: exact
    Counts, statuses and states to the golden bar; NumPy out."""
    import fastode_amd as fa
    from fastode_amd.system import ChemSystem
    g = load("mech20_dense")
    wl = workload("mech20_dense")
    ref_mech = reference_shaped(wl.mech)
    alg = fa.Algorithm(fa.ROS3)
    Y = g["Y0"].copy()
    mems = [alg.new_memory() for _ in range(len(g["groups"]))]
    for it in range(int(g["n_intervals"])):
        for gi in range(len(g["groups"])):
            sl = slice(gi * 32, gi * 32 + 32)
            system = ChemSystem(ref_mech, wl.k(it, g["cells"][sl]))
            c = fa.Counts()
            Yn, st = alg.solve(system, Y[sl], wl.dt, wl.atol, wl.rtol, c, mems[gi])
            assert isinstance(Yn, np.ndarray)
            assert st == g["status"][it, gi]
            assert counts_list(c) == list(g["counts"][it, gi]), (it, gi)
            Y[sl] = Yn
    last = list(g["snap_interval"]).index(int(g["n_intervals"]))
    assert state_error(Y, g["snaps"][last])[0] <= 1.0
    assert getattr(ref_mech, "_fastode_amd_mech", None) is not None   # compiled once, kept


def test_numpy_solver_drives_device_systems():
    """A NumPy solver written as the reference's is (the oracle's
    restatement of fastode/solvers.py:305-415) drives the device ChemSystem
    through rhs / jac / factor / LU.solve on NumPy arrays, and takes the
    golden step sequence of the 200-species workload."""
    import fastode_oracle as O
    from fastode_amd.system import ChemSystem
    g = load("mech200_diurnal")
    wl = workload("mech200_diurnal")
    ref_mech = reference_shaped(wl.mech)
    alg = O.Algorithm(O.ROS3)
    gi = 0
    sl = slice(0, 32)
    Y = g["Y0"][sl].copy()
    mem = alg.new_memory()
    for it in range(3):
        system = ChemSystem(ref_mech, wl.k(it, g["cells"][sl]))
        assert isinstance(system.rhs(Y), np.ndarray)
        c = O.Counts()
        Y, st = alg.solve(system, Y, wl.dt, wl.atol, wl.rtol, c, mem)
        assert st == g["status"][it, gi]
        assert list(c.as_dict().values()) == list(g["counts"][it, gi]), (it, c.as_dict())


def test_dense_system_takes_numpy_and_torch_callables():
    import fastode_amd as fa
    a = decay_np(0.5)
    b = fa.DenseSystem(1, lambda Y: -0.5 * Y,
                       lambda Y: torch.full((Y.shape[0], 1, 1), -0.5, dtype=torch.float64, device=Y.device))
    c = fa.DenseSystem(1, lambda Y: np.asarray(-0.5 * np.asarray(Y)),
                       lambda Y: np.full((np.asarray(Y).shape[0], 1, 1), -0.5))
    outs = []
    for s in (a, b, c):
        cnt = fa.Counts()
        Y, st = fa.Algorithm(fa.ROS3).solve(s, np.array([[1.0], [2.0]]), 1.0, np.array([1e-12]), 1e-6, cnt)
        assert st == "converged" and isinstance(Y, np.ndarray)
        outs.append((Y, counts_list(cnt)))
    assert all(np.array_equal(outs[0][0], o[0]) and outs[0][1] == o[1] for o in outs[1:])
    assert abs(outs[0][0][0, 0] - np.exp(-0.5)) < 1e-5
