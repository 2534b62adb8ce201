"""GPU: the plain drop-in ``Algorithm.solve`` against the reference's golden
vectors (exact Counts and step sequence, states within 1e-9 relative above
the floor 1e-17) and the reference's own known-answer tests, restated on
device systems."""
import math
import time

import numpy as np
import pytest
import torch

from golden_util import COUNT_NAMES, GOLDEN, load, state_error, workload

pytestmark = pytest.mark.gpu



def dev(x):
    return torch.as_tensor(x, dtype=torch.float64, device="cuda")


def replay_plain(name, groups=None, n_intervals=None):
    import fastode_amd as fa
    from fastode_amd.system import ChemSystem
    g = load(name)
    wl = workload(name)
    tab = fa.ROS3 if wl.tableau == "ros3" else fa.RODAS3
    alg = fa.Algorithm(tab)
    groups = list(range(len(g["groups"]))) if groups is None else groups
    nint = int(g["n_intervals"]) if n_intervals is None else n_intervals
    Y = g["Y0"].copy()
    mems = {gi: alg.new_memory() for gi in groups}
    worst = 0.0
    near_ties = 0
    for it in range(nint):
        for gi in groups:
            sl = slice(gi * 32, gi * 32 + 32)
            k = wl.k(it, g["cells"][sl])
            system = ChemSystem(wl.mech, k, mass=wl.mass)
            c = fa.Counts()
            alg.trace = []
            Yn, st = alg.solve(system, Y[sl], wl.dt, wl.atol, wl.rtol, c, mems[gi])
            Y[sl] = Yn
            sel = (g["att_group"] == gi) & (g["att_interval"] == it)
            e_ref = g["att_err"][sel]
            near_ties += int(np.sum(np.abs(e_ref - 1.0) < 1e-10))
            assert st == g["status"][it, gi], (it, gi, st)
            assert [getattr(c, f) for f in COUNT_NAMES] == list(g["counts"][it, gi]), \
                (it, gi, c.as_dict(), list(g["counts"][it, gi]))
            H_ref = 1.0 / (g["att_alpha"][sel] * tab.gamma)
            H = np.array([t[0] for t in alg.trace])
            e = np.array([t[1] for t in alg.trace])
            acc = np.array([t[2] for t in alg.trace])
            assert len(H) == len(H_ref)
            assert np.allclose(H, H_ref, rtol=1e-12, atol=0), (it, gi)
            assert np.array_equal(acc, e_ref < 1.0)
            # the embedded estimate is a difference of stage vectors: its
            # relative agreement is looser than the states'
            assert np.allclose(e, e_ref, rtol=1e-4, atol=0), (it, gi, np.abs(e / e_ref - 1).max())
        if (it + 1) in list(g["snap_interval"]):
            s = list(g["snap_interval"]).index(it + 1)
            for gi in groups:
                sl = slice(gi * 32, gi * 32 + 32)
                ref = g["snaps"][s][sl]
                scaled, rel = state_error(Y[sl], ref, wl.atol)
                print(f"{name} interval {it + 1} group {gi}: scaled error {scaled:.3e}, "
                      f"relative above the floor {rel:.3e}")
                worst = max(worst, scaled)
    assert near_ties == 0, "an attempt's reference error lies within 1e-10 of 1"
    return worst


def test_plain_matches_golden_mech20():
    worst = replay_plain("mech20_dense")
    print("mech20 worst rel", worst)
    assert worst <= 1.0


def test_plain_matches_golden_mech200_diurnal():
    t = time.time()
    worst = replay_plain("mech200_diurnal")
    print(f"mech200 diurnal worst rel {worst:.3e} in {time.time() - t:.1f} s")
    assert worst <= 1.0


@pytest.mark.parametrize("name", ["mech200_uneven", "mech100_constrained"])
def test_plain_matches_golden_other(name):
    if not (GOLDEN / f"{name}.npz").exists():
        pytest.skip(f"{name} golden vectors not generated")
    worst = replay_plain(name)
    print(name, "worst rel", worst)
    assert worst <= 1.0


# ---- the reference's own known-answer tests (tests/test_fastode.py:44-207),
# restated with device systems whose rhs/jac are torch callables.

def decay(rate):
    from fastode_amd import DenseSystem
    return DenseSystem(1, lambda Y: -rate * Y,
                       lambda Y: torch.full((Y.shape[0], 1, 1), -rate, dtype=torch.float64, device=Y.device))


@pytest.mark.parametrize("tabname", ["ROS3", "RODAS3"])
def test_order_three_on_a_nonlinear_ode(tabname):
    import fastode_amd as fa
    from fastode_amd.solvers import fixed_steps
    tab = getattr(fa, tabname)
    sys_ = fa.DenseSystem(1, lambda Y: -Y ** 3, lambda Y: (-3 * Y ** 2)[:, :, None])
    exact = 1.0 / math.sqrt(3.0)
    errs = [abs(float(fixed_steps(sys_, tab, dev([[1.0]]), 1.0, n)[0, 0]) - exact) for n in (20, 40, 80)]
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert min(orders) > 2.8


def test_robertson_linear_invariant():
    import fastode_amd as fa
    from fastode_amd.mechanism import Mechanism
    from fastode_amd.solvers import fixed_steps
    mech = Mechanism(["A", "B", "C"], [(["A"], [("B", 1.0)]), (["B", "B"], [("B", 1.0), ("C", 1.0)]),
                                       (["B", "C"], [("A", 1.0), ("C", 1.0)])])
    system = fa.ChemSystem(mech, np.array([[0.04, 3.0e7, 1.0e4]]))
    for tab in (fa.ROS3, fa.RODAS3):
        Y = fixed_steps(system, tab, dev([[1.0, 0.0, 0.0]]), 1.0e-2, 50)
        assert abs(float(Y.sum()) - 1.0) < 1e-14


def test_carried_memory_takes_one_step():
    import fastode_amd as fa
    sys_ = decay(1e-6)
    alg = fa.Algorithm(fa.ROS3)
    mem = alg.new_memory()
    Y = np.array([[1.0]])
    c = fa.Counts()
    for _ in range(5):
        Y, _ = alg.solve(sys_, Y, 600.0, np.array([1e-10]), 1e-3, c, mem)
    c3 = fa.Counts()
    alg.solve(sys_, Y, 600.0, np.array([1e-10]), 1e-3, c3, mem)
    assert c3.accepted == 1


def test_quantized_steps_land_on_the_interval_end():
    import fastode_amd as fa
    sys_ = decay(1e-9)
    alg = fa.Algorithm(fa.ROS3, fac_max=1.0)
    mem = alg.new_memory()
    mem.h_next = 600.0 / 7
    Y = np.array([[1.0]])
    for _ in range(3):
        c = fa.Counts()
        Y, st = alg.solve(sys_, Y, 600.0, np.array([1e-10]), 1e-3, c, mem)
        assert st == "converged" and c.steps == 7
        assert abs(mem.h_next - 600.0 / 7) < 1e-9


def test_a_rejection_shrinks_the_step():
    import fastode_amd as fa
    sys_ = decay(1.0)
    alg = fa.Algorithm(fa.ROS3)
    mem = alg.new_memory()
    mem.h_next, mem.err_prev, mem.h_prev = 1.0, 1e-8, 1.0
    c = fa.Counts()
    alg.solve(sys_, np.array([[1.0]]), 1.0, np.array([1e-12]), 4e-3, c, mem)
    assert c.rejected == 1


def test_no_forced_acceptance_below_minimum_step():
    import fastode_amd as fa
    sys_ = decay(1e15)
    alg = fa.Algorithm(fa.ROS3, max_steps=10)
    mem = alg.new_memory()
    mem.h_next = 1e-15
    c = fa.Counts()
    Y, st = alg.solve(sys_, np.array([[1.0]]), 1.0, np.array([1e-30]), 1e-12, c, mem)
    assert st == "step too small"
    assert c.steps == c.rejected == 1 and c.accepted == 0
    assert Y[0, 0] == 1.0


@pytest.mark.parametrize("budget", [0, 1, 2])
def test_attempt_budget_counts_rejections(budget):
    import fastode_amd as fa
    sys_ = decay(1.0)
    alg = fa.Algorithm(fa.ROS3, max_steps=budget)
    mem = alg.new_memory()
    mem.h_next = 1.0
    c = fa.Counts()
    Y, st = alg.solve(sys_, np.array([[1.0]]), 1.0, np.array([1e-30]), 1e-13, c, mem)
    assert st == "too many steps"
    assert c.steps == c.rejected == budget and c.accepted == 0
    assert Y[0, 0] == 1.0


@pytest.mark.parametrize("budget, expected", [(2, "too many steps"), (10, "converged")])
def test_attempt_budget_counts_accepted(budget, expected):
    import fastode_amd as fa
    sys_ = decay(1e-6)
    alg = fa.Algorithm(fa.ROS3, max_steps=budget, fac_max=1.0)
    mem = alg.new_memory()
    mem.h_next = 0.1
    c = fa.Counts()
    _, st = alg.solve(sys_, np.array([[1.0]]), 1.0, np.array([1e-10]), 1e-3, c, mem)
    assert st == expected
    assert c.steps == c.accepted == budget and c.rejected == 0


def test_nan_rhs_returns_nan_or_inf():
    import fastode_amd as fa
    from fastode_amd import DenseSystem
    sys_ = DenseSystem(1, lambda Y: Y * float("nan"),
                       lambda Y: torch.full((Y.shape[0], 1, 1), -1.0, dtype=torch.float64, device=Y.device))
    alg = fa.Algorithm(fa.ROS3)
    c = fa.Counts()
    Y, st = alg.solve(sys_, np.array([[1.0]]), 1.0, np.array([1e-10]), 1e-3, c)
    assert st == "nan or inf" and c.steps == 1 and c.accepted == 0
    assert Y[0, 0] == 1.0


def test_oracle_and_plain_agree_on_a_constrained_dae():
    """The triples DAE with RODAS3 (fastode/problems.py:140-181): the device
    run takes exactly the oracle's steps over 100 intervals."""
    import fastode_amd as fa
    import fastode_oracle as O
    N = 1
    mass = np.tile([0.0, 0.0, 1.0], N)

    def rhs_np(Y):
        F = np.empty_like(Y)
        A, B, C = Y[:, 0::3], Y[:, 1::3], Y[:, 2::3]
        F[:, 0::3] = A - B
        F[:, 1::3] = A + B + C - 1.0
        F[:, 2::3] = 1.0 * B
        return F

    def jac_np(Y):
        J = np.zeros((Y.shape[0], 3, 3))
        J[:, 0, 0] = 1.0; J[:, 0, 1] = -1.0
        J[:, 1, 0] = 1.0; J[:, 1, 1] = 1.0; J[:, 1, 2] = 1.0
        J[:, 2, 1] = 1.0
        return J

    osys = O.DenseSystem(3, rhs_np, jac_np, mass=mass)
    dsys = fa.DenseSystem(3, lambda Y: dev(rhs_np(Y.cpu().numpy())), lambda Y: dev(jac_np(Y.cpu().numpy())),
                          mass=mass)
    oalg, alg = O.Algorithm(O.RODAS3), fa.Algorithm(fa.RODAS3)
    om, m = oalg.new_memory(), alg.new_memory()
    Yo = Y = np.tile([0.5, 0.5, 0.0], (1, N))
    oc, c = O.Counts(), fa.Counts()
    for _ in range(100):
        Yo, sto = oalg.solve(osys, Yo, 0.2, np.full(3, 1e-18), 1e-4, oc, om)
        Y, st = alg.solve(dsys, Y, 0.2, np.full(3, 1e-18), 1e-4, c, m)
        assert st == sto == "converged"
    # The step sequence is exact.  project's Newton work is not: its exit test
    # is an exactly zero algebraic residual (fastode/solvers.py:132), which a
    # rounding difference of 1e-18 in the residual turns into one more pass.
    for f in ("steps", "accepted", "rejected", "rhs", "proj_passes", "proj_trials", "calls"):
        assert getattr(c, f) == getattr(oc, f), f
    assert c.steps < 250
    assert np.allclose(Y, Yo, rtol=1e-9, atol=1e-17)


@pytest.mark.parametrize("case", ["backtracks", "no_root", "already_consistent"])
def test_project_newton_and_backtracking_match_the_oracle(case):
    """project (fastode/solvers.py:100-178) on the device against the oracle's
    restatement, on a DenseSystem whose algebraic row is atan(y1 - y0) (a full
    Newton step from y1 - y0 = 3 overshoots, so the line search halves it),
    exp(y1 - y0) (no root: the projection fails and restores Y), or
    an already consistent state (no Newton pass).  Status, every counter and
    the state agree; four cells take different starting offsets."""
    import fastode_amd as fa
    import fastode_oracle as O
    from fastode_amd.solvers import project

    mass = np.array([1.0, 0.0])

    def g_of(x):
        return np.exp(x) if case == "no_root" else np.arctan(x)

    def dg_of(x):
        return np.exp(x) if case == "no_root" else 1.0 / (1.0 + x * x)

    def rhs_np(Y):
        F = np.empty_like(Y)
        F[:, 0] = -Y[:, 0]
        F[:, 1] = g_of(Y[:, 1] - Y[:, 0])
        return F

    def jac_np(Y):
        x = Y[:, 1] - Y[:, 0]
        J = np.zeros((Y.shape[0], 2, 2))
        J[:, 0, 0] = -1.0
        J[:, 1, 0] = -dg_of(x)
        J[:, 1, 1] = dg_of(x)
        return J

    off = {"backtracks": [3.0, 2.5, -3.5, 1.2], "no_root": [0.5, 1.0, -2.0, 3.0],
           "already_consistent": [0.0, 0.0, 0.0, 0.0]}[case]
    Y0 = np.stack([np.full(4, 0.7), 0.7 + np.array(off)], axis=1)
    atol, rtol = np.full(2, 1e-12), 1e-6
    osys = O.DenseSystem(2, rhs_np, jac_np, mass=mass)
    dsys = fa.DenseSystem(2, lambda Y: dev(rhs_np(Y.cpu().numpy())), lambda Y: dev(jac_np(Y.cpu().numpy())),
                          mass=mass)
    oc, c = O.Counts(), fa.Counts()
    Yo, sto = O.project(osys, Y0.copy(), atol, rtol, oc)
    Yd, st = project(dsys, dev(Y0), atol, rtol, c)
    print(case, st, c.as_dict())
    assert st == sto
    assert c.as_dict() == oc.as_dict()
    Yn, stn = project(dsys, Y0.copy(), atol, rtol, fa.Counts())     # NumPy in, NumPy out
    assert stn == st and isinstance(Yn, np.ndarray) and np.array_equal(Yn, Yd.cpu().numpy())
    if case == "backtracks":
        assert st == "converged" and c.proj_trials > c.proj_passes   # the line search halved a step
        assert np.allclose(Yd.cpu().numpy(), Yo, rtol=1e-12, atol=1e-15)
    elif case == "no_root":
        assert st == "projection failed"
        assert np.array_equal(Yd.cpu().numpy(), Y0)                  # restored
    else:
        assert st == "converged" and c.proj_passes == 1 and c.jac == 0
        assert np.array_equal(Yd.cpu().numpy(), Y0)
