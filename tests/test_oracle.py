"""CPU checks that pin the oracle (and the host tableaus) to the reference:
closed forms, invariants, the reference's own known-answer tests
(tests/test_fastode.py of the reference, restated) and the golden vectors the
reference wrote for the seeded workloads."""
import math

import numpy as np
import pytest

import fastode_oracle as O
from golden_util import COUNT_NAMES, load, workload

import fastode_amd.tableaus as T


def robertson_oracle():
    # fastode/problems.py:16-24
    mech = O.OracleMechanism(["A", "B", "C"], [
        (["A"], [("B", 1.0)]),
        (["B", "B"], [("B", 1.0), ("C", 1.0)]),
        (["B", "C"], [("A", 1.0), ("C", 1.0)]),
    ])
    k = np.array([[0.04, 3.0e7, 1.0e4]])
    return mech, k


@pytest.mark.parametrize("tab", [O.ROS3, O.RODAS3])
def test_order_conditions(tab):
    A, Gam, b, bhat = tab.standard_form()
    beta = A + Gam
    ap, bp = A.sum(axis=1), beta.sum(axis=1)
    assert abs(b.sum() - 1) < 1e-12
    assert abs(b @ bp - 0.5) < 1e-12
    assert abs(b @ ap ** 2 - 1 / 3) < 1e-12
    assert abs(b @ (beta @ bp) - 1 / 6) < 1e-12
    assert abs(bhat.sum() - 1) < 1e-12 and abs(bhat @ bp - 0.5) < 1e-12


@pytest.mark.parametrize("mod", [O, T])
def test_index1_and_forcing_flags(mod):
    if mod is O:   # the package's tableaus equal the oracle's (test_host_tableaus_match_oracle)
        assert abs(mod.RODAS3.index1(2) - 1) < 1e-12 and abs(mod.RODAS3.index1(3) - 1) < 1e-12
        assert abs(mod.ROS3.index1(2) - 1.0908767) < 1e-6
    assert mod.ROS3.new_forcing == [True, True, False]
    assert mod.RODAS3.new_forcing == [True, False, True, True]


def test_host_tableaus_match_oracle():
    for a, b in ((T.ROS3, O.ROS3), (T.RODAS3, O.RODAS3)):
        assert a.gamma == b.gamma
        for f in ("a", "c", "m", "e"):
            assert np.array_equal(getattr(a, f), getattr(b, f))


def test_stability_function_is_l_stable_for_rodas3():
    # R(z) = 1 + z b^T (I - z B)^{-1} 1 of the standard form; RODAS3 is
    # stiffly accurate (R(inf) = 0), ROS3 has |R(inf)| < 1.
    for tab, rinf in ((O.RODAS3, 0.0), (O.ROS3, None)):
        A, Gam, b, _ = tab.standard_form()
        Bm = A + Gam
        z = -1e12
        R = 1 + z * b @ np.linalg.solve(np.eye(tab.stages) - z * Bm, np.ones(tab.stages))
        if rinf is not None:
            assert abs(R - rinf) < 1e-6
        assert abs(R) < 1.0


def test_quantize_divides_the_interval():
    for h in (1.0, 7.3, 33.0, 600.0, 900.0, 600.0 / 7, 1e-9):
        q = O.quantize(h, 600.0)
        n = round(600.0 / q)
        assert abs(n * q - 600.0) < 1e-9
        assert h >= 600.0 or h / 2 < q <= h


def test_error_norm_closed_form():
    Y = np.array([[1.0, -2.0], [0.0, 4.0]])
    Yn = np.array([[2.0, 1.0], [0.0, -1.0]])
    E = np.array([[1e-3, 2e-3], [0.0, 4e-3]])
    atol = np.array([1e-6, 1e-5])
    s = [1e-3 / (1e-6 + 1e-2 * 2), 2e-3 / (1e-5 + 1e-2 * 2), 0.0, 4e-3 / (1e-5 + 1e-2 * 4)]
    want = math.sqrt(sum(x * x for x in s) / 4)
    assert abs(O.error_norm(Y, Yn, E, atol, 1e-2) - want) < 1e-15 * want
    assert O.error_norm(Y, Y, np.zeros_like(Y), atol, 1e-2) == 1e-10
    assert math.isnan(O.error_norm(Y, Y, E * np.nan, atol, 1e-2))


def test_controller_factor_closed_form():
    alg = O.Algorithm(O.ROS3)
    mem = O.Memory()
    assert alg.factor(0.5, 10.0, mem) == min(6.0, max(0.2, 0.9 / 0.5 ** (1 / 3)))
    mem.err_prev, mem.h_prev = 0.25, 5.0
    want = 0.9 * (0.5 * 0.25) ** (-1 / 12) * (10.0 / 5.0) ** (-1 / 4)
    assert alg.factor(0.5, 10.0, mem) == min(6.0, max(0.2, want))
    assert alg.factor(1e-12, 1.0, O.Memory()) == 6.0


@pytest.mark.parametrize("tab", [O.ROS3, O.RODAS3])
def test_order_three_on_a_nonlinear_ode(tab):
    sys_ = O.DenseSystem(1, lambda Y: -Y ** 3, lambda Y: (-3 * Y ** 2)[:, :, None])
    exact = 1.0 / math.sqrt(3.0)
    errs = [abs(O.fixed_steps(sys_, tab, np.array([[1.0]]), 1.0, n)[0, 0] - exact) for n in (20, 40, 80)]
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert min(orders) > 2.8


def test_robertson_rhs_and_invariant():
    mech, k = robertson_oracle()
    Y = np.array([[0.7, 2e-5, 0.3]])
    F = mech.rhs(k, Y)
    A, B, C = Y[0]
    want = [-0.04 * A + 1e4 * B * C, 0.04 * A - 3e7 * B * B - 1e4 * B * C, 3e7 * B * B]
    assert np.allclose(F[0], want, rtol=1e-14, atol=0)
    # dense Jacobian against central differences
    J = mech.dense(mech.jac_values(k, Y))[0]
    for j in range(3):
        h = 1e-7 * max(1e-5, abs(Y[0, j]))
        d = np.zeros_like(Y)
        d[0, j] = h
        fd = (mech.rhs(k, Y + d) - mech.rhs(k, Y - d))[0] / (2 * h)
        assert np.allclose(J[:, j], fd, rtol=1e-6, atol=1e-6 * np.abs(fd).max())
    system = O.ChemSystem(mech, k)
    for tab in (O.ROS3, O.RODAS3):
        Yf = O.fixed_steps(system, tab, np.array([[1.0, 0.0, 0.0]]), 1.0e-2, 50)
        assert abs(Yf.sum() - 1.0) < 1e-14


def decay(rate):
    return O.DenseSystem(1, lambda Y: -rate * Y, lambda Y: np.full((Y.shape[0], 1, 1), -rate))


def test_carried_memory_takes_one_step():
    sys_ = decay(1e-6)
    alg = O.Algorithm(O.ROS3)
    mem = alg.new_memory()
    Y = np.array([[1.0]])
    c = O.Counts()
    for _ in range(5):
        Y, _ = alg.solve(sys_, Y, 600.0, np.array([1e-10]), 1e-3, c, mem)
    c3 = O.Counts()
    alg.solve(sys_, Y, 600.0, np.array([1e-10]), 1e-3, c3, mem)
    assert c3.accepted == 1


def test_quantized_steps_land_on_the_interval_end():
    sys_ = decay(1e-9)
    alg = O.Algorithm(O.ROS3, fac_max=1.0)
    mem = alg.new_memory()
    mem.h_next = 600.0 / 7
    Y = np.array([[1.0]])
    for _ in range(3):
        c = O.Counts()
        Y, st = alg.solve(sys_, Y, 600.0, np.array([1e-10]), 1e-3, c, mem)
        assert st == "converged" and c.steps == 7
        assert abs(mem.h_next - 600.0 / 7) < 1e-9


def test_a_rejection_shrinks_the_step():
    sys_ = decay(1.0)
    alg = O.Algorithm(O.ROS3)
    mem = alg.new_memory()
    mem.h_next, mem.err_prev, mem.h_prev = 1.0, 1e-8, 1.0
    c = O.Counts()
    alg.solve(sys_, np.array([[1.0]]), 1.0, np.array([1e-12]), 4e-3, c, mem)
    assert c.rejected == 1


def test_no_forced_acceptance_below_minimum_step():
    sys_ = decay(1e15)
    alg = O.Algorithm(O.ROS3, max_steps=10)
    mem = alg.new_memory()
    mem.h_next = 1e-15
    c = O.Counts()
    Y, st = alg.solve(sys_, np.array([[1.0]]), 1.0, np.array([1e-30]), 1e-12, c, mem)
    assert st == "step too small"
    assert c.steps == c.rejected == 1 and c.accepted == 0
    assert Y[0, 0] == 1.0


@pytest.mark.parametrize("budget", [0, 1, 2])
def test_attempt_budget_counts_rejections(budget):
    sys_ = decay(1.0)
    alg = O.Algorithm(O.ROS3, max_steps=budget)
    mem = alg.new_memory()
    mem.h_next = 1.0
    c = O.Counts()
    Y, st = alg.solve(sys_, np.array([[1.0]]), 1.0, np.array([1e-30]), 1e-13, c, mem)
    assert st == "too many steps"
    assert c.steps == c.rejected == budget and c.accepted == 0


@pytest.mark.parametrize("budget, expected", [(2, "too many steps"), (10, "converged")])
def test_attempt_budget_counts_accepted(budget, expected):
    sys_ = decay(1e-6)
    alg = O.Algorithm(O.ROS3, max_steps=budget, fac_max=1.0)
    mem = alg.new_memory()
    mem.h_next = 0.1
    c = O.Counts()
    _, st = alg.solve(sys_, np.array([[1.0]]), 1.0, np.array([1e-10]), 1e-3, c, mem)
    assert st == expected
    assert c.steps == c.accepted == budget and c.rejected == 0


def _constraints(mech):
    return {mech.species[r]: ({mech.species[s]: mech.constraint_coef[i, s]
                               for s in np.nonzero(mech.constraint_coef[i])[0]},
                              mech.constraint_total[i])
            for i, r in enumerate(mech.constraint_rows)} or None


def run_oracle_on_golden(name, groups=None, n_intervals=None):
    """Replay a golden workload through the oracle; return the worst relative
    state difference above the floor after asserting Counts and traces."""
    g = load(name)
    wl = workload(name)
    assert np.array_equal(g["react"], wl.mech.react) and np.array_equal(g["S"], wl.mech.S)
    assert np.array_equal(g["Y0"], wl.Y0[g["cells"]])
    om = O.OracleMechanism(wl.mech.species, wl.mech.reactions_list(), _constraints(wl.mech))
    tab = O.ROS3 if wl.tableau == "ros3" else O.RODAS3
    alg = O.Algorithm(tab)
    groups = list(range(len(g["groups"]))) if groups is None else groups
    nint = int(g["n_intervals"]) if n_intervals is None else n_intervals
    Y = g["Y0"].copy()
    mems = {gi: O.Memory() for gi in groups}
    worst = 0.0
    for it in range(nint):
        for gi in groups:
            sl = slice(gi * 32, gi * 32 + 32)
            k = wl.k(it, g["cells"][sl])
            dense = wl.dense or wl.mech.has_constraints
            system = O.ChemSystem(om, k, mass=wl.mass, dense=dense)
            c = O.Counts()
            alg.trace = []
            Yn, st = alg.solve(system, Y[sl], wl.dt, wl.atol, wl.rtol, c, mems[gi])
            Y[sl] = Yn
            assert st == g["status"][it, gi]
            assert [getattr(c, f) for f in COUNT_NAMES] == list(g["counts"][it, gi]), (it, gi)
            sel = (g["att_group"] == gi) & (g["att_interval"] == it)
            H_ref = 1.0 / (g["att_alpha"][sel] * tab.gamma)
            H = np.array([t[0] for t in alg.trace])
            e = np.array([t[1] for t in alg.trace])
            assert len(H) == len(H_ref)
            assert np.allclose(H, H_ref, rtol=1e-12, atol=0)
            assert np.allclose(e, g["att_err"][sel], rtol=1e-6, atol=0)
        if (it + 1) in list(g["snap_interval"]):
            s = list(g["snap_interval"]).index(it + 1)
            for gi in groups:
                sl = slice(gi * 32, gi * 32 + 32)
                ref = g["snaps"][s][sl]
                big = np.abs(ref) > 1e-17
                worst = max(worst, float((np.abs(Y[sl] - ref)[big] / np.abs(ref)[big]).max()))
    return worst


def test_oracle_reproduces_golden_mech20():
    assert run_oracle_on_golden("mech20_dense") < 1e-12


def test_oracle_reproduces_golden_mech200_first_intervals():
    run_oracle_on_golden("mech200_diurnal", groups=[0], n_intervals=4)


@pytest.mark.slow
def test_oracle_reproduces_golden_mech200_full():
    assert run_oracle_on_golden("mech200_diurnal") < 1e-12
