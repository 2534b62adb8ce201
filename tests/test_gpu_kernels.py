"""GPU: the HIP kernels against the oracle and NumPy, through the C ABI."""
import math

import numpy as np
import pytest
import torch

import fastode_oracle as O
from fastode_amd import workloads as W
from fastode_amd.mechanism import synthetic

pytestmark = pytest.mark.gpu


def dev(x):
    return torch.as_tensor(x, dtype=torch.float64, device="cuda").contiguous()


def oracle_mech(mech):
    cons = {mech.species[r]: ({mech.species[s]: mech.constraint_coef[i, s]
                               for s in np.nonzero(mech.constraint_coef[i])[0]}, mech.constraint_total[i])
            for i, r in enumerate(mech.constraint_rows)} or None
    return O.OracleMechanism(mech.species, mech.reactions_list(), cons)


def test_device_info():
    import fastode_amd._lib as L
    info = L.device_info(0)
    print("device:", info, torch.cuda.get_device_name(0))
    assert info["warp_size"] == 64 and info["cus"] > 0 and info["lds_per_block"] >= 65536


@pytest.mark.parametrize("which", ["small", "diurnal200", "constrained100"])
def test_rhs_and_jacobian_match_oracle(which):
    from fastode_amd.system import ChemSystem
    wl = {"small": lambda: W.small(C=64), "diurnal200": lambda: W.diurnal200(C=64),
          "constrained100": lambda: W.constrained100(C=64)}[which]()
    om = oracle_mech(wl.mech)
    k = wl.k(0)
    system = ChemSystem(wl.mech, k, mass=wl.mass)
    Y = wl.Y0
    F = system.rhs(dev(Y)).cpu().numpy()
    Fo = om.rhs(k, Y)
    chem = ~wl.mech.is_constraint
    scale = np.abs(Fo[:, chem]).max(axis=1, keepdims=True)
    assert (np.abs(F - Fo)[:, chem] <= 1e-12 * scale + 1e-300).all()
    # conservation rows: sum_s y_s - total cancels to rounding of the total
    con = wl.mech.constraint_rows
    if len(con):
        assert (np.abs(F - Fo)[:, con] <= 8 * np.finfo(float).eps * wl.mech.constraint_total).all()
    J = system.jac(dev(Y)).cpu().numpy()
    Jo = om.jac_values(k, Y)
    assert J.shape == Jo.shape
    assert np.allclose(J, Jo, rtol=1e-13, atol=0.0)


def test_rhs_and_jacobian_of_a_mechanism_beyond_the_cell_lds():
    """40 species and 8,200 reactions: the rates and the rate partials exceed
    the 64 KB a per-cell workgroup holds, so the per-row and per-entry
    kernels evaluate (rows of about 800 terms)."""
    from fastode_amd.mechanism import log_uniform_state
    from fastode_amd.system import ChemSystem
    syn = synthetic(40, 8200, 3)
    assert (40 + 8200 + 1) * 8 > 65536
    rng = np.random.default_rng(3)
    C = 6
    Y = log_uniform_state(C, 40, rng)
    k = syn.k * rng.uniform(0.5, 2.0, size=(C, 8200))
    system = ChemSystem(syn.mech, k)
    om = oracle_mech(syn.mech)
    F = system.rhs(dev(Y)).cpu().numpy()
    Fo = om.rhs(k, Y)
    # a row's rounding is bounded by a multiple of eps times the sum of its
    # terms' magnitudes (rates are positive here)
    terms = (k * om._gather(Y).prod(axis=2)) @ np.abs(om.ST)
    assert (np.abs(F - Fo) <= 1e-12 * terms + 1e-300).all()
    J = system.jac(dev(Y)).cpu().numpy()
    Jo = om.jac_values(k, Y)
    assert np.allclose(J, Jo, rtol=1e-12, atol=0.0)


@pytest.mark.parametrize("n,R,seed,alpha", [(20, 40, 1, 1e-2), (100, 250, 4, 1e-3),
                                              (200, 500, 2, 1e-3), (200, 500, 2, 10.0)])
def test_factor_and_solve_match_numpy(n, R, seed, alpha):
    """alpha*I - J of a mass-action Jacobian at a log-uniform state, factored
    without pivoting on the device, solved and compared with LAPACK."""
    from fastode_amd.mechanism import log_uniform_state
    from fastode_amd.system import ChemSystem
    syn = synthetic(n, R, seed)
    mech = syn.mech
    rng = np.random.default_rng(seed)
    C = 8
    Y = log_uniform_state(C, n, rng)
    k = np.broadcast_to(syn.k, (C, R)).copy()
    system = ChemSystem(mech, k)
    J = system.jac(dev(Y))
    LU = system.factor(alpha, J)
    B = rng.standard_normal((C, n))
    X = LU.solve(dev(B)).cpu().numpy()
    om = oracle_mech(mech)
    Jd = om.dense(J.cpu().numpy())
    for c in range(C):
        A = alpha * np.eye(n) - Jd[c]
        xr = np.linalg.solve(A, B[c])
        res = np.abs(A @ X[c] - B[c]).max() / (np.abs(A).max() * np.abs(X[c]).max())
        assert res < 1e-13, (c, res)
        assert np.abs(X[c] - xr).max() <= 1e-8 * np.abs(xr).max()


@pytest.mark.parametrize("n", [1, 3, 17, 40])
def test_dense_system_factor(n):
    from fastode_amd.system import DenseSystem
    rng = np.random.default_rng(n)
    C = 5
    Jm = rng.standard_normal((C, n, n))
    system = DenseSystem(n, lambda Y: Y, lambda Y: dev(Jm))
    alpha = 3.0 * n + 1.0
    LU = system.factor(alpha, dev(Jm))
    B = rng.standard_normal((C, n))
    X = LU.solve(dev(B)).cpu().numpy()
    for c in range(C):
        xr = np.linalg.solve(alpha * np.eye(n) - Jm[c], B[c])
        assert np.allclose(X[c], xr, rtol=1e-12, atol=1e-14)


def test_algebraic_rows_only_factor():
    from fastode_amd.system import DenseSystem
    n, C = 4, 3
    rng = np.random.default_rng(5)
    Jm = rng.standard_normal((C, n, n))
    mass = np.array([1.0, 0.0, 1.0, 0.0])
    system = DenseSystem(n, lambda Y: Y, lambda Y: dev(Jm), mass=mass)
    LU = system.factor(1.0, dev(Jm), algebraic_rows_only=True)
    B = rng.standard_normal((C, n))
    X = LU.solve(dev(B)).cpu().numpy()
    A = O.stage_matrix_dense(1.0, mass, Jm, algebraic_rows_only=True)
    for c in range(C):
        assert np.allclose(X[c], np.linalg.solve(A[c], B[c]), rtol=1e-12, atol=1e-14)


def test_stage_kernels_match_the_reference_formulas():
    import ctypes as ct
    import fastode_amd._lib as L
    from fastode_amd import Algorithm, Memory, ROS3, RODAS3
    from fastode_amd.solvers import GroupControl, ctrl_params
    rng = np.random.default_rng(3)
    C, n = 32, 7
    for tab in (ROS3, RODAS3):
        p = ctrl_params(Algorithm(tab), tab, False)
        H = 0.37
        ctrl = GroupControl(1.0, Memory())
        ctrl.dev.view(torch.float64)[1] = H
        Y = rng.standard_normal((C, n))
        F = rng.standard_normal((C, n))
        K = [rng.standard_normal((C, n)) for _ in range(tab.stages)]
        Kd = [dev(x) for x in K]
        mass = np.array([1, 1, 0, 1, 1, 1, 0], dtype=float)
        # device copies held for the kernels' lifetime
        Yd, Fd, md = dev(Y), dev(F), dev(mass)
        out = torch.empty((C, n), dtype=torch.float64, device="cuda")
        s = L.stream_handle()
        for r in range(1, tab.stages):
            L.check(L.lib.fode_stage_point(ct.byref(p), r, C, n, L.ptr(Yd), L.ptr_array(Kd[:r]),
                                           L.ptr(out), s))
            want = Y.copy()
            for q in range(r):
                if tab.a[r, q] != 0.0:
                    want += tab.a[r, q] * K[q]
            assert np.array_equal(out.cpu().numpy(), want)
            for dae in (0, 1):
                L.check(L.lib.fode_stage_forcing(ct.byref(p), r, C, n, ctrl.ptr, L.ptr(Fd),
                                                 L.ptr_array(Kd[:r]), L.ptr(md), dae,
                                                 L.ptr(out), s))
                want = F.copy()
                for q in range(r):
                    coef = tab.c[r, q] / H
                    if coef != 0.0:
                        want += coef * (K[q] * mass[None, :] if dae else K[q])
                assert np.array_equal(out.cpu().numpy(), want)
        Ynew = torch.empty_like(out)
        E = torch.empty_like(out)
        L.check(L.lib.fode_stage_final(ct.byref(p), C, n, L.ptr(Yd), L.ptr_array(Kd), L.ptr(Ynew),
                                       L.ptr(E), s))
        wy, we = Y.copy(), np.zeros_like(Y)
        for r in range(tab.stages):
            if tab.m[r] != 0.0:
                wy += tab.m[r] * K[r]
            if tab.e[r] != 0.0:
                we += tab.e[r] * K[r]
        assert np.array_equal(Ynew.cpu().numpy(), wy) and np.array_equal(E.cpu().numpy(), we)
        _ = math


@pytest.mark.parametrize("loop", ["algorithm", "restarting"])
def test_norm_control_per_cell_shares_equal_one_workgroup(loop):
    """fode_norm_control with a work array (one workgroup per cell, the
    shares added in cell order, then the controller and the accept/clamp
    over a grid) against the one-workgroup form: the same control block and
    the same state, bit for bit, across accepting, rejecting and finishing
    attempts.  A restarting loop with a constraint row exercises the
    step-change rows of the error."""
    import ctypes as ct
    import fastode_amd._lib as L
    from fastode_amd import Algorithm, Memory, RestartingLoop, ROS3
    from fastode_amd.solvers import GroupControl, ctrl_params
    rng = np.random.default_rng(11)
    C, n = 40, 300                      # more species than threads, a ragged cell count
    mass = np.ones(n)
    constrained = loop == "restarting"
    if constrained:
        mass[[3, 77]] = 0.0
    solver = RestartingLoop(ROS3, step_change_rows=True) if constrained else Algorithm(ROS3)
    p = ctrl_params(solver, ROS3, constrained)
    atol, rtol = dev(np.full(n, 1e-8)), 1e-3
    md = dev(mass)
    s = L.stream_handle()
    seen = set()
    for trial, scale in enumerate((1e-6, 1e3, 1e-6, 1e-2)):
        Y = rng.uniform(-1e-3, 1.0, (C, n))      # some negative entries for the clamp
        Ynew = Y + scale * rng.standard_normal((C, n))
        E = scale * rng.standard_normal((C, n))
        outs = []
        for use_work in (False, True):
            mem = Memory()
            if trial == 2:
                mem.h_next = 1.0        # the carried step covers the interval: the attempt finishes
            ctrl = GroupControl(1.0, mem)
            L.check(L.lib.fode_ctrl_init(ct.byref(p), ctrl.ptr, 1, 0, s))
            Yd, Ynd, Ed = dev(Y), dev(Ynew), dev(E)
            work = torch.empty(C, dtype=torch.float64, device="cuda") if use_work else None
            L.check(L.lib.fode_norm_control(ct.byref(p), ctrl.ptr, C, n, L.ptr(Yd), L.ptr(Ynd), L.ptr(Ed),
                                            L.ptr(atol), rtol, L.ptr(md), L.ptr(work), s))
            outs.append((ctrl.dev.cpu().numpy().copy(), Yd.cpu().numpy()))
            rec = ctrl.read()
        (g0, y0), (g1, y1) = outs
        assert np.array_equal(g0, g1), (loop, trial)
        assert np.array_equal(y0, y1), (loop, trial)
        seen.add("finish" if rec["action"] == L.FINISH else ("accept" if rec["accepted"] else "reject"))
    assert seen >= ({"accept", "reject", "finish"} if loop == "algorithm" else {"accept", "reject"}), seen


def test_be_update_per_cell_flags_equal_one_workgroup():
    """fode_be_update with a work array (one workgroup per cell writes its
    flags, one thread takes the decision, a grid commits, resets and clamps)
    against the one-workgroup form: the same control block, iterate and step
    start, bit for bit, for a converging pass (committed; the call ends and
    clamps), a pass that misses its bound, and one with a value not finite."""
    import ctypes as ct
    import fastode_amd._lib as L
    from fastode_amd import BackwardEuler, Memory
    from fastode_amd.solvers import GroupControl, ctrl_params
    rng = np.random.default_rng(5)
    C, n = 40, 300
    p = ctrl_params(BackwardEuler(), None, False)
    atol, rtol = dev(np.full(n, 1e-8)), 1e-3
    s = L.stream_handle()
    seen = set()
    for case in ("converge", "miss", "nonfinite"):
        Yn = rng.uniform(-1e-3, 1.0, (C, n))
        Yn1 = Yn + 1e-3 * rng.standard_normal((C, n))
        delta = {"converge": 1e-30, "miss": 1.0, "nonfinite": 1e-30}[case] * rng.standard_normal((C, n))
        if case == "nonfinite":
            delta[7, 11] = np.inf
        forms = []
        for use_work in (False, True):
            ctrl = GroupControl(1.0, Memory())
            L.check(L.lib.fode_ctrl_init(ct.byref(p), ctrl.ptr, 1, 0, s))
            work = torch.empty(2 * C, dtype=torch.float64, device="cuda") if use_work else None
            forms.append((ctrl, dev(Yn), dev(Yn1), work))
        dd = dev(delta)
        for step in range(14):                   # passes of one call, both forms side by side
            outs = []
            for ctrl, Ynd, Yn1d, work in forms:
                if ctrl.read()["action"] == L.FINISH:
                    continue
                L.check(L.lib.fode_be_update(ct.byref(p), ctrl.ptr, C, n, L.ptr(Yn1d), L.ptr(Ynd), L.ptr(dd),
                                             L.ptr(atol), rtol, L.ptr(work), s))
                outs.append((ctrl.dev.cpu().numpy().copy(), Yn1d.cpu().numpy(), Ynd.cpu().numpy()))
                rec = ctrl.read()
            if not outs:
                break
            (g0, a0, b0), (g1, a1, b1) = outs
            assert np.array_equal(g0, g1), (case, step)
            assert np.array_equal(a0, a1) and np.array_equal(b0, b1), (case, step)
            seen.add((int(rec["accepted"]), int(rec["action"])))
    assert len(seen) >= 3, seen
