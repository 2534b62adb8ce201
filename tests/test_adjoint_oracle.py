"""CPU checks that pin the oracle's tangent-linear model, the reference the
GPU adjoint is compared with: complex-step derivatives of the rates and the
Jacobian, central differences of Algorithm.solve itself, and the linearity
of a linear mechanism's flow.  Also the host side of the adjoint: the plan of
the transposed pattern."""
import numpy as np
import pytest

import fastode_oracle as O

import fastode_amd.workloads as W
from fastode_amd.mechanism import synthetic
from fastode_amd.symbolic import build_plan, emulate_factor, emulate_solve


def small_case(C=6, seed=7):
    syn = synthetic(20, 40, seed)
    om = O.OracleMechanism(syn.mech.species, syn.mech.reactions_list())
    rng = np.random.default_rng(seed)
    Y = 10.0 ** rng.uniform(-12, -4, size=(C, 20))
    k = np.broadcast_to(syn.k, (C, 40)).copy()
    return om, k, Y, rng


def test_forcing_tangent_matches_complex_step():
    om, k, Y, rng = small_case()
    dY = Y * rng.normal(size=Y.shape)
    dlogk = rng.normal(size=k.shape)
    h = 1e-30
    Fc = om.rhs(k * np.exp(1j * h * dlogk), Y + 1j * h * dY)
    ref = Fc.imag / h
    got = O.forcing_tangent(om, k, Y, dY, dlogk)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())


def test_jacobian_tangent_matches_complex_step():
    om, k, Y, rng = small_case()
    dY = Y * rng.normal(size=Y.shape)
    dlogk = rng.normal(size=k.shape)
    V = rng.normal(size=Y.shape) * Y
    h = 1e-30
    vals = om.jac_values(k * np.exp(1j * h * dlogk), Y + 1j * h * dY)
    Jc = np.zeros((Y.shape[0], om.n, om.n), dtype=complex)
    Jc[:, om.rows, om.cols] = vals
    ref = np.einsum("cij,cj->ci", Jc.imag / h, V)
    got = O.jacobian_tangent(om, k, Y, dY, dlogk, V)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())


def test_tangent_step_matches_central_differences():
    om, k, Y, rng = small_case()
    dY = Y * rng.normal(size=Y.shape)
    dlogk = rng.normal(size=k.shape)
    H = 5.0
    y1, dy = O.tangent_step(om, k, O.ROS3, Y, dY, dlogk, H)
    eps = 1e-6
    yp, _ = O.tangent_step(om, k * np.exp(eps * dlogk), O.ROS3, Y + eps * dY, dY, dlogk, H)
    ym, _ = O.tangent_step(om, k * np.exp(-eps * dlogk), O.ROS3, Y - eps * dY, dY, dlogk, H)
    fd = (yp - ym) / (2 * eps)
    scale = np.abs(dy).max()
    assert np.abs(fd - dy).max() < 1e-6 * scale
    # the state part is the arithmetic of stages()
    sysm = O.ChemSystem(om, k)
    LU = sysm.factor(1.0 / (H * O.ROS3.gamma), sysm.jac(Y))
    yn, _ = O.stages(sysm, O.ROS3, Y, sysm.rhs(Y), LU, H, O.Counts())
    assert np.array_equal(yn, y1)


def test_linear_mechanism_flow_is_its_own_tangent():
    # All-first-order chemistry: the interval map Y0 -> Y is linear for fixed
    # steps, so the tangent along dY0 = Y0 is the final state itself.
    species = ["A", "B", "C", "D"]
    mech = O.OracleMechanism(species, [
        (["A"], [("B", 1.0)]), (["B"], [("C", 1.0)]), (["C"], [("A", 1.0), ("D", 1.0)]),
        (["D"], [("B", 1.0)])])
    k = np.array([[0.3, 2.0, 50.0, 1e-3]] * 2)
    Y0 = np.array([[1e-3, 2e-4, 3e-6, 1e-9], [4e-5, 1e-3, 2e-7, 3e-4]])
    kmuls = np.ones((3, 4))
    atol = np.full(4, 1e-20)
    Y, dY, steps = O.tangent_linear(mech, k, kmuls, O.ROS3, Y0, Y0.copy(), np.zeros_like(k), 10.0, atol,
                                    1e-4, group_size=1)
    assert all(len(h) > 1 for row in steps for h in row)
    assert np.allclose(dY, Y, rtol=1e-12, atol=0)


def test_tangent_linear_matches_solve_differences():
    """The tangent of whole controlled intervals (replayed accepted steps)
    against central differences of Algorithm.solve on the 20-species
    workload, 2 groups of 4 cells, 3 intervals."""
    wl = W.small(C=8)
    om = O.OracleMechanism(wl.mech.species, wl.mech.reactions_list())
    rng = np.random.default_rng(11)
    k = np.broadcast_to(wl.k_base, (8, wl.R)).copy()
    kmuls = np.stack([wl.kmul(i) for i in range(3)])
    dlogk = rng.normal(size=k.shape)
    dY0 = wl.Y0 * rng.normal(size=wl.Y0.shape) * 0.5
    Y, dY, steps = O.tangent_linear(om, k, kmuls, O.ROS3, wl.Y0, dY0, dlogk, wl.dt, wl.atol, wl.rtol,
                                    group_size=4)

    def run(s):
        Yp, _, st = O.tangent_linear(om, k * np.exp(s * dlogk), kmuls, O.ROS3, wl.Y0 + s * dY0,
                                     np.zeros_like(dY0), np.zeros_like(dlogk), wl.dt, wl.atol, wl.rtol,
                                     group_size=4)
        return Yp, st

    eps = 1e-6
    yp, sp = run(eps)
    ym, sm = run(-eps)
    assert sp == steps and sm == steps          # the step sequence does not move
    fd = (yp - ym) / (2 * eps)
    w = np.random.default_rng(3).uniform(size=wl.n)
    a = 1e-17
    g_tl = float((w * dY / (Y + a)).sum())
    g_fd = float((w * fd / (Y + a)).sum())
    assert abs(g_tl - g_fd) <= 1e-5 * abs(g_tl)
    big = np.abs(Y) > 1e-12
    assert np.abs(fd - dY)[big].max() <= 1e-5 * np.abs(dY[big]).max()


@pytest.mark.parametrize("wl_name", ["mech20", "mech200"])
def test_transposed_plan_factors_the_transpose(wl_name):
    wl = W.small() if wl_name == "mech20" else W.diurnal200(C=32)
    mech = wl.mech
    fwd = build_plan(mech.n, mech.rows, mech.cols)
    tr = build_plan(mech.n, mech.cols, mech.rows, order=fwd.perm, tile=fwd.tile)
    assert (tr.c, tr.m, tr.tile, tr.ns, tr.nsch, tr.nnz_lu) == (fwd.c, fwd.m, fwd.tile, fwd.ns, fwd.nsch,
                                                                  fwd.nnz_lu)
    rng = np.random.default_rng(5)
    vals = rng.normal(size=mech.nnz)
    vals[: mech.n] += 50.0                      # diagonal first: well conditioned
    A = np.zeros((mech.n, mech.n))
    A[mech.rows, mech.cols] = vals
    b = rng.normal(size=mech.n)
    lds, core = emulate_factor(tr, A.T)
    x = emulate_solve(tr, lds, core, b)
    assert np.allclose(A.T @ x, b, rtol=1e-10, atol=1e-10)
    # assembly sources: the transposed plan reads A^T's entries from the same
    # Jacobian values array
    ar = tr.arrays
    si, sj = np.nonzero(tr.F)
    keep = np.minimum(si, sj) < tr.c
    si, sj = si[keep], sj[keep]
    src = ar["sp_src"]
    has = src >= 0
    At = A.T[np.ix_(tr.perm, tr.perm)]
    assert np.array_equal(vals[src[has]], At[si[has], sj[has]])
