"""CPU checks of the host side: the symbolic LU plan (through a CPU model of
exactly what the kernels do with it), the mechanism tables against the oracle,
and the C ABI library's exported symbols."""
import ctypes as ct
import re
from pathlib import Path

import numpy as np
import pytest

import fastode_oracle as O
from fastode_amd import workloads as W
from fastode_amd.mechanism import Mechanism, synthetic
from fastode_amd.symbolic import G, build_plan, emulate_factor, emulate_solve

ROOT = Path(__file__).resolve().parents[1]


def random_matrix(n, rows, cols, rng, dominance=20.0):
    A = np.zeros((n, n))
    A[rows, cols] = rng.standard_normal(len(rows))
    A[np.arange(n), np.arange(n)] += dominance
    return A


@pytest.mark.parametrize("n,R,seed", [(20, 40, 1), (100, 250, 4), (200, 500, 2), (200, 500, 7)])
def test_plan_factor_and_solve_match_numpy(n, R, seed):
    mech = synthetic(n, R, seed).mech
    plan = build_plan(n, mech.rows, mech.cols)
    assert plan.c + plan.m == n and plan.m <= plan.m_pad
    rng = np.random.default_rng(seed)
    A = random_matrix(n, mech.rows, mech.cols, rng)
    lds, core = emulate_factor(plan, A)
    for _ in range(2):
        b = rng.standard_normal(n)
        x = emulate_solve(plan, lds, core, b)
        xr = np.linalg.solve(A, b)
        assert np.abs(x - xr).max() <= 1e-12 * np.abs(xr).max()


@pytest.mark.parametrize("n", [1, 2, 3, 16, 17, 40])
def test_dense_plans(n):
    rows = np.repeat(np.arange(n), n)
    cols = np.tile(np.arange(n), n)
    plan = build_plan(n, rows, cols, order=np.arange(n))
    assert plan.c == 0 and plan.ns == 0 and plan.tile == -(-n // G)
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 3 * n * np.eye(n)
    lds, core = emulate_factor(plan, A)
    b = rng.standard_normal(n)
    assert np.allclose(emulate_solve(plan, lds, core, b), np.linalg.solve(A, b), rtol=1e-12, atol=1e-14)


def test_plan_fill_is_a_superset_and_levels_are_causal():
    mech = synthetic(200, 500, 2).mech
    plan = build_plan(mech.n, mech.rows, mech.cols)
    inv = np.empty(mech.n, dtype=np.int64)
    inv[plan.perm] = np.arange(mech.n)
    assert plan.F[inv[mech.rows], inv[mech.cols]].all()
    # every factor task's pairs read entries final before its stage
    a = plan.arrays
    tasks = a["f_tasks"].reshape(-1, 2).astype(np.int64)
    ptr = a["f_ptr"]
    done_stage = {}
    for s in range(len(ptr) - 1):
        for t in range(ptr[s], ptr[s + 1]):
            done_stage[int(tasks[t, 0] & 0xFFFF)] = s
    for s in range(len(ptr) - 1):
        for t in range(ptr[s], ptr[s + 1]):
            start, count = tasks[t, 1] & 0xFFFFF, tasks[t, 1] >> 20
            for q in range(start, start + count):
                w = int(a["f_pairs"][q])
                for src in (w >> 16, w & 0xFFFF):
                    assert done_stage.get(src, -1) < s
    # the division source too, and its task is the entry's last
    for s in range(len(ptr) - 1):
        for t in range(ptr[s], ptr[s + 1]):
            d = int(tasks[t, 0] >> 16)
            if d != 0xFFFF:
                assert done_stage.get(d, -1) < s
                assert done_stage[int(tasks[t, 0] & 0xFFFF)] == s
    # Schur sums: each used by exactly one core slot, closed by exactly one
    # pair of its owner's list, whose factors are sparse entries (final after
    # the factor tasks, which the Schur update follows)
    mask = a["sch_mask"].reshape(256, 4)
    per_thread = np.array([sum(bin(int(x)).count("1") for x in row) for row in mask])
    assert per_thread.sum() == plan.nsch and per_thread.max() == plan.sch_depth
    base = a["sch_base"].astype(np.int64)
    assert base[0] == 0 and base[-1] == len(a["sch_pairs"]) and (base % 4 == 0).all()
    pr = a["sch_pairs"].astype(np.int64)
    for t in range(256):
        assert int((pr[base[t]:base[t + 1]] >> 31).sum()) == per_thread[t]
    assert (((pr >> 16) & 0x7FFF) < plan.ns).all() and ((pr & 0xFFFF) < plan.ns).all()


@pytest.mark.parametrize("which", ["ts1", "synthetic200"])
def test_factor_chains_keep_their_order_across_stages(which):
    """An entry's factor chain may be cut into runs over several stages; the
    runs, taken in stage order, are exactly its multiply-adds in ascending
    pivot order (so the bits are those of one task), each stage updates an
    entry at most once, and the division comes with the last run."""
    if which == "ts1":
        from fastode_amd.ts1 import TS1
        mech = TS1().mech
    else:
        mech = synthetic(200, 500, 2).mech
    plan = build_plan(mech.n, mech.rows, mech.cols)
    F, c = plan.F, plan.c
    si, sj = np.nonzero(F)
    keep = np.minimum(si, sj) < c
    si, sj = si[keep], sj[keep]
    index = {(int(i), int(j)): e for e, (i, j) in enumerate(zip(si, sj))}
    a = plan.arrays
    tasks = a["f_tasks"].reshape(-1, 2).astype(np.int64)
    ptr = a["f_ptr"]
    chain, divs, cut = {}, {}, 0
    for s in range(len(ptr) - 1):
        seen = set()
        for t in range(ptr[s], ptr[s + 1]):
            e = int(tasks[t, 0] & 0xFFFF)
            assert e not in seen
            seen.add(e)
            assert e not in divs, "a run after the division"
            start, count = int(tasks[t, 1] & 0xFFFFF), int(tasks[t, 1] >> 20)
            cut += e in chain
            chain.setdefault(e, []).extend(int(w) for w in a["f_pairs"][start:start + count])
            if tasks[t, 0] >> 16 != 0xFFFF:
                divs[e] = int(tasks[t, 0] >> 16)
    for e, (i, j) in enumerate(zip(si, sj)):
        i, j = int(i), int(j)
        want = [(index[(i, p)] << 16) | index[(p, j)] for p in range(min(i, j)) if F[i, p] and F[p, j]]
        assert chain.get(e, []) == want
        if i > j:
            assert divs[e] == index[(j, j)]
        else:
            assert e not in divs
    if which == "ts1":
        assert cut > 0          # TS1's long chains are cut


def test_mechanism_tables_match_the_oracle():
    wl = W.small()
    mech = wl.mech
    om = O.OracleMechanism(mech.species, mech.reactions_list())
    rng = np.random.default_rng(0)
    Y = wl.Y0[:4]
    k = wl.k(0)[:4]
    # forcing through the gather tables
    rates = k * np.prod(np.concatenate([Y, np.ones((4, 1))], axis=1)[:, mech.react], axis=2)
    F = np.zeros_like(Y)
    for i in range(mech.n):
        for q in range(mech.rhs_ptr[i], mech.rhs_ptr[i + 1]):
            F[:, i] += rates[:, mech.rhs_rxn[q]] * mech.rhs_coef[q]
    assert np.allclose(F, om.rhs(k, Y), rtol=1e-13, atol=1e-30)
    # Jacobian values through the gather tables, same pattern order
    assert np.array_equal(mech.rows, om.rows) and np.array_equal(mech.cols, om.cols)
    Jo = om.jac_values(k, Y)
    X = np.concatenate([Y, np.ones((4, 1))], axis=1)[:, mech.react]
    d = np.stack([k * X[:, :, 1 - p] for p in range(2)], axis=2).reshape(4, -1)
    J = np.zeros_like(Jo)
    for e in range(mech.nnz):
        for q in range(mech.jac_ptr[e], mech.jac_ptr[e + 1]):
            J[:, e] += mech.jac_coef[q] * d[:, mech.jac_src[q]]
    assert np.allclose(J, Jo, rtol=1e-13, atol=1e-30)
    _ = rng


def test_constrained_mechanism_rows():
    wl = W.constrained100(C=8)
    mech = wl.mech
    assert len(mech.constraint_rows) == 3
    # the initial state satisfies each conservation row
    g = wl.Y0 @ mech.constraint_coef.T - mech.constraint_total[None, :]
    assert np.abs(g).max() < 1e-15
    assert (wl.mass[mech.constraint_rows] == 0).all() and wl.mass.sum() == 97


def test_reactions_roundtrip():
    mech = synthetic(30, 60, 3).mech
    again = Mechanism(mech.species, mech.reactions_list())
    assert np.array_equal(again.react, mech.react) and np.array_equal(again.S, mech.S)


def test_abi_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "fastode_hip.h").read_text()
    declared = set(re.findall(r"\b(fode_[a-z_]+)\s*\(", header))
    import fastode_amd._lib as L
    for name in declared:
        assert hasattr(L.lib, name), name
        assert isinstance(getattr(L.lib, name), ct._CFuncPtr)
    assert declared == set(L.EXPORTS)


@pytest.mark.parametrize("bad", ["react", "rhs_rxn", "jac_src", "rhs_ptr", "con_rows"])
def test_mech_create_refuses_out_of_range_tables(bad):
    """fode_mech_create checks every index the kernels follow before it
    touches the device, so a corrupt table is an error message, not an
    out-of-bounds read on the GPU (no device is needed to see it)."""
    import fastode_amd._lib as L
    mech = synthetic(12, 20, seed=3, constraints=1).mech

    def arr(a, dt):
        return np.ascontiguousarray(np.asarray(a), dtype=dt)

    t = dict(react=arr(mech.react, np.int32), rhs_ptr=arr(mech.rhs_ptr, np.int32),
             rhs_rxn=arr(mech.rhs_rxn, np.int32), rhs_coef=arr(mech.rhs_coef, np.float64),
             jac_ptr=arr(mech.jac_ptr, np.int32), jac_src=arr(mech.jac_src, np.int32),
             jac_coef=arr(mech.jac_coef, np.float64), jac_const=arr(mech.jac_const, np.float64),
             con_rows=arr(mech.constraint_rows, np.int32), con_coef=arr(mech.constraint_coef, np.float64),
             con_total=arr(mech.constraint_total, np.float64))
    if bad == "react":
        t["react"][3, 0] = mech.n + 1
    elif bad == "rhs_rxn":
        t["rhs_rxn"][-1] = mech.R
    elif bad == "jac_src":
        t["jac_src"][0] = mech.R * mech.P
    elif bad == "rhs_ptr":
        t["rhs_ptr"][2], t["rhs_ptr"][3] = t["rhs_ptr"][3] + 1, t["rhs_ptr"][3]
    else:
        t["con_rows"][0] = -1
    d = L.MechDesc(n=mech.n, R=mech.R, P=mech.P, nnz=mech.nnz, ncon=len(mech.constraint_rows),
                   **{k: v.ctypes.data_as(ct.c_void_p) for k, v in t.items()})
    h = ct.c_void_p()
    assert L.lib.fode_mech_create(ct.byref(d), ct.byref(h)) == -1
    assert h.value is None
    msg = L.lib.fode_last_error().decode()
    assert msg.startswith("fode_mech_create: ") and ("bad" in msg or "monotone" in msg), msg


def test_ctrl_record_layout_matches_header():
    import fastode_amd._lib as L
    # 12 doubles, 10 int32, 13 int64 (include/fastode_hip.h: fode_group_ctrl)
    assert L.CTRL_DTYPE.itemsize == 12 * 8 + 10 * 4 + 13 * 8
    # fode_ctrl_params: 6 doubles, 8 int32, a/c (16 doubles each), m/e (4 each),
    # new_forcing (4 int32), loop..n_reductions (4 int32), reductions (8), small
    assert ct.sizeof(L.CtrlParams) == 6 * 8 + 8 * 4 + 16 * 8 * 2 + 4 * 8 * 2 + 4 * 4 + 4 * 4 + 8 * 8 + 8


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """Every field offset of the ctypes / NumPy mirrors equals offsetof() in
    the header as the C compiler lays it out."""
    import shutil
    import subprocess
    import fastode_amd._lib as L
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"fode_ctrl_params": L.CtrlParams, "fode_batch_desc": L.BatchDesc,
               "fode_mech_desc": L.MechDesc, "fode_plan_desc": L.PlanDesc,
               "fode_adjoint_desc": L.AdjointDesc}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fastode_hip.h"', "int main(void) {"]
    want = []
    for cname, cls in structs.items():
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        want.append(ct.sizeof(cls))
        for f, _ in cls._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {f}));')
            want.append(getattr(cls, f).offset)
    lines.append('printf("%zu\\n", sizeof(fode_group_ctrl));')
    want.append(L.CTRL_DTYPE.itemsize)
    for f in L.CTRL_DTYPE.names:
        lines.append(f'printf("%zu\\n", offsetof(fode_group_ctrl, {f}));')
        want.append(L.CTRL_DTYPE.fields[f][1])
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == want
