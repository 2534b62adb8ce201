"""ctypes binding of the HIP library's C ABI (include/fastode_hip.h).

The library is built in-tree by ``build.py``; when it is missing this module
raises at import, so that no product path can silently run without it.
"""
from __future__ import annotations

import ctypes as ct
import os
from pathlib import Path

import numpy as np
# torch first: its HIP runtime (soname libamdhip64.so.7) must be the one the
# library binds to, so that torch's streams and allocations and the library's
# kernels live in one runtime instance.
import torch  # noqa: F401

HERE = Path(__file__).resolve().parent
# the product library, or (FASTODE_AMD_LIB=libfastode_hip_timed.so) the same
# code built with the batched kernel's phase timers
LIB_NAMES = ("libfastode_hip.so", "libfastode_hip_timed.so")
_name = os.environ.get("FASTODE_AMD_LIB", LIB_NAMES[0])
if _name not in LIB_NAMES:
    raise ImportError(f"FASTODE_AMD_LIB must be one of {LIB_NAMES}")
if _name != LIB_NAMES[0]:
    raise ImportError("the phase-timer library is switched off: both of its runs on an MI355X went silent "
                      "for 180 s and were killed, with no cause found in the code")
LIB_PATH = HERE / _name
PHASES = ("load", "rates", "jacobian", "assemble", "sparse_factor", "schur", "core_factor", "inverses",
          "stage_setup", "sparse_solve", "core_solve", "combine", "control")

if not LIB_PATH.exists():
    raise ImportError(f"{LIB_PATH} is not built: run __graft_entry__.build() (hipcc, gfx950)"
                      + (" with phase timers: build.py --phase-timers" if _name != LIB_NAMES[0] else ""))

lib = ct.CDLL(str(LIB_PATH))

NCOUNTS = 13
RUNNING, CONVERGED, TOO_MANY_STEPS, STEP_TOO_SMALL, NAN_OR_INF, PROJECTION_FAILED, ACCEPTED_UNCONVERGED = range(7)
FINISH, NEW_STEP, RETRY, NEWTON = range(4)
LOOP_ALGORITHM, LOOP_RESTARTING, LOOP_BACKWARD_EULER = range(3)
STATUS = {CONVERGED: "converged", TOO_MANY_STEPS: "too many steps",
          STEP_TOO_SMALL: "step too small", NAN_OR_INF: "nan or inf",
          PROJECTION_FAILED: "projection failed", RUNNING: "running",
          ACCEPTED_UNCONVERGED: "accepted unconverged"}

# fode_group_ctrl as a NumPy record (232 bytes, 8-byte aligned fields).
CTRL_DTYPE = np.dtype([
    ("t", "f8"), ("H", "f8"), ("H_ctrl", "f8"), ("proposed", "f8"), ("dt", "f8"),
    ("h_min", "f8"), ("h_max", "f8"), ("err", "f8"), ("alpha", "f8"),
    ("h_next", "f8"), ("err_prev", "f8"), ("h_prev", "f8"),
    ("nsteps", "i4"), ("status", "i4"), ("last", "i4"), ("shortened", "i4"),
    ("rej_last", "i4"), ("rej_more", "i4"), ("action", "i4"), ("accepted", "i4"),
    ("fails", "i4"), ("first_fail", "i4"),
    ("counts", "i8", (NCOUNTS,)),
])
assert CTRL_DTYPE.itemsize == 240
CTRL_WORDS = CTRL_DTYPE.itemsize // 8
ALPHA_OFFSET = CTRL_DTYPE.fields["alpha"][1]
NSTEPS_WORD = CTRL_DTYPE.fields["nsteps"][1] // 4


class CtrlParams(ct.Structure):
    _fields_ = [("b", ct.c_double), ("safety", ct.c_double), ("fac_min", ct.c_double),
                ("fac_max", ct.c_double), ("k_exp", ct.c_double), ("gamma", ct.c_double),
                ("h211b", ct.c_int32), ("carry", ct.c_int32), ("quantize", ct.c_int32),
                ("carry_memory", ct.c_int32), ("max_steps", ct.c_int32), ("stages", ct.c_int32),
                ("n_new_forcing", ct.c_int32), ("constrained", ct.c_int32),
                ("a", ct.c_double * 16), ("c", ct.c_double * 16), ("m", ct.c_double * 4),
                ("e", ct.c_double * 4), ("new_forcing", ct.c_int32 * 4),
                ("loop", ct.c_int32), ("step_change_rows", ct.c_int32), ("max_iterations", ct.c_int32),
                ("n_reductions", ct.c_int32), ("reductions", ct.c_double * 8), ("small", ct.c_double)]


class MechDesc(ct.Structure):
    _fields_ = [("n", ct.c_int32), ("R", ct.c_int32), ("P", ct.c_int32), ("nnz", ct.c_int32),
                ("ncon", ct.c_int32),
                ("react", ct.c_void_p), ("rhs_ptr", ct.c_void_p), ("rhs_rxn", ct.c_void_p),
                ("rhs_coef", ct.c_void_p), ("jac_ptr", ct.c_void_p), ("jac_src", ct.c_void_p),
                ("jac_coef", ct.c_void_p), ("jac_const", ct.c_void_p), ("con_rows", ct.c_void_p),
                ("con_coef", ct.c_void_p), ("con_total", ct.c_void_p)]


class PlanDesc(ct.Structure):
    _fields_ = [("n", ct.c_int32), ("c", ct.c_int32), ("m", ct.c_int32), ("tile", ct.c_int32),
                ("ns", ct.c_int32), ("nsch", ct.c_int32), ("nnz", ct.c_int32), ("slot_stride", ct.c_int32),
                ("n_f_tasks", ct.c_int32), ("f_nst", ct.c_int32), ("n_f_pairs", ct.c_int32),
                ("n_sf_tasks", ct.c_int32), ("sf_nst", ct.c_int32), ("n_sb_tasks", ct.c_int32),
                ("sb_nst", ct.c_int32), ("n_s_pairs", ct.c_int32), ("n_sch_pairs", ct.c_int32),
                ("sp_src", ct.c_void_p), ("sp_diag", ct.c_void_p), ("sp_row", ct.c_void_p),
                ("core_src", ct.c_void_p), ("sch_mask", ct.c_void_p), ("sch_base", ct.c_void_p),
                ("sch_pairs", ct.c_void_p), ("core_row", ct.c_void_p),
                ("f_tasks", ct.c_void_p), ("f_ptr", ct.c_void_p), ("f_pairs", ct.c_void_p),
                ("sf_tasks", ct.c_void_p), ("sf_ptr", ct.c_void_p),
                ("sb_tasks", ct.c_void_p), ("sb_ptr", ct.c_void_p),
                ("s_pairs", ct.c_void_p), ("perm", ct.c_void_p)]


class BatchDesc(ct.Structure):
    _fields_ = [("C", ct.c_int32), ("group_size", ct.c_int32), ("ngroups", ct.c_int32),
                ("k_stride", ct.c_int32), ("dt", ct.c_double), ("rtol", ct.c_double),
                ("Y", ct.c_void_p), ("k", ct.c_void_p), ("kmul", ct.c_void_p), ("atol", ct.c_void_p),
                ("mass", ct.c_void_p), ("g", ct.c_void_p), ("work", ct.c_void_p),
                ("queue", ct.c_void_p), ("hlog", ct.c_void_p), ("nacc", ct.c_void_p),
                ("hcap", ct.c_int32), ("options", ct.c_int32), ("order", ct.c_void_p)]


class AdjointDesc(ct.Structure):
    _fields_ = [("C", ct.c_int32), ("group_size", ct.c_int32), ("ngroups", ct.c_int32),
                ("k_stride", ct.c_int32), ("hcap", ct.c_int32), ("pad_", ct.c_int32),
                ("ws_stride", ct.c_int64),
                ("Y", ct.c_void_p), ("k", ct.c_void_p), ("kmul", ct.c_void_p), ("mass", ct.c_void_p),
                ("hlog", ct.c_void_p), ("nacc", ct.c_void_p), ("lam", ct.c_void_p), ("kbar", ct.c_void_p),
                ("ws", ct.c_void_p), ("queue", ct.c_void_p)]


vp = ct.c_void_p
i32 = ct.c_int32
P = ct.POINTER

_SIGS = {
    "fode_last_error": (ct.c_char_p, []),
    "fode_device_info": (i32, [i32, P(i32), P(i32), P(i32)]),
    "fode_mech_create": (i32, [P(MechDesc), P(vp)]),
    "fode_mech_destroy": (i32, [vp]),
    "fode_plan_create": (i32, [P(PlanDesc), P(vp), P(ct.c_int64)]),
    "fode_plan_destroy": (i32, [vp]),
    "fode_rhs": (i32, [vp, i32, vp, vp, vp, vp]),
    "fode_jac": (i32, [vp, i32, vp, vp, vp, vp]),
    "fode_factor": (i32, [vp, i32, vp, vp, vp, i32, vp, vp]),
    "fode_solve": (i32, [vp, i32, vp, vp, vp, vp]),
    "fode_ctrl_init": (i32, [P(CtrlParams), vp, i32, i32, vp]),
    "fode_ctrl_begin": (i32, [P(CtrlParams), vp, i32, vp]),
    "fode_stage_point": (i32, [P(CtrlParams), i32, i32, i32, vp, P(vp), vp, vp]),
    "fode_stage_forcing": (i32, [P(CtrlParams), i32, i32, i32, vp, vp, P(vp), vp, i32, vp, vp]),
    "fode_stage_final": (i32, [P(CtrlParams), i32, i32, vp, P(vp), vp, vp, vp]),
    "fode_norm_control": (i32, [P(CtrlParams), vp, i32, i32, vp, vp, vp, vp, ct.c_double, vp, vp, vp]),
    "fode_clamp": (i32, [i32, i32, vp, vp, i32, vp]),
    "fode_proj_mask": (i32, [i32, i32, vp, vp, vp, vp]),
    "fode_proj_candidate": (i32, [i32, i32, vp, vp, ct.c_double, vp, vp, vp]),
    "fode_proj_measure": (i32, [i32, i32, vp, vp, vp, vp, ct.c_double, vp, vp]),
    "fode_be_residual": (i32, [i32, i32, vp, vp, vp, vp, vp, vp]),
    "fode_scale_columns": (i32, [i32, i32, vp, i32, vp, vp, vp, vp, vp]),
    "fode_be_update": (i32, [P(CtrlParams), vp, i32, i32, vp, vp, vp, vp, ct.c_double, vp, vp]),
    "fode_batched_interval": (i32, [vp, vp, P(CtrlParams), P(BatchDesc), i32, vp]),
    "fode_batched_lds_bytes": (i32, [vp, vp, i32, P(ct.c_int64)]),
    "fode_batched_interval_noalloc": (i32, [vp, vp, P(CtrlParams), P(BatchDesc), i32, vp]),
    "fode_phase_timers": (i32, [P(ct.c_uint64), i32, i32]),
    "fode_batched_graph_create": (i32, [vp, vp, P(CtrlParams), P(BatchDesc), i32, P(vp)]),
    "fode_graph_launch": (i32, [vp, vp]),
    "fode_graph_destroy": (i32, [vp]),
    "fode_adjoint_interval": (i32, [vp, vp, vp, P(CtrlParams), P(AdjointDesc), vp]),
    "fode_adjoint_ws_blocks": (i32, [vp, vp, vp, i32, i32, P(i32)]),
}

for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args

EXPORTS = tuple(_SIGS)


class HipError(RuntimeError):
    pass


def check(rc, what=""):
    if rc != 0:
        msg = lib.fode_last_error().decode(errors="replace")
        raise HipError(f"{what}: {msg}" if what else msg)


def ptr(t):
    """Device (or host) address of a torch tensor, or None."""
    return None if t is None else ct.c_void_p(t.data_ptr())


def stream_handle():
    import torch
    return ct.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr_array(tensors):
    arr = (ct.c_void_p * 4)()
    for i, t in enumerate(tensors):
        arr[i] = t.data_ptr()
    return arr


def device_info(device=0):
    lds, cus, ws = i32(), i32(), i32()
    check(lib.fode_device_info(device, ct.byref(lds), ct.byref(cus), ct.byref(ws)), "device_info")
    return {"lds_per_block": lds.value, "cus": cus.value, "warp_size": ws.value}
