# -*- coding: utf-8 -*-
"""ctypes binding of libwds.so (include/wds.h).

The library is built in-tree (``python -m wdsgpu_amd.build``).  Loading it
fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwds.so")

WDS_OK, WDS_EINVAL, WDS_EUNSUPPORTED, WDS_ENOTSPD = 0, -1, -2, -3
WDS_EUNCONVERGED, WDS_EDEVICE, WDS_ESTALE = -4, -5, -6

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class NetworkDesc(C.Structure):
    _fields_ = [("N", C.c_int32), ("L", C.c_int32),
                ("node_type", _ip), ("elev", _dp),
                ("link_n1", _ip), ("link_n2", _ip), ("link_type", _ip),
                ("init_status", _ip), ("diam", _dp), ("km", _dp),
                ("km_tcv", _dp), ("km_valve", _dp), ("kc", _dp), ("q0", _dp),
                ("setting", _dp), ("pump_kind", _ip), ("pump_h0", _dp),
                ("pump_r", _dp), ("pump_n", _dp), ("pump_hmax", _dp),
                ("ntank", C.c_int32), ("tank_node", _ip), ("tank_hmin", _dp),
                ("tank_hmax", _dp), ("tank_overflow", _ip),
                ("headloss", C.c_int32), ("hexp", C.c_double), ("qexp", C.c_double),
                ("rqtol", C.c_double), ("viscos", C.c_double), ("htol", C.c_double),
                ("qtol", C.c_double), ("checkfreq", C.c_int32),
                ("maxcheck", C.c_int32), ("extra_iter", C.c_int32),
                ("pump_curve_ptr", _ip), ("pump_curve_x", _dp), ("pump_curve_y", _dp),
                ("ucf_flow", C.c_double), ("ucf_head", C.c_double)]


class PlanStats(C.Structure):
    _fields_ = [("N", C.c_int32), ("L", C.c_int32), ("Nj", C.c_int32),
                ("nnzL", C.c_int32), ("nlevels", C.c_int32), ("npcv", C.c_int32),
                ("nnz_csr", C.c_int64), ("nnz_lower", C.c_int64),
                ("factor_updates", C.c_int64), ("factor_steps", C.c_int32),
                ("fwd_steps", C.c_int32), ("bwd_steps", C.c_int32),
                ("factor_instr", C.c_int64), ("fwd_instr", C.c_int64),
                ("bwd_instr", C.c_int64), ("lds_bytes", C.c_int64),
                ("arena_bytes", C.c_int64), ("ordering", C.c_int32),
                ("iteration_cost", C.c_int32)]


class Counters(C.Structure):
    _fields_ = [("factorize", C.c_int64), ("solve", C.c_int64),
                ("bwd_solve", C.c_int64), ("bwd_reuse", C.c_int64),
                ("bwd_refactorize", C.c_int64), ("launches", C.c_int64)]


class SolveArgs(C.Structure):
    _fields_ = [("B", C.c_int32), ("demand", C.c_void_p), ("res_head", C.c_void_p),
                ("ke", C.c_void_p), ("ke_stride", C.c_int32),
                ("r_hw", C.c_void_p), ("r_stride", C.c_int32),
                ("q0", C.c_void_p), ("e0", C.c_void_p),
                ("max_iter", C.c_int32), ("accuracy", C.c_double),
                ("status_machine", C.c_int32), ("n_refine", C.c_int32),
                ("polish_steps", C.c_int32),
                ("head", C.c_void_p), ("flow", C.c_void_p), ("emitter", C.c_void_p),
                ("iters", C.c_void_p), ("relerr", C.c_void_p), ("status", C.c_void_p),
                ("diag_ratio", C.c_void_p), ("info", C.c_void_p), ("resid", C.c_void_p),
                ("factor", C.c_void_p), ("workspace", C.c_void_p), ("status0", C.c_void_p),
                ("sched_flags", C.c_void_p), ("sched_events", C.c_void_p),
                ("sched_count", C.c_void_p), ("sched_cap", C.c_int32),
                ("res_head_shared", C.c_int32)]


class AdjointArgs(C.Structure):
    _fields_ = [("B", C.c_int32), ("demand", C.c_void_p), ("ke", C.c_void_p),
                ("ke_stride", C.c_int32), ("r_hw", C.c_void_p), ("r_stride", C.c_int32),
                ("head", C.c_void_p), ("flow", C.c_void_p), ("emitter", C.c_void_p),
                ("status", C.c_void_p), ("factor", C.c_void_p),
                ("gH", C.c_void_p), ("gQ", C.c_void_p), ("gE", C.c_void_p),
                ("n_refine", C.c_int32),
                ("g_demand", C.c_void_p), ("g_res_head", C.c_void_p),
                ("g_ke", C.c_void_p), ("g_r", C.c_void_p), ("info", C.c_void_p),
                ("workspace", C.c_void_p), ("g_pump", C.c_void_p)]


# every exported symbol of include/wds.h with its signature
SIGNATURES = {
    "wds_plan_create": (C.c_int, [C.POINTER(NetworkDesc), C.c_int32, C.POINTER(C.c_void_p)]),
    "wds_plan_destroy": (None, [C.c_void_p]),
    "wds_plan_stats_get": (C.c_int, [C.c_void_p, C.POINTER(PlanStats)]),
    "wds_plan_export": (C.c_int64, [C.c_void_p, C.c_int32, _ip, C.c_int64]),
    "wds_plan_upload": (C.c_int, [C.c_void_p]),
    "wds_workspace_bytes": (C.c_int64, [C.c_void_p, C.c_int32]),
    "wds_factor_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                   C.c_void_p, C.c_void_p]),
    "wds_factor": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                             C.c_void_p, C.c_void_p]),
    "wds_trisolve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "wds_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(SolveArgs)]),
    "wds_plan_set_pumps": (C.c_int, [C.c_void_p, C.c_void_p, _dp, _dp, _dp, _dp, _dp]),
    "wds_solve_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(SolveArgs)]),
    "wds_multi_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_int32),
                                   C.c_int64, C.POINTER(C.c_void_p)]),
    "wds_multi_destroy": (None, [C.c_void_p]),
    "wds_multi_workspace_bytes": (C.c_int64, [C.c_void_p, C.c_int32]),
    "wds_multi_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int64)]),
    "wds_multi_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(SolveArgs)]),
    "wds_multi_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(AdjointArgs)]),
    "wds_newton_iter": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(SolveArgs),
                                  C.c_int32, C.c_void_p]),
    "wds_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(AdjointArgs)]),
    "wds_counters_get": (C.c_int, [C.c_void_p, C.POINTER(Counters), C.c_int32]),
    "wds_last_error": (C.c_char_p, []),
    "wds_link_coeffs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                  C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "wds_hw_resistance": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p]),
    "wds_hw_resistance_grad": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p]),
    "wds_emitter_ke": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_double, C.c_double,
                                 C.c_void_p]),
    "wds_batch_sums": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_void_p),
                                 C.POINTER(C.c_int64), C.POINTER(C.c_void_p)]),
    "wds_emitter_ke_grad": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                      C.c_double, C.c_void_p, C.c_void_p]),
    "wds_detmath_eval": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p,
                                   C.c_void_p]),
}

_lib = None


def load():
    """Load libwds.so once; raise ImportError naming the build command if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with "
                          "`python -m wdsgpu_amd.build` (there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, what):
    """Map a wds_status to the reference's exception classes
    (ValueError / NotImplementedError / RuntimeError)."""
    if rc >= 0:
        return rc
    msg = f"{what}: {load().wds_last_error().decode(errors='replace')}"
    if rc == WDS_EINVAL:
        raise ValueError(msg)
    if rc == WDS_EUNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


def ptr(t):
    """Device/host address of a torch tensor (None -> NULL)."""
    return None if t is None else C.c_void_p(t.data_ptr())
