"""ctypes binding of libhdg (include/hdg.h).

The library is built in-tree by build.py.  Importing this module never falls
back to anything else: if libhdg.so is missing, `lib()` raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhdg.so")

HDG_OK = 0
STATUS = {
    1: "HDG_ERR_ARG", 2: "HDG_ERR_NEG_JACOBIAN", 3: "HDG_ERR_QUAD", 4: "HDG_ERR_LOCAL_SINGULAR",
    5: "HDG_ERR_NOT_CONVERGED", 6: "HDG_ERR_HBM_BUDGET", 7: "HDG_ERR_HIP", 8: "HDG_ERR_STATE",
    9: "HDG_ERR_UNSUPPORTED", 10: "HDG_ERR_BOUNDARY",
}
HDG_FP64, HDG_MIXED = 0, 1
PC_AUTO, PC_BLOCK_JACOBI, PC_MULTIGRID = 0, 1, 2
SCHEDULE_EAGER = 1
SCHEDULE_AUTOTUNE = 2
SCHEDULE_LU = 4
SRC_REFERENCE, SRC_KAPPA, SRC_SINE_SERIES, SRC_VALUES = 0, 1, 2, 3
UEX_SINSIN, UEX_LINEAR = 0, 1
G_LINEAR, G_SINSIN = 0, 1
DIRICHLET_REFERENCE = 1

# every symbol include/hdg.h declares
EXPORTS = (
    "hdg_create", "hdg_destroy", "hdg_last_error", "hdg_memory_used", "hdg_rectangle_nfaces",
    "hdg_rectangle_mesh", "hdg_set_mesh", "hdg_setup", "hdg_set_source", "hdg_set_dirichlet",
    "hdg_query_sizes", "hdg_assemble_condense", "hdg_get_system", "hdg_get_local_ops",
    "hdg_get_condensed", "hdg_get_element_blocks", "hdg_solve", "hdg_set_trace", "hdg_recover",
    "hdg_errornorm", "hdg_run", "hdg_dubiner", "hdg_quadrature", "hdg_time_kernel", "hdg_get_trace",
    "hdg_server_create", "hdg_server_submit", "hdg_server_wait", "hdg_server_trace", "hdg_server_footprint",
    "hdg_server_stats", "hdg_server_destroy", "hdg_server_last_error", "hdg_server_set_batching",
    "hdg_batch_create", "hdg_batch_run", "hdg_batch_trace", "hdg_batch_info", "hdg_batch_destroy",
    "hdg_batch_last_error", "hdg_autotune", "hdg_set_schedule", "hdg_set_rectangle_mesh",
    "hdg_set_dirichlet_boundary", "hdg_errornorm_noalloc", "hdg_set_dirichlet_function",
    "hdg_get_applied_system", "hdg_refine_probe",
)


class HDGError(RuntimeError):
    """A non-zero hdg_status; `.code` and `.name` identify it."""

    def __init__(self, code, msg):
        self.code = code
        self.name = STATUS.get(code, f"status {code}")
        super().__init__(f"{self.name}: {msg}")


class ArgumentError(HDGError, ValueError):
    """Raised where the reference throws ArgumentError."""


class SingularException(HDGError, ArithmeticError):
    """Local (q,u) matrix singular (LinearAlgebra.SingularException)."""


class Options(C.Structure):
    _fields_ = [("k", C.c_int32), ("quad_degree", C.c_int32), ("tau", C.c_double),
                ("precision", C.c_int32), ("rtol", C.c_double), ("maxit", C.c_int32),
                ("schedule", C.c_int32), ("precond", C.c_int32)]


class Source(C.Structure):
    _fields_ = [("kind", C.c_int32), ("kappa_on", C.c_int32), ("ka", C.c_double),
                ("kb", C.c_double), ("coef", C.c_double * 16), ("f_qp", C.c_void_p)]


class Sizes(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("n_cells", C.c_int64), ("n_faces", C.c_int64),
                ("n_dofs", C.c_int64), ("nnz", C.c_int64), ("nu", C.c_int32), ("nq", C.c_int32),
                ("nt", C.c_int32), ("m", C.c_int32), ("ntr", C.c_int32), ("nqp", C.c_int32),
                ("nfq", C.c_int32), ("sym_blocks", C.c_int32)]


class SolveInfo(C.Structure):
    _fields_ = [("iters", C.c_int32), ("refinements", C.c_int32), ("relres", C.c_double),
                ("true_relres", C.c_double), ("seconds", C.c_double)]


class Request(C.Structure):
    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64), ("jitter", C.c_double), ("seed", C.c_uint64),
                ("opt", Options), ("src", Source), ("uex_kind", C.c_int32), ("want_trace", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int32), ("iters", C.c_int32), ("relres", C.c_double),
                ("true_relres", C.c_double), ("errornorm", C.c_double), ("n_cells", C.c_int64),
                ("n_dofs", C.c_int64), ("t_submit", C.c_double), ("t_start", C.c_double),
                ("t_done", C.c_double), ("device_bytes", C.c_size_t), ("slot", C.c_int32)]


class Schedule(C.Structure):
    _fields_ = [("condense_seg", C.c_int32), ("condense_grid", C.c_int32), ("condense_us", C.c_double),
                ("default_us", C.c_double), ("condense_form", C.c_int32), ("condense_waves", C.c_int32)]


class ServerSummary(C.Structure):
    _fields_ = [("completed", C.c_int64), ("failed", C.c_int64), ("cells", C.c_int64),
                ("p50_latency", C.c_double), ("p99_latency", C.c_double), ("mean_latency", C.c_double),
                ("mean_queue_wait", C.c_double), ("busy_seconds", C.c_double),
                ("cells_per_second", C.c_double), ("peak_reserved", C.c_size_t)]


_lib = None
_P = C.POINTER
_d, _i64, _i8 = _P(C.c_double), _P(C.c_int64), _P(C.c_int8)
_ctx = C.c_void_p


def lib():
    """The loaded libhdg (raises FileNotFoundError if it was never built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with hdiscontinuousgalerkin.jl_amd/build.py")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    sig = {
        "hdg_create": (C.c_int, [_P(_ctx), C.c_int, C.c_size_t]),
        "hdg_destroy": (None, [_ctx]),
        "hdg_last_error": (C.c_char_p, [_ctx]),
        "hdg_memory_used": (C.c_int, [_ctx, _P(C.c_size_t)]),
        "hdg_rectangle_nfaces": (C.c_int64, [C.c_int64, C.c_int64]),
        "hdg_rectangle_mesh": (C.c_int, [C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.c_double, C.c_uint64, _d, _i64, _i64, _i8]),
        "hdg_set_mesh": (C.c_int, [_ctx, C.c_int64, _d, C.c_int64, _i64, C.c_int64, _i64]),
        "hdg_setup": (C.c_int, [_ctx, _P(Options)]),
        "hdg_set_source": (C.c_int, [_ctx, _P(Source)]),
        "hdg_set_dirichlet": (C.c_int, [_ctx, C.c_int64, _i64, _d]),
        "hdg_query_sizes": (C.c_int, [_ctx, _P(Sizes)]),
        "hdg_assemble_condense": (C.c_int, [_ctx]),
        "hdg_get_system": (C.c_int, [_ctx, _i64, _i64, _d, _d]),
        "hdg_get_local_ops": (C.c_int, [_ctx, _d, _d]),
        "hdg_get_condensed": (C.c_int, [_ctx, _d, _d]),
        "hdg_get_element_blocks": (C.c_int, [_ctx, C.c_int64, C.c_int64] + [_d] * 7),
        "hdg_solve": (C.c_int, [_ctx, _d, _P(SolveInfo)]),
        "hdg_set_trace": (C.c_int, [_ctx, _d]),
        "hdg_recover": (C.c_int, [_ctx, _d, _d, _d]),
        "hdg_errornorm": (C.c_int, [_ctx, C.c_int, _d, _d]),
        "hdg_errornorm_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_int, _d, C.c_void_p, C.c_void_p]),
        "hdg_run": (C.c_int, [_ctx, _P(SolveInfo)]),
        "hdg_dubiner": (C.c_int, [C.c_int, C.c_double, C.c_double, _d, _d]),
        "hdg_quadrature": (C.c_int, [C.c_int, C.c_int, C.c_int, _d, _d]),
        "hdg_time_kernel": (C.c_int, [_ctx, C.c_int, C.c_int, _d]),
        "hdg_refine_probe": (C.c_int, [_ctx, C.c_int, C.c_double, _d, C.c_void_p]),
        "hdg_get_trace": (C.c_int, [_ctx, _d]),
        "hdg_server_create": (C.c_int, [_P(C.c_void_p), C.c_int, C.c_size_t, C.c_int]),
        "hdg_server_submit": (C.c_int, [C.c_void_p, _P(Request), _i64]),
        "hdg_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.c_double, _P(Result)]),
        "hdg_server_trace": (C.c_int, [C.c_void_p, C.c_int64, _d, C.c_int64]),
        "hdg_server_footprint": (C.c_size_t, [_P(Request)]),
        "hdg_server_stats": (C.c_int, [C.c_void_p, _P(ServerSummary)]),
        "hdg_server_destroy": (None, [C.c_void_p]),
        "hdg_server_last_error": (C.c_char_p, [C.c_void_p]),
        "hdg_server_set_batching": (C.c_int, [C.c_void_p, C.c_int, C.c_int64]),
        "hdg_autotune": (C.c_int, [C.c_void_p, _P(Schedule)]),
        "hdg_set_rectangle_mesh": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double,
                                             C.c_double, C.c_double, C.c_uint64]),
        "hdg_set_dirichlet_boundary": (C.c_int, [C.c_void_p, C.c_double]),
        "hdg_set_dirichlet_function": (C.c_int, [C.c_void_p, C.c_int, _d, C.c_int]),
        "hdg_get_applied_system": (C.c_int, [_ctx, _i64, _i64, _d, _d, _d]),
        "hdg_set_schedule": (C.c_int, [C.c_void_p, _P(Schedule)]),
        "hdg_batch_create": (C.c_int, [_P(C.c_void_p), C.c_int, C.c_size_t, C.c_int, _P(Request)]),
        "hdg_batch_run": (C.c_int, [C.c_void_p, _P(Result)]),
        "hdg_batch_trace": (C.c_int, [C.c_void_p, C.c_int, _d, C.c_int64]),
        "hdg_batch_info": (C.c_int, [C.c_void_p, _P(C.c_int), _P(C.c_size_t)]),
        "hdg_batch_destroy": (None, [C.c_void_p]),
        "hdg_batch_last_error": (C.c_char_p, [C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def ptr(a, t=_d):
    return None if a is None else a.ctypes.data_as(t)


def check(code, ctx=None):
    if code == HDG_OK:
        return
    msg = ""
    if ctx is not None:
        raw = lib().hdg_last_error(ctx)
        msg = raw.decode() if raw else ""
    if code in (1, 2, 3, 10):
        raise ArgumentError(code, msg)
    if code == 4:
        raise SingularException(code, msg)
    raise HDGError(code, msg)
