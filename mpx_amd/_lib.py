"""ctypes binding of ``include/mpx_kkt.h`` (``mpx_amd/lib/libmpx_kkt.so``).

The structs below mirror the header field by field.  Loading fails loudly when
the library has not been built: there is no CPU fallback for the product path.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmpx_kkt.so")

OK, ERR_ARG, ERR_UNSUPPORTED, ERR_WORKSPACE, ERR_DEVICE = 0, 1, 2, 3, 4
F32, F64 = 0, 1
MODEL_LTV, MODEL_TANH, MODEL_CALLER = 0, 1, 2
PHASES = 4          # MPXKKT_PHASES
NOALLOC_WORKSPACE_BYTES = 64 << 20   # MPXKKT_NOALLOC_WORKSPACE_BYTES

EXPORTED = (
    "mpxkkt_default_params",
    "mpxkkt_workspace_size",
    "mpxkkt_sqp_iterate",
    "mpxkkt_sqp_iterate_noalloc",
    "mpxkkt_sqp_direction",
    "mpxkkt_sqp_accept",
    "mpxkkt_srbd_iterate",
    "mpxkkt_srbd_reference",
    "mpxkkt_default_solve_opts",
    "mpxkkt_solve",
    "mpxkkt_kkt_residual",
    "mpxkkt_model_eval",
    "mpxkkt_shift_warmstart",
    "mpxkkt_profile_enable",
    "mpxkkt_profile_collect",
    "mpxkkt_last_error",
    "mpxkkt_version",
)


# views of the caller-linearised form only (MODEL_CALLER)
CALLER_VIEWS = ("fbar", "M", "q", "r", "g")


class View(C.Structure):
    _fields_ = [("ptr", C.c_void_p), ("batch_stride", C.c_int64), ("stage_stride", C.c_int64)]


class Problem(C.Structure):
    _fields_ = [
        ("dtype", C.c_int32), ("model", C.c_int32),
        ("batch", C.c_int32), ("n", C.c_int32), ("m", C.c_int32), ("horizon", C.c_int32),
        ("n_eq", C.c_int32), ("has_box", C.c_int32),
        ("h", C.c_double), ("kappa", C.c_double),
        ("A", View), ("B", View), ("W", View),
        ("Q", View), ("R", View), ("QN", View), ("xref", View), ("uref", View),
        ("D", View), ("E", View), ("e", View),
        ("ulo", View), ("uhi", View),
        ("x0", View),
        ("fbar", View), ("M", View), ("q", View), ("r", View), ("g", View),
    ]


class Iterate(C.Structure):
    _fields_ = [("X", C.c_void_p), ("U", C.c_void_p), ("V", C.c_void_p),
                ("Y_eq", C.c_void_p), ("Y_box", C.c_void_p)]


class Params(C.Structure):
    _fields_ = [
        ("rho0", C.c_double), ("rho_max", C.c_double), ("rho_c", C.c_double),
        ("mu_b", C.c_double), ("tau", C.c_double), ("beta", C.c_double), ("c1", C.c_double),
        ("armijo_slack", C.c_double),
        ("max_trials", C.c_int32), ("limited_memory", C.c_int32), ("segments", C.c_int32),
        ("block_threads", C.c_int32), ("trial_group", C.c_int32), ("flags", C.c_int32),
        ("hess_floor", C.c_double),
    ]


class Stats(C.Structure):
    _fields_ = [("alpha", C.c_void_p), ("rho", C.c_void_p), ("merit0", C.c_void_p),
                ("dmerit", C.c_void_p), ("status", C.c_void_p), ("trials", C.c_void_p)]


class Direction(C.Structure):
    _fields_ = [("dX", C.c_void_p), ("dU", C.c_void_p), ("alpha_max", C.c_void_p)]


class Srbd(C.Structure):
    _fields_ = [("mass", C.c_double), ("dt", C.c_double), ("inertia", C.c_double * 9),
                ("inertia_inv", C.c_double * 9), ("W", View), ("reference", View), ("parameter", View)]


class SrbdGait(C.Structure):
    _fields_ = [("dt", C.c_double), ("mass", C.c_double), ("duty_factor", C.c_double),
                ("step_freq", C.c_double), ("step_height", C.c_double), ("clearance_speed", C.c_double),
                ("foot0", C.c_double * 12), ("advance_dt", C.c_double)]


class SolveOpts(C.Structure):
    _fields_ = [("max_iter", C.c_int32), ("check_every", C.c_int32),
                ("dcost_tol", C.c_double), ("cc_tol", C.c_double)]


class KKTError(RuntimeError):
    pass


_lib = None


def load():
    """The loaded library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise KKTError(f"{LIB_PATH} not found: build it with __graft_entry__.build() "
                       "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    lib.mpxkkt_default_params.argtypes = [P(Params)]
    lib.mpxkkt_workspace_size.argtypes = [P(Problem), P(Params), P(C.c_size_t)]
    lib.mpxkkt_sqp_iterate.argtypes = [C.c_void_p, P(Problem), P(Params), P(Iterate), P(Stats),
                                       C.c_void_p, C.c_size_t]
    lib.mpxkkt_sqp_iterate_noalloc.argtypes = [C.c_void_p, P(Problem), P(Params), P(Iterate), P(Stats)]
    lib.mpxkkt_sqp_direction.argtypes = [C.c_void_p, P(Problem), P(Params), P(Iterate), P(Direction),
                                         C.c_void_p, C.c_size_t]
    lib.mpxkkt_sqp_accept.argtypes = [C.c_void_p, P(Problem), P(Params), P(Iterate), C.c_int32, C.c_void_p,
                                      C.c_void_p, P(Stats), C.c_void_p, C.c_size_t]
    lib.mpxkkt_srbd_iterate.argtypes = [C.c_void_p, P(Problem), P(Srbd), P(Params), P(Iterate), P(Stats),
                                        C.c_void_p, C.c_size_t]
    lib.mpxkkt_srbd_reference.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, P(SrbdGait)] + \
        [C.c_void_p] * 9
    lib.mpxkkt_default_solve_opts.argtypes = [P(SolveOpts)]
    lib.mpxkkt_solve.argtypes = [C.c_void_p, P(Problem), P(Params), P(SolveOpts), P(Iterate), P(Stats),
                                 C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                 P(C.c_int32)]
    lib.mpxkkt_kkt_residual.argtypes = [C.c_void_p, P(Problem), P(Params), P(Iterate), C.c_void_p,
                                        C.c_void_p, C.c_size_t]
    lib.mpxkkt_model_eval.argtypes = [C.c_void_p, P(Problem), C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p]
    lib.mpxkkt_shift_warmstart.argtypes = [C.c_void_p, P(Problem), C.c_int32, P(Iterate),
                                           C.c_void_p, C.c_void_p]
    lib.mpxkkt_profile_enable.argtypes = [C.c_int32]
    lib.mpxkkt_profile_collect.argtypes = [P(C.c_double * PHASES), P(C.c_int32)]
    lib.mpxkkt_last_error.restype = C.c_char_p
    lib.mpxkkt_version.restype = C.c_int32
    for name in EXPORTED[:-2]:
        getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def check(status):
    if status != OK:
        msg = load().mpxkkt_last_error().decode()
        raise KKTError(f"mpx_kkt status {status}: {msg}")


def default_params(**over):
    p = Params()
    check(load().mpxkkt_default_params(C.byref(p)))
    for k, v in over.items():
        setattr(p, k, v)
    return p


def default_solve_opts(**over):
    o = SolveOpts()
    check(load().mpxkkt_default_solve_opts(C.byref(o)))
    for k, v in over.items():
        setattr(o, k, v)
    return o


def profile_enable(on=True):
    check(load().mpxkkt_profile_enable(1 if on else 0))


def profile_collect():
    """(ms per phase [linearise, LQR recursion, costates, line search] summed
    over calls, calls)."""
    ms = (C.c_double * PHASES)()
    calls = C.c_int32(0)
    check(load().mpxkkt_profile_collect(C.byref(ms), C.byref(calls)))
    return list(ms), calls.value
