"""ctypes binding of liblra_hip.so (the C ABI declared in include/lra.h).

The shared library is built in-tree by build.py.  Loading fails loudly when it is missing:
there is no CPU fallback in the product path.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import threading
import weakref

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblra_hip.so")

i64 = C.c_int64
P_D = C.c_void_p  # device pointers are passed as raw addresses
P_I = C.c_void_p

LRA_OK = 0
SKETCH = {"none": 0, "randn": 1, "srft": 2, "sub": 3, "sprn": 4}
PREC = {"f64": 0, "f32": 1, "bf16": 2, "bf16x3": 3}
RET_Q, RET_R, RET_T = 1, 2, 4
TRANS = {"n": 0, "c": 1}
PRANGE_TRANS = {"n": 0, "c": 1, "b": 2}
SIDE = {"left": 0, "right": 1}


class lra_opts(C.Structure):
    """Mirror of `lra_opts` (include/lra.h); field order matters."""
    _fields_ = [
        ("atol", C.c_double),
        ("maxdet_niter", i64),
        ("maxdet_tol", C.c_double),
        ("nb", i64),
        ("pheig_orthtol", C.c_double),
        ("pqrfact_retval", C.c_int32),
        ("rank", i64),
        ("rtol", C.c_double),
        ("sketch", C.c_int32),
        ("sketch_randn_niter", i64),
        ("sketchfact_adap", C.c_int32),
        ("randn_samp_mul", i64), ("randn_samp_add", i64),
        ("srft_samp_mul", i64), ("srft_samp_add", i64),
        ("sub_samp_mul", i64), ("sub_samp_add", i64),
        ("snorm_niter", i64),
        ("verb", C.c_int32),
        ("precision", C.c_int32),
        ("seed", C.c_uint64),
    ]


class lra_id_out(C.Structure):
    _fields_ = [("sk", P_I), ("rd", P_I), ("T", P_D), ("ldt", i64), ("k", i64),
                ("rounds", i64)]


class lra_pqr_out(C.Structure):
    _fields_ = [("Q", P_D), ("ldq", i64), ("R", P_D), ("ldr", i64), ("p", P_I), ("T", P_D),
                ("ldt", i64), ("k", i64), ("rounds", i64)]


class lra_psvd_out(C.Structure):
    _fields_ = [("U", P_D), ("ldu", i64), ("S", P_D), ("Vt", P_D), ("ldvt", i64), ("k", i64)]


class lra_pheig_out(C.Structure):
    _fields_ = [("values", P_D), ("vectors", P_D), ("ldv", i64), ("k", i64), ("kn", i64),
                ("kp", i64)]


class lra_server_out(C.Structure):
    _fields_ = [("kind", C.c_int32), ("status", C.c_int32), ("m", i64), ("n", i64), ("k", i64),
                ("sk", P_I), ("rd", P_I), ("T", P_D), ("ldt", i64), ("U", P_D), ("ldu", i64),
                ("S", P_D), ("Vt", P_D), ("ldvt", i64), ("t_submit", C.c_double),
                ("t_start", C.c_double), ("t_done", C.c_double)]


APPLY_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, i64, C.c_void_p, i64, C.c_void_p, i64,
                       C.c_void_p)

# (name, restype, argtypes) of every exported symbol declared in include/lra.h
SIGNATURES = [
    ("lra_opts_default", None, [C.POINTER(lra_opts)]),
    ("lra_ctx_create", C.c_int, [C.c_int, C.c_size_t, C.POINTER(C.c_void_p)]),
    ("lra_ctx_destroy", C.c_int, [C.c_void_p]),
    ("lra_ctx_set_stream", C.c_int, [C.c_void_p, C.c_void_p]),
    ("lra_sync", C.c_int, [C.c_void_p]),
    ("lra_last_error", C.c_char_p, []),
    ("lra_device_info", C.c_int, [C.c_void_p, C.POINTER(i64), C.POINTER(i64)]),
    ("lra_gaussian_fill", C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, i64, i64, P_D, i64]),
    ("lra_sketch_draw", C.c_int, [C.c_void_p, C.c_int, C.c_uint64, C.c_uint64, i64, i64, P_D,
                                  P_I]),
    ("lra_sketch_gaussian", C.c_int, [C.c_void_p, C.c_int, C.c_int, i64, i64, P_D, i64, i64,
                                      P_D, C.c_uint64, C.c_uint64, C.c_int, P_D, i64]),
    ("lra_sketch_srft", C.c_int, [C.c_void_p, C.c_int, i64, i64, P_D, i64, i64, P_D, P_I,
                                  C.c_uint64, C.c_uint64, P_D, i64]),
    ("lra_sketch_subset", C.c_int, [C.c_void_p, C.c_int, i64, i64, P_D, i64, i64, P_I,
                                    C.c_uint64, C.c_uint64, P_D, i64]),
    ("lra_sketch_sprn", C.c_int, [C.c_void_p, C.c_int, i64, i64, P_D, i64, i64, P_I, P_D,
                                  C.c_uint64, C.c_uint64, P_D, i64]),
    ("lra_laqps_f64", C.c_int, [C.c_void_p, i64, i64, i64, i64, C.POINTER(i64), P_D, i64,
                                P_I, P_D, P_D, P_D, P_D, P_D, i64]),
    ("lra_geqp3_adap_f64", C.c_int, [C.c_void_p, i64, i64, P_D, i64, i64, C.c_double,
                                     C.c_double, i64, P_I, P_D, C.POINTER(i64)]),
    ("lra_id_T_f64", C.c_int, [C.c_void_p, i64, i64, P_D, i64, P_D, i64]),
    ("lra_maxdet_swapcols_f64", C.c_int, [C.c_void_p, i64, i64, P_I, P_D, i64, C.c_double,
                                          i64, C.POINTER(i64)]),
    ("lra_snorm", C.c_int, [C.c_void_p, i64, i64, P_D, i64, C.c_int, C.POINTER(lra_opts),
                            P_D, C.POINTER(C.c_double)]),
    ("lra_snormdiff", C.c_int, [C.c_void_p, i64, i64, P_D, i64, P_D, i64, C.c_int,
                                C.POINTER(lra_opts), P_D, C.POINTER(C.c_double)]),
    ("lra_snormdiff_id", C.c_int, [C.c_void_p, i64, i64, P_D, i64, i64, P_I, P_I, P_D, i64,
                                   C.POINTER(lra_opts), P_D, C.POINTER(C.c_double)]),
    ("lra_snormdiff_id_noalloc", C.c_int, [C.c_void_p, i64, i64, P_D, i64, i64, P_I, P_I, P_D,
                                           i64, C.POINTER(lra_opts), P_D,
                                           C.POINTER(C.c_double)]),
    ("lra_snorm_op", C.c_int, [C.c_void_p, i64, i64, APPLY_FN, C.c_void_p, C.c_int,
                               C.POINTER(lra_opts), P_D, C.POINTER(C.c_double)]),
    ("lra_geqrf_f64", C.c_int, [C.c_void_p, i64, i64, P_D, i64, P_D]),
    ("lra_orgqr_f64", C.c_int, [C.c_void_p, i64, i64, P_D, i64, P_D, P_D, i64]),
    ("lra_idfact", C.c_int, [C.c_void_p, C.c_int, i64, i64, P_D, i64, C.POINTER(lra_opts),
                             C.POINTER(lra_id_out)]),
    ("lra_idfact_op", C.c_int, [C.c_void_p, C.c_int, i64, i64, APPLY_FN, C.c_void_p, C.c_int,
                                C.POINTER(lra_opts), C.POINTER(lra_id_out)]),
    ("lra_pqrfact_op", C.c_int, [C.c_void_p, C.c_int, i64, i64, APPLY_FN, C.c_void_p,
                                 C.c_int, C.POINTER(lra_opts), C.POINTER(lra_pqr_out)]),
    ("lra_psvdfact_op", C.c_int, [C.c_void_p, i64, i64, APPLY_FN, C.c_void_p, C.c_int,
                                  C.POINTER(lra_opts), C.POINTER(lra_psvd_out)]),
    ("lra_psvdvals_op", C.c_int, [C.c_void_p, i64, i64, APPLY_FN, C.c_void_p, C.c_int,
                                  C.POINTER(lra_opts), P_D, C.POINTER(i64)]),
    ("lra_pheigfact_op", C.c_int, [C.c_void_p, i64, APPLY_FN, C.c_void_p,
                                   C.POINTER(lra_opts), C.POINTER(lra_pheig_out)]),
    ("lra_pheigvals_op", C.c_int, [C.c_void_p, i64, APPLY_FN, C.c_void_p,
                                   C.POINTER(lra_opts), P_D, C.POINTER(i64)]),
    ("lra_curfact_op", C.c_int, [C.c_void_p, i64, i64, APPLY_FN, C.c_void_p, C.c_int,
                                 C.POINTER(lra_opts), P_I, P_I, C.POINTER(i64)]),
    ("lra_sketchfact", C.c_int, [C.c_void_p, C.c_int, i64, i64, P_D, i64, C.POINTER(lra_opts),
                                 C.c_uint64, P_I, C.POINTER(i64)]),
    ("lra_curfact", C.c_int, [C.c_void_p, i64, i64, P_D, i64, C.POINTER(lra_opts), P_I, P_I,
                              C.POINTER(i64)]),
    ("lra_pqrfact", C.c_int, [C.c_void_p, C.c_int, i64, i64, P_D, i64, C.POINTER(lra_opts),
                              C.POINTER(lra_pqr_out)]),
    ("lra_prange", C.c_int, [C.c_void_p, C.c_int, i64, i64, P_D, i64, C.POINTER(lra_opts),
                             C.POINTER(lra_pqr_out)]),
    ("lra_cur_build", C.c_int, [C.c_void_p, i64, i64, P_D, i64, C.c_void_p, C.c_void_p, i64,
                                P_D, i64, P_D, i64, P_D, i64, P_D, P_D, i64]),
    ("lra_psvdfact", C.c_int, [C.c_void_p, i64, i64, P_D, i64, C.POINTER(lra_opts),
                               C.POINTER(lra_psvd_out)]),
    ("lra_psvdvals", C.c_int, [C.c_void_p, i64, i64, P_D, i64, C.POINTER(lra_opts), P_D,
                               C.POINTER(i64)]),
    ("lra_pheigfact", C.c_int, [C.c_void_p, i64, P_D, i64, C.POINTER(lra_opts),
                                C.POINTER(lra_pheig_out)]),
    ("lra_pheigvals", C.c_int, [C.c_void_p, i64, P_D, i64, C.POINTER(lra_opts), P_D,
                                C.POINTER(i64)]),
    ("lra_gesvj_f64", C.c_int, [C.c_void_p, i64, i64, P_D, i64, P_D, i64, P_D, P_D, i64]),
    ("lra_gemm_choice", C.c_int, [C.c_int, C.c_int, C.c_int, i64, i64, i64,
                                  C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_int)]),
    ("lra_dgemm", C.c_int, [C.c_void_p, C.c_int, C.c_int, i64, i64, i64, C.c_double, P_D, i64,
                            P_D, i64, C.c_double, P_D, i64]),
    ("lra_id_apply", C.c_int, [C.c_void_p, C.c_int, i64, i64, P_I, P_I, P_D, i64, i64, P_D, i64,
                               P_D, i64]),
    ("lra_scale_rows", C.c_int, [C.c_void_p, i64, i64, P_D, P_D, i64]),
    ("lra_zsketch_draw", C.c_int, [C.c_void_p, C.c_int, C.c_uint64, C.c_uint64, i64, i64, P_D,
                                   P_I]),
    ("lra_zsketch", C.c_int, [C.c_void_p, C.c_int, C.c_int, i64, i64, P_D, i64, i64, P_D, P_I,
                              C.c_uint64, C.c_uint64, P_D, i64]),
    ("lra_zgeqp3_adap", C.c_int, [C.c_void_p, i64, i64, P_D, i64, i64, C.c_double, C.c_double,
                                  i64, P_I, P_D, C.POINTER(i64)]),
    ("lra_zid_T", C.c_int, [C.c_void_p, i64, i64, P_D, i64, P_D, i64]),
    ("lra_zidfact", C.c_int, [C.c_void_p, C.c_int, i64, i64, P_D, i64, C.POINTER(lra_opts),
                              C.POINTER(lra_id_out)]),
    ("lra_idfact_batched", C.c_int, [C.c_void_p, i64, P_I, P_I, C.c_void_p, P_I,
                                     C.POINTER(lra_opts), C.c_void_p, C.c_void_p, C.c_void_p,
                                     P_I, P_I, C.POINTER(i64)]),
    ("lra_idfact_batched_plan", C.c_int, [C.c_void_p, i64, P_I, P_I, C.c_void_p, P_I,
                                          C.POINTER(lra_opts), C.c_void_p, C.c_void_p,
                                          C.c_void_p, P_I, C.POINTER(C.c_void_p)]),
    ("lra_idfact_batched_capture", C.c_int, [C.c_void_p]),
    ("lra_idfact_batched_launch", C.c_int, [C.c_void_p]),
    ("lra_idfact_batched_finish", C.c_int, [C.c_void_p, P_I, C.POINTER(i64)]),
    ("lra_idfact_batched_plan_destroy", C.c_int, [C.c_void_p]),
    ("lra_idfact_batched_plan_widths", C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_int]),
    ("lra_server_create", C.c_int, [C.c_int, C.c_size_t, C.c_int, C.POINTER(C.c_void_p)]),
    ("lra_server_submit", C.c_int, [C.c_void_p, C.c_int, i64, i64, P_D, i64,
                                    C.POINTER(lra_opts), C.POINTER(i64)]),
    ("lra_server_poll", C.c_int, [C.c_void_p, i64, C.POINTER(C.c_int32)]),
    ("lra_server_wait", C.c_int, [C.c_void_p, i64, C.c_double, C.POINTER(C.c_int32)]),
    ("lra_server_result", C.c_int, [C.c_void_p, i64, C.POINTER(lra_server_out)]),
    ("lra_server_release", C.c_int, [C.c_void_p, i64]),
    ("lra_server_stats", C.c_int, [C.c_void_p, C.POINTER(i64), C.POINTER(i64),
                                   C.POINTER(C.c_size_t)]),
    ("lra_server_destroy", C.c_int, [C.c_void_p]),
]


class LraError(RuntimeError):
    pass


_lib = None
_lock = threading.Lock()


def load(path: str = LIB_PATH):
    """Load liblra_hip.so and declare every signature.  Raises if the library is missing."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise LraError(
                f"{path} is missing: build it with `python lowrankapprox.jl_amd/build.py` "
                "(the HIP path has no CPU fallback)")
        L = C.CDLL(path)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
        return L


E_HIP, E_ALLOC, E_UNSUPPORTED, E_TIMEOUT, E_STATE = 1, 2, 3, 4, 5


def check(status: int, what: str = "") -> None:
    if status != LRA_OK:
        msg = load().lra_last_error()
        msg = msg.decode() if msg else ""
        raise LraError(f"{what} failed with status {status}: {msg}")


def default_opts() -> lra_opts:
    o = lra_opts()
    load().lra_opts_default(C.byref(o))
    return o


# Objects holding library handles (servers, batch plans), closed before the contexts at exit.
_OWNERS = weakref.WeakSet()


def register_owner(obj) -> None:
    """Track an object with a close() method that releases library handles."""
    _OWNERS.add(obj)


def _teardown() -> None:
    """Release every library object while the HIP runtime is still loaded: servers and batch
    plans first (their worker threads and graphs use the device), then the per-device
    contexts (streams, events, workspaces).  Runs at interpreter exit, before module and
    runtime teardown, so no stream or event outlives the runtime that owns it."""
    if _lib is None:
        return
    for obj in list(_OWNERS):
        try:
            obj.close()
        except Exception:
            pass
    for dev, ctx in list(Context._ctx.items()):
        try:
            ctx.close()
        except Exception:
            pass
    Context._ctx.clear()


atexit.register(_teardown)


class Context:
    """One lra_ctx per (process, device); work is enqueued on torch's current stream."""

    _ctx: dict = {}

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = C.c_void_p()
        check(load().lra_ctx_create(device, hbm_budget, C.byref(h)), "lra_ctx_create")
        self.h = h
        self.device = device

    @classmethod
    def get(cls, device: int = 0) -> "Context":
        c = cls._ctx.get(device)
        if c is None:
            c = cls._ctx[device] = Context(device)
        return c

    def bind_stream(self, stream_ptr: int) -> None:
        check(load().lra_ctx_set_stream(self.h, C.c_void_p(stream_ptr)), "lra_ctx_set_stream")

    def sync(self) -> None:
        check(load().lra_sync(self.h), "lra_sync")

    def close(self) -> None:
        """Wait for the context's work and destroy it (lra_ctx_destroy)."""
        if self.h:
            load().lra_ctx_destroy(self.h)
            self.h = C.c_void_p()
