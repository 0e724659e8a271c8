"""ctypes binding of libcsk.so (the C ABI declared in include/csk.h).

The product path has no CPU fallback: if the HIP library is missing or no
GPU is visible, calls raise CskError.  Device pointers are plain integers.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcsk.so")

CSK_OK = 0
CSK_ERR_ARG = -1
CSK_ERR_HIP = -2
CSK_ERR_NOMEM = -3
CSK_ERR_NODEV = -4
BATCH_CTX_WORKSPACE = 1

OP_MULM, OP_ADDM, OP_SUBM, OP_TOMONT, OP_FROMMONT, OP_INVM, OP_MULM_EXT = range(7)


class CskError(RuntimeError):
    """Raised for every failure of the native library (never exits the process)."""


class MsmCfg(ctypes.Structure):
    _fields_ = [("window_bits", ctypes.c_uint32), ("chunk", ctypes.c_uint32),
                ("reduce_seg", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


_lib = None
_lock = threading.Lock()

P = ctypes.c_void_p
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
INT = ctypes.c_int

_SIGS = {
    "csk_ctx_create": ([INT, ctypes.POINTER(P)], INT),
    "csk_ctx_destroy": ([P], None),
    "csk_last_error": ([], ctypes.c_char_p),
    "csk_device_count": ([ctypes.POINTER(INT)], INT),
    "csk_ctx_set_timing": ([P, INT], INT),
    "csk_last_stage_times": ([P, ctypes.POINTER(ctypes.c_float), INT], INT),
    "csk_last_msm_plan": ([P, ctypes.POINTER(ctypes.c_uint32), INT], INT),
    "csk_malloc": ([P, ctypes.POINTER(P), U64], INT),
    "csk_free": ([P, P], INT),
    "csk_memcpy_h2d": ([P, P, P, U64, P], INT),
    "csk_memcpy_d2h": ([P, P, P, U64, P], INT),
    "csk_memcpy_d2d": ([P, P, P, U64, P], INT),
    "csk_stream_sync": ([P, P], INT),
    "csk_host_alloc": ([P, ctypes.POINTER(P), U64], INT),
    "csk_host_free": ([P, P], INT),
    "csk_msm_g1": ([P, P, P, U64, P, P, P], INT),
    "csk_msm_g2": ([P, P, P, U64, P, P, P], INT),
    "csk_msm_g1_host": ([P, P, P, U64, P, P], INT),
    "csk_msm_g2_host": ([P, P, P, U64, P, P], INT),
    "csk_msm_g1_noalloc": ([P, P, P, U64, P, P, P], INT),
    "csk_field_op": ([P, INT, U32, P, P, P, U64, P], INT),
    "csk_ec_jac2aff": ([P, U32, P, U64, P, U32, P], INT),
    "csk_ec_check": ([P, U32, P, U64, U32, ctypes.POINTER(U64), P], INT),
    "csk_fixed_base_g1": ([P, P, U64, P, P], INT),
    "csk_fixed_base_g2": ([P, P, U64, P, P], INT),
    "csk_ntt": ([P, P, P, U32, INT, INT, P], INT),
    "csk_interpolandmul": ([P, P, P, U32, P, P], INT),
    "csk_interpolandmul_batch": ([P, U32, P, P, U32, P, P], INT),
    "csk_mpoly_eval_csr": ([P, P, P, P, P, U32, P, P], INT),
    "csk_valu_probe": ([P, U32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)], INT),
    "csk_msm_set_tuning": ([P, INT, P, U32], INT),
    "csk_msm_batch_create": ([P, INT, U32, P, P, U32, ctypes.POINTER(P)], INT),
    "csk_msm_batch_bytes": ([P, INT, U32, P, P, ctypes.POINTER(U64)], INT),
    "csk_stream_create": ([P, ctypes.POINTER(P)], INT),
    "csk_stream_destroy": ([P, P], INT),
    "csk_stream_wait_event": ([P, P, P], INT),
    "csk_event_create": ([P, ctypes.POINTER(P)], INT),
    "csk_event_destroy": ([P, P], INT),
    "csk_event_record": ([P, P, P], INT),
    "csk_event_query": ([P, P], INT),
    "csk_event_sync": ([P, P], INT),
    "csk_event_elapsed": ([P, P, P, ctypes.POINTER(ctypes.c_float)], INT),
    "csk_msm_batch_run": ([P, P, P, P, INT, P], INT),
    "csk_msm_batch_info": ([P, P, INT], INT),
    "csk_msm_table_create": ([P, INT, P, U64, U32, ctypes.POINTER(P), P], INT),
    "csk_msm_table_run": ([P, P, P, P], INT),
    "csk_msm_table_run_batch": ([P, P, U32, P, P, P], INT),
    "csk_msm_table_run_multi": ([P, U32, P, P, P, P], INT),
    "csk_msm_table_bytes": ([INT, U64, U32, ctypes.POINTER(U64)], INT),
    "csk_msm_table_info": ([P, ctypes.POINTER(U64), INT], INT),
    "csk_msm_table_destroy": ([P], None),
    "csk_msm_batch_destroy": ([P], None),
}

# symbols of subsystems that may be absent from an older build
_OPTIONAL: dict = {}


def lib():
    """Load libcsk.so once; raise CskError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise CskError("libcsk.so is not built (run cusnarks_amd.build.build()); "
                           "there is no CPU fallback for the GPU path")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        for name, (args, res) in _OPTIONAL.items():
            fn = getattr(L, name, None)
            if fn is not None:
                fn.argtypes = args
                fn.restype = res
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    """Names of the C-ABI symbols this binding knows about that the library exports."""
    L = lib()
    return [n for n in list(_SIGS) + list(_OPTIONAL) if hasattr(L, n)]


def check(rc: int, what: str) -> None:
    if rc != CSK_OK:
        msg = lib().csk_last_error()
        raise CskError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))


class Context:
    """Owns a csk_ctx (device binding + scratch memory)."""

    def __init__(self, device: int = 0):
        self._ptr = P()
        check(lib().csk_ctx_create(device, ctypes.byref(self._ptr)), "csk_ctx_create")
        self.device = device

    @property
    def ptr(self):
        return self._ptr

    def close(self) -> None:
        if self._ptr:
            lib().csk_ctx_destroy(self._ptr)
            self._ptr = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: dict[int, Context] = {}


def default_context(device: int = 0) -> Context:
    """The shared context of a device, with the persisted multiexp tuning installed."""
    ctx = _default_ctx.get(device)
    if ctx is None:
        ctx = Context(device)
        from .tune import apply
        apply(ctx)
        _default_ctx[device] = ctx
    return ctx
