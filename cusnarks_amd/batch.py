"""Batched multi-exponentiation: many uneven multiexps in one planned pipeline.

Replaces the reference's chunked multiexp dispatch over GPUs and streams
(findECPointsDispatch / init_EC_P, cusnarks src/python/groth_prover.py:1548-1688;
buildDispatchTable, src/python/cuda_wrapper.py:1643-1677): the sizes are
planned once (members grouped by window width, one launch per stage and group,
groups on concurrent streams), and the enqueue sequence is captured once into
a HIP graph and replayed.  Every member's result equals the single-call
result bit for bit.
"""
from __future__ import annotations

import ctypes

from . import _capi
from ._capi import P, check, lib
from .device import dptr

G1W, G2W = 16, 32


def _ptr_array(items):
    arr = (ctypes.c_void_p * len(items))(*[dptr(x) or None for x in items])
    return arr


class MsmBatch:
    """A planned batch of multiexps of fixed sizes on one curve.

    run() takes per-member device scalars / points / outputs (DeviceBuffers or
    integer device pointers); the buffers may change between runs.
    """

    def __init__(self, sizes, g2: bool = False, window_bits=None, ctx=None, shared_workspace: bool = False):
        self.ctx = ctx or _capi.default_context()
        self.sizes = [int(x) for x in sizes]
        if not self.sizes:
            raise ValueError("empty batch")
        self.g2 = bool(g2)
        cnt = len(self.sizes)
        n = (ctypes.c_uint64 * cnt)(*self.sizes)
        wb = (ctypes.c_uint32 * cnt)(*[int(x) for x in window_bits]) if window_bits is not None else None
        if window_bits is not None and len(window_bits) != cnt:
            raise ValueError("window_bits length differs from sizes")
        self._ptr = P()
        flags = _capi.BATCH_CTX_WORKSPACE if shared_workspace else 0
        check(lib().csk_msm_batch_create(self.ctx.ptr, int(self.g2), cnt, ctypes.cast(n, P),
                                         ctypes.cast(wb, P) if wb is not None else None, flags,
                                         ctypes.byref(self._ptr)), "csk_msm_batch_create")

    @staticmethod
    def workspace_bytes(sizes, g2: bool = False, ctx=None) -> int:
        """Workspace bytes a batch of these sizes needs (planning only)."""
        ctx = ctx or _capi.default_context()
        sizes = [int(x) for x in sizes]
        n = (ctypes.c_uint64 * len(sizes))(*sizes)
        out = ctypes.c_uint64(0)
        check(lib().csk_msm_batch_bytes(ctx.ptr, int(bool(g2)), len(sizes), ctypes.cast(n, P), None,
                                        ctypes.byref(out)), "csk_msm_batch_bytes")
        return int(out.value)

    def __len__(self):
        return len(self.sizes)

    def run(self, scalars, points, outs, use_graph: bool = True, stream=None) -> None:
        """Enqueue all members (no synchronisation)."""
        cnt = len(self.sizes)
        if not (len(scalars) == len(points) == len(outs) == cnt):
            raise ValueError("need one scalar, point and output buffer per member")
        s, p, o = _ptr_array(scalars), _ptr_array(points), _ptr_array(outs)
        check(lib().csk_msm_batch_run(self._ptr, ctypes.cast(s, P), ctypes.cast(p, P), ctypes.cast(o, P),
                                      int(bool(use_graph)), stream), "csk_msm_batch_run")

    def info(self) -> dict:
        """{'bytes': device bytes, 'groups': [(c, members, chunk, entry bound), ...]}"""
        buf = (ctypes.c_uint64 * (2 + 4 * len(self.sizes)))()
        check(lib().csk_msm_batch_info(self._ptr, ctypes.cast(buf, P), len(buf)), "csk_msm_batch_info")
        ng = int(buf[0])
        groups = [tuple(int(buf[2 + 4 * g + k]) for k in range(4)) for g in range(ng)]
        return {"bytes": int(buf[1]), "groups": groups}

    def close(self) -> None:
        if self._ptr:
            lib().csk_msm_batch_destroy(self._ptr)
            self._ptr = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
