"""Multi-exponentiation front end (G1 / G2 Pippenger on the GPU).

Layouts follow cusnarks: scalars (n, 8) uint32 normal form; G1 points (n, 16)
uint32 = affine Montgomery [X | Y]; G2 points (n, 32) = [X.c0|X.c1|Y.c0|Y.c1].
Results are affine Montgomery points, infinity encoded (0, 1*R).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi
from ._capi import MsmCfg, P, check, lib
from .device import dptr


def _cfg(window_bits=0, chunk=0, reduce_seg=0):
    if not (window_bits or chunk or reduce_seg):
        return None
    return ctypes.byref(MsmCfg(window_bits, chunk, reduce_seg, 0))


def msm_host(scalars: np.ndarray, points: np.ndarray, g2: bool = False, ctx=None, **cfg) -> np.ndarray:
    """sum k_i P_i for host arrays (copies in and out); returns (16,) or (32,) uint32."""
    ctx = ctx or _capi.default_context()
    nw = 32 if g2 else 16
    s = np.ascontiguousarray(scalars, dtype=np.uint32).reshape(-1, 8)
    p = np.ascontiguousarray(points, dtype=np.uint32).reshape(-1, nw)
    if s.shape[0] != p.shape[0]:
        raise ValueError("scalars and points differ in length")
    out = np.zeros(nw, dtype=np.uint32)
    fn = lib().csk_msm_g2_host if g2 else lib().csk_msm_g1_host
    check(fn(ctx.ptr, s.ctypes.data_as(P), p.ctypes.data_as(P), s.shape[0], out.ctypes.data_as(P),
             _cfg(**cfg)), "csk_msm_host")
    return out


def msm_device(d_scalars, d_points, n: int, d_out, g2: bool = False, ctx=None, stream=None, **cfg) -> None:
    """Enqueue sum k_i P_i on device-resident data (no synchronisation)."""
    ctx = ctx or _capi.default_context()
    fn = lib().csk_msm_g2 if g2 else lib().csk_msm_g1
    check(fn(ctx.ptr, dptr(d_scalars), dptr(d_points), int(n), dptr(d_out), _cfg(**cfg), stream), "csk_msm")


def fixed_base(d_scalars, n: int, d_out, g2: bool = False, ctx=None, stream=None) -> None:
    """k_i * G for the G1 generator (1, 2) or the standard G2 generator."""
    ctx = ctx or _capi.default_context()
    fn = lib().csk_fixed_base_g2 if g2 else lib().csk_fixed_base_g1
    check(fn(ctx.ptr, dptr(d_scalars), int(n), dptr(d_out), stream), "csk_fixed_base")


def valu_probe(kind: int = 0, ctx=None) -> tuple[float, float]:
    """(ops/s, ms): kind 0 = v_mad_u64_u32 rate, kind 1 = Fq Montgomery products/s."""
    ctx = ctx or _capi.default_context()
    rate, ms = ctypes.c_double(0), ctypes.c_double(0)
    check(lib().csk_valu_probe(ctx.ptr, kind, ctypes.byref(rate), ctypes.byref(ms)), "csk_valu_probe")
    return rate.value, ms.value


class MsmTable:
    """Fixed-base table of n device-resident points (csk_msm_table_*): the
    window-shifted copies 2^(c w) P_i, so every multiexp over these points
    accumulates all windows into one set of buckets and reduces it once.
    Tables of proving-key arrays replace the reference's precomputed EC tables
    (cusnarks src/python/groth_setup.py:802-945, src/cuda/ec.cpp:1837-1979)."""

    def __init__(self, d_points, n: int, g2: bool = False, window_bits: int = 0, ctx=None, stream=None):
        self.ctx = ctx or _capi.default_context()
        self.n, self.g2 = int(n), bool(g2)
        self._t = P()
        check(lib().csk_msm_table_create(self.ctx.ptr, 1 if g2 else 0, dptr(d_points), self.n, int(window_bits),
                                         ctypes.byref(self._t), stream), "csk_msm_table_create")

    def run(self, d_scalars, d_out, stream=None) -> None:
        """Enqueue out = sum_i k_i P_i (no synchronisation)."""
        check(lib().csk_msm_table_run(self._t, dptr(d_scalars), dptr(d_out), stream), "csk_msm_table_run")

    def run_batch(self, scalars, outs, ctx=None, stream=None) -> None:
        """Enqueue outs[p] = sum_i scalars[p]_i P_i for up to 64 scalar
        vectors as one pipeline (csk_msm_table_run_batch), in the workspace of
        ctx (default: the table's context); no synchronisation."""
        if len(scalars) != len(outs):
            raise ValueError("one output per scalar vector")
        cnt = len(scalars)
        s = (ctypes.c_void_p * max(cnt, 1))(*[dptr(x) for x in scalars])
        o = (ctypes.c_void_p * max(cnt, 1))(*[dptr(x) for x in outs])
        check(lib().csk_msm_table_run_batch(ctx.ptr if ctx is not None else None, self._t, cnt,
                                            ctypes.cast(s, P), ctypes.cast(o, P), stream),
              "csk_msm_table_run_batch")

    @staticmethod
    def run_multi(tables, scalars, outs, ctx=None, stream=None) -> None:
        """Enqueue outs[p] = multiexp of scalars[p] over tables[p] for up to
        64 runs over tables of one curve and window width, as one pipeline
        (csk_msm_table_run_multi), in ctx's workspace (default: the first
        table's context)."""
        if not (len(tables) == len(scalars) == len(outs)):
            raise ValueError("one table, scalar vector and output per run")
        cnt = len(tables)
        if cnt == 0:
            return
        t = (ctypes.c_void_p * cnt)(*[tb._t.value for tb in tables])
        s = (ctypes.c_void_p * cnt)(*[dptr(x) for x in scalars])
        o = (ctypes.c_void_p * cnt)(*[dptr(x) for x in outs])
        c = ctx or tables[0].ctx
        check(lib().csk_msm_table_run_multi(c.ptr, cnt, ctypes.cast(t, P), ctypes.cast(s, P), ctypes.cast(o, P),
                                            stream), "csk_msm_table_run_multi")

    def info(self) -> dict:
        v = (_capi.U64 * 5)()
        check(lib().csk_msm_table_info(self._t, v, 5), "csk_msm_table_info")
        return {"n": v[0], "window_bits": v[1], "windows": v[2], "table_bytes": v[3], "workspace_bytes": v[4]}

    @staticmethod
    def bytes_for(n: int, window_bits: int, g2: bool = False) -> int:
        b = _capi.U64(0)
        check(lib().csk_msm_table_bytes(1 if g2 else 0, int(n), int(window_bits), ctypes.byref(b)),
              "csk_msm_table_bytes")
        return b.value

    def close(self) -> None:
        if self._t:
            lib().csk_msm_table_destroy(self._t)
            self._t = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
