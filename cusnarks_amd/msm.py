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
