"""csk_msm_g1_noalloc: the G1 multiexp in the library's fixed scratch array.

Called through the C ABI and compared bit for bit with csk_msm_g1 on two calls
in a row (the largest allowed size, then a smaller one, so the second call
runs over the first call's leftovers in the shared array), against the oracle
on the smaller one, with free device memory unchanged across the calls; sizes
above the limit and windows that do not fit are refused with an error."""
import ctypes

import numpy as np
import pytest

import oracle
from oracle import pyoracle as P
from cusnarks_amd import _capi, inputs
from cusnarks_amd.device import DeviceBuffer
from cusnarks_amd.msm import fixed_base

pytestmark = pytest.mark.gpu

MAX_N = 1 << 15


def _free_bytes() -> int:
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _inputs(ctx, n, seed):
    k = DeviceBuffer.from_array(ctx, inputs.field_elements(seed, n, nonzero=True))
    pts = DeviceBuffer(ctx, n * 64)
    fixed_base(k, n, pts, ctx=ctx)
    _capi.check(_capi.lib().csk_stream_sync(ctx.ptr, None), "sync")
    k.close()
    return DeviceBuffer.from_array(ctx, inputs.field_elements(seed + 1, n)), pts


def _call(fn, ctx, sc, pts, n, out, cfg=None):
    return fn(ctx.ptr, sc.ptr if sc else None, pts.ptr if pts else None, n, out.ptr,
              ctypes.byref(cfg) if cfg is not None else None, None)


def test_noalloc_matches_the_allocating_entry_point():
    L = _capi.lib()
    ctx = _capi.Context(0)
    big = _inputs(ctx, MAX_N, 900)
    small = _inputs(ctx, 1000, 902)
    ref_out = DeviceBuffer(ctx, 64)
    out = DeviceBuffer(ctx, 64)
    # reference results first: their workspace and kernels are loaded before the free-memory reading
    want = []
    for (sc, pts), n in ((big, MAX_N), (small, 1000)):
        assert _call(L.csk_msm_g1, ctx, sc, pts, n, ref_out) == 0
        want.append(ref_out.download((16,)))
    _capi.check(L.csk_stream_sync(ctx.ptr, None), "sync")
    free0 = _free_bytes()
    for ((sc, pts), n), w in zip(((big, MAX_N), (small, 1000)), want):
        assert _call(L.csk_msm_g1_noalloc, ctx, sc, pts, n, out) == 0, L.csk_last_error()
        _capi.check(L.csk_stream_sync(ctx.ptr, None), "sync")
        assert np.array_equal(out.download((16,)), w), n
    assert _free_bytes() == free0
    ref, _ = oracle.ec_jacreduce(small[0].download((1000, 8)), small[1].download((1000, 16)))
    assert np.array_equal(want[1], ref)
    # the empty sum is the point at infinity (0, 1*R)
    assert _call(L.csk_msm_g1_noalloc, ctx, None, None, 0, out) == 0
    _capi.check(L.csk_stream_sync(ctx.ptr, None), "sync")
    inf = out.download((16,))
    assert not inf[:8].any() and list(inf[8:]) == list(inputs.from_ints([(1 << 256) % P.Q])[0])


def test_noalloc_refuses_what_does_not_fit():
    L = _capi.lib()
    ctx = _capi.default_context()
    sc, pts = _inputs(ctx, 64, 904)
    out = DeviceBuffer(ctx, 64)
    assert _call(L.csk_msm_g1_noalloc, ctx, sc, pts, MAX_N + 1, out) == _capi.CSK_ERR_ARG
    assert b"limit" in L.csk_last_error()
    cfg = _capi.MsmCfg(20, 0, 0, 0)  # 2^19 buckets per window: far beyond the array
    assert _call(L.csk_msm_g1_noalloc, ctx, sc, pts, 64, out, cfg) == _capi.CSK_ERR_ARG
    assert b"scratch" in L.csk_last_error()
    # an explicit width that fits is honoured and gives the same point
    cfg = _capi.MsmCfg(9, 0, 0, 0)
    ref = DeviceBuffer(ctx, 64)
    assert _call(L.csk_msm_g1, ctx, sc, pts, 64, ref) == 0
    assert _call(L.csk_msm_g1_noalloc, ctx, sc, pts, 64, out, cfg) == 0
    _capi.check(L.csk_stream_sync(ctx.ptr, None), "sync")
    assert np.array_equal(out.download((16,)), ref.download((16,)))
