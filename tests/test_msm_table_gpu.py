"""Fixed-base tables (csk_msm_table_*): the multiexp over a table of
window-shifted points equals the plain multiexp, the reference fixtures, the
CPU oracle and, at 2^22 points, the closed form (sum k_i k'_i) G, bit for bit.
Covers both infinity encodings among the points, zero and full-width
scalars, sizes around the sort tile, every window width class, reuse of one
table for many scalar vectors, and G2."""
import ctypes

import numpy as np
import pytest

import oracle as O
from cusnarks_amd import _capi, inputs
from cusnarks_amd.device import DeviceBuffer
from cusnarks_amd.msm import MsmTable, fixed_base, msm_host
from oracle import pyoracle as P
from tests.gpu_util import dot_mod_r, g1_fixture, g2_fixture, gen_mul

pytestmark = pytest.mark.gpu


def _table_msm(scl, pts, g2=False, window_bits=0, table=None):
    ctx = _capi.default_context()
    nw = 32 if g2 else 16
    dp = DeviceBuffer.from_array(ctx, np.ascontiguousarray(pts, dtype=np.uint32).reshape(-1, nw))
    ds = DeviceBuffer.from_array(ctx, np.ascontiguousarray(scl, dtype=np.uint32).reshape(-1, 8))
    out = DeviceBuffer(ctx, nw * 4)
    t = table or MsmTable(dp, scl.shape[0], g2=g2, window_bits=window_bits, ctx=ctx)
    t.run(ds, out)
    res = out.download((nw,))
    if table is None:
        t.close()
    return res


def _rand_points(seed, n, g2=False):
    k = inputs.field_elements(seed, n, inputs.R, nonzero=True)
    gen = np.array(P.g2_to_words(P.G2_GEN) if g2 else P.g1_to_words(P.G1_GEN), np.uint32)
    return O.ec_scmul_aff(k, gen, ec2=g2)


def test_g1_fixture_every_window_class():
    scl, pts, exp = g1_fixture()  # infinity at 1023, 815 scalars >= r
    for c in (0, 2, 5, 8, 11, 13, 16, 18, 20):
        assert np.array_equal(_table_msm(scl, pts, window_bits=c), exp), c


def test_g2_fixture():
    scl, pts, exp = g2_fixture()
    for c in (0, 4, 12, 16):
        assert np.array_equal(_table_msm(scl, pts, g2=True, window_bits=c), exp), c


@pytest.mark.parametrize("n", [1, 2, 33, 4095, 4096, 4097, 3 * 4096 + 77, 70001])
def test_g1_random_vs_plain_and_oracle(n):
    pts = _rand_points(500 + n, n)
    pts[n // 2] = np.array(P.g1_to_words(None), np.uint32)  # (0, 1*R)
    if n > 3:
        pts[1] = 0  # (0, 0) zero padding
    scl = inputs.field_elements(600 + n, n)
    scl[0] = 0
    want = msm_host(scl, pts)
    if n <= 5000:
        # the CPU oracle (the reference's CPU path) reads infinity as (0, 1*R) only;
        # the GPU paths also take the reference GPU kernel's (0, 0) padding
        opts = pts.copy()
        if n > 3:
            opts[1] = np.array(P.g1_to_words(None), np.uint32)
        ref, _ = O.ec_jacreduce(scl, opts)
        assert np.array_equal(want, ref)
    assert np.array_equal(_table_msm(scl, pts), want)


def test_one_table_many_scalar_vectors_g1_and_g2():
    ctx = _capi.default_context()
    for g2, n in ((False, 20000), (True, 3000)):
        nw = 32 if g2 else 16
        pts = _rand_points(700 + n, n, g2=g2)
        dp = DeviceBuffer.from_array(ctx, pts)
        t = MsmTable(dp, n, g2=g2, ctx=ctx)
        info = t.info()
        assert info["n"] == n and info["windows"] == 254 // info["window_bits"] + 1
        assert info["table_bytes"] == MsmTable.bytes_for(n, info["window_bits"], g2=g2)
        for rep in range(3):
            scl = inputs.field_elements(800 + 10 * n + rep, n)
            if rep == 2:
                scl[:] = inputs.from_ints([inputs.R - 1])  # all scalars equal: every entry in one bucket
            got = _table_msm(scl, pts, g2=g2, table=t)
            assert np.array_equal(got, msm_host(scl, pts, g2=g2)), (g2, rep)
            assert got.shape == (nw,)
        t.close()


def test_g1_2e22_closed_form():
    ctx = _capi.default_context()
    n = 1 << 22
    kp = inputs.field_elements(901, n, nonzero=True)
    k = inputs.field_elements(902, n)
    dk = DeviceBuffer.from_array(ctx, kp)
    dp = DeviceBuffer(ctx, n * 64)
    fixed_base(dk, n, dp, ctx=ctx)
    _capi.check(_capi.lib().csk_stream_sync(ctx.ptr, None), "sync")
    dk.close()
    ds = DeviceBuffer.from_array(ctx, k)
    out = DeviceBuffer(ctx, 64)
    t = MsmTable(dp, n, ctx=ctx)
    t.run(ds, out)
    assert np.array_equal(out.download((16,)), gen_mul(dot_mod_r(k, kp)))
    t.close()


def test_bad_arguments():
    L = _capi.lib()
    ctx = _capi.default_context()
    b = _capi.U64(0)
    assert L.csk_msm_table_bytes(0, 0, 16, ctypes.byref(b)) == _capi.CSK_ERR_ARG
    assert L.csk_msm_table_bytes(0, 10, 21, ctypes.byref(b)) == _capi.CSK_ERR_ARG
    assert L.csk_msm_table_bytes(1, 4097, 16, ctypes.byref(b)) == 0 and b.value == 16 * 8192 * 128 + 8192 // 8
    t = _capi.P()
    d = DeviceBuffer(ctx, 64)
    assert L.csk_msm_table_create(ctx.ptr, 0, d.ptr, 1, 1, ctypes.byref(t), None) == _capi.CSK_ERR_ARG
    assert L.csk_msm_table_create(ctx.ptr, 0, None, 1, 8, ctypes.byref(t), None) == _capi.CSK_ERR_ARG


@pytest.mark.parametrize("g2,n,count", [(False, 4097, 1), (False, 5000, 3), (True, 3000, 5), (False, 600, 64)])
def test_run_batch_equals_single_runs(g2, n, count):
    """count scalar vectors over one table as one pipeline give the single
    runs' results, in the table's workspace and in another context's."""
    ctx = _capi.default_context()
    nw = 32 if g2 else 16
    pts = _rand_points(1100 + n, n, g2=g2)
    pts[n // 3] = np.array(P.g2_to_words(None) if g2 else P.g1_to_words(None), np.uint32)
    dp = DeviceBuffer.from_array(ctx, pts)
    t = MsmTable(dp, n, g2=g2, ctx=ctx)
    scl = [inputs.field_elements(1200 + 7 * p + n, n) for p in range(count)]
    if count > 1:
        scl[1][:] = 0  # an all-zero vector: the point at infinity
    ds = [DeviceBuffer.from_array(ctx, s) for s in scl]
    single = []
    for d in ds:
        o = DeviceBuffer(ctx, nw * 4)
        t.run(d, o)
        single.append(o.download((nw,)))
        o.close()
    other = _capi.Context(ctx.device)
    for c in (None, other):
        outs = [DeviceBuffer(ctx, nw * 4) for _ in range(count)]
        t.run_batch(ds, outs, ctx=c)
        for p, o in enumerate(outs):
            assert np.array_equal(o.download((nw,)), single[p]), (c, p)
            o.close()
    if count > 1:
        assert np.array_equal(single[1], np.array(P.g2_to_words(None) if g2 else P.g1_to_words(None), np.uint32))
    if count <= 3:
        for p in range(count):
            assert np.array_equal(single[p], msm_host(scl[p], pts, g2=g2))
    t.close()


def test_run_batch_arguments():
    ctx = _capi.default_context()
    pts = _rand_points(1300, 100)
    dp = DeviceBuffer.from_array(ctx, pts)
    t = MsmTable(dp, 100, ctx=ctx)
    d = DeviceBuffer(ctx, 100 * 32)
    o = DeviceBuffer(ctx, 64)
    t.run_batch([], [])  # nothing to do
    with pytest.raises(_capi.CskError):
        t.run_batch([d] * 65, [o] * 65)
    with pytest.raises(ValueError):
        t.run_batch([d], [])
    L = _capi.lib()
    arr = (ctypes.c_void_p * 1)(None)
    assert L.csk_msm_table_run_batch(ctx.ptr, t._t, 1, ctypes.cast(arr, _capi.P), ctypes.cast(arr, _capi.P),
                                     None) == _capi.CSK_ERR_ARG
    t.close()


@pytest.mark.parametrize("g2", [False, True])
@pytest.mark.parametrize("c", [12, 20])
def test_run_multi_over_different_tables(g2, c):
    """Runs over tables of different sizes (one window width) as one pipeline
    equal the single runs; tables of different widths are refused.  At c = 20
    nearly every bucket is empty and each run's top window (2^14 reachable
    buckets of 2^19) is followed by the next run's keys."""
    ctx = _capi.default_context()
    nw = 32 if g2 else 16
    sizes = (5000, 4096, 9000) if not g2 else (3000, 5000)
    tabs, scl, want = [], [], []
    for k, n in enumerate(sizes):
        pts = _rand_points(1400 + 10 * k + n, n, g2=g2)
        if k == 0:
            pts[7] = np.array(P.g2_to_words(None) if g2 else P.g1_to_words(None), np.uint32)
        t = MsmTable(DeviceBuffer.from_array(ctx, pts), n, g2=g2, window_bits=c, ctx=ctx)
        s = inputs.field_elements(1500 + k, n)
        tabs.append(t)
        scl.append(DeviceBuffer.from_array(ctx, s))
        want.append(msm_host(s, pts, g2=g2))
    # every table twice, in an interleaved order, with the other context's workspace
    order = list(range(len(sizes))) + list(reversed(range(len(sizes))))
    outs = [DeviceBuffer(ctx, nw * 4) for _ in order]
    MsmTable.run_multi([tabs[i] for i in order], [scl[i] for i in order], outs, ctx=_capi.Context(ctx.device))
    for i, o in zip(order, outs):
        assert np.array_equal(o.download((nw,)), want[i]), i
    other = MsmTable(DeviceBuffer.from_array(ctx, _rand_points(1600, 100, g2=g2)), 100, g2=g2, window_bits=10,
                     ctx=ctx)
    with pytest.raises(_capi.CskError):
        MsmTable.run_multi([tabs[0], other], [scl[0], scl[0]], outs[:2])
    for t in tabs + [other]:
        t.close()
