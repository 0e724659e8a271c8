"""Parity at the benchmark sizes through size-independent properties.

The multiexp bench points are k'_i * G with known k'_i, so
sum_i k_i (k'_i G) = (sum_i k_i k'_i mod r) * G: the scalar is formed on the
host from the oracle's Montgomery products (a * b * R^-1) summed limb-wise,
and the one-point product comes from the oracle.  This pins the full 2^24 G1
multiexp (and a 2^20 G2 one) to a closed form, bit for bit."""
import numpy as np
import pytest

import oracle as O
from cusnarks_amd import _capi, inputs
from cusnarks_amd.device import DeviceBuffer
from cusnarks_amd.msm import fixed_base, msm_device
from cusnarks_amd.poly import ntt_device
from oracle import pyoracle as P
from tests.gpu_util import dot_mod_r

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("g2,lg", [(False, 24), (True, 20)])
def test_msm_full_size_closed_form(g2, lg):
    ctx = _capi.default_context()
    n = 1 << lg
    w = 32 if g2 else 16
    kp = inputs.field_elements(0, n, nonzero=True)     # the bench's point seeds
    k = inputs.field_elements(1, n)                    # and scalar seeds
    d_kp = DeviceBuffer.from_array(ctx, kp)
    d_p = DeviceBuffer(ctx, n * w * 4)
    fixed_base(d_kp, n, d_p, g2=g2, ctx=ctx)
    d_kp.close()
    d_k = DeviceBuffer.from_array(ctx, k)
    out = DeviceBuffer(ctx, w * 4)
    msm_device(d_k, d_p, n, out, g2=g2, ctx=ctx)
    got = out.download((w,))
    s = dot_mod_r(k, kp)
    gen = np.array(P.g2_to_words(P.G2_GEN) if g2 else P.g1_to_words(P.G1_GEN), np.uint32)
    exp = O.ec_scmul_aff(inputs.from_ints([s]), gen, ec2=g2)[0]
    assert np.array_equal(got, exp)
    # split-sum: the two halves add up to the whole (one 2-point multiexp)
    h = n // 2
    parts = DeviceBuffer(ctx, 2 * w * 4)
    msm_device(d_k, d_p, h, parts.ptr, g2=g2, ctx=ctx)
    msm_device(d_k.ptr + h * 32, d_p.ptr + h * w * 4, n - h, parts.ptr + w * 4, g2=g2, ctx=ctx)
    ones = DeviceBuffer.from_array(ctx, inputs.from_ints([1, 1]))
    msm_device(ones, parts, 2, out, g2=g2, ctx=ctx)
    assert np.array_equal(out.download((w,)), exp)


def test_ntt_2e24_roundtrip_and_linearity():
    ctx = _capi.default_context()
    lg = 24
    n = 1 << lg
    a = inputs.field_elements(2, n)                    # Montgomery data: any element of [0, r)
    da = DeviceBuffer.from_array(ctx, a)
    fa = DeviceBuffer(ctx, n * 32)
    back = DeviceBuffer(ctx, n * 32)
    ntt_device(da, fa, lg, ctx=ctx)
    ntt_device(fa, back, lg, inverse=True, fmt=1, ctx=ctx)
    assert np.array_equal(back.download((n, 8)), a)
    # X[0] = sum a_i and, for a delta at index 1, X[k] = w^k: spot-check both
    fa_h = fa.download((n, 8))
    tot = sum(int(v) << (32 * i) for i, v in enumerate(a.astype(np.uint64).sum(axis=0))) % P.RR
    assert inputs.to_ints(fa_h[:1])[0] == tot
    delta = np.zeros((n, 8), np.uint32)
    delta[1] = inputs.from_ints([P.to_mont(1, P.RR)])[0]
    dd = DeviceBuffer.from_array(ctx, delta)
    ntt_device(dd, fa, lg, ctx=ctx)
    w = P.root_of_unity(lg)
    idx = [0, 1, 2, 12345, n // 2, n - 1]
    got = fa.download((n, 8))[idx]
    assert [P.from_mont(v, P.RR) for v in inputs.to_ints(got)] == [pow(w, i, P.RR) for i in idx]


@pytest.mark.parametrize("g2,lg,widths", [(False, 24, (16, 18, 19, 20)), (True, 20, (14, 16, 18))])
def test_msm_wide_windows_closed_form(g2, lg, widths):
    """Window widths whose top window keeps few scalar bits (254 mod c: 2 at
    c = 18 and 14, 7 at 19) or sorts by fewer low bits than the others (16,
    20) at the benchmark sizes, where the sort's first pass bins by 8 bits:
    the plain multiexp and table runs (the default 20-bit table of 2^24 G1
    points, an 18-bit one of 2^20) give the closed form."""
    from cusnarks_amd.msm import MsmTable
    ctx = _capi.default_context()
    n = 1 << lg
    w = 32 if g2 else 16
    kp = inputs.field_elements(0, n, nonzero=True)
    k = inputs.field_elements(1, n)
    d_kp = DeviceBuffer.from_array(ctx, kp)
    d_p = DeviceBuffer(ctx, n * w * 4)
    fixed_base(d_kp, n, d_p, g2=g2, ctx=ctx)
    d_kp.close()
    d_k = DeviceBuffer.from_array(ctx, k)
    out = DeviceBuffer(ctx, w * 4)
    gen = np.array(P.g2_to_words(P.G2_GEN) if g2 else P.g1_to_words(P.G1_GEN), np.uint32)
    exp = O.ec_scmul_aff(inputs.from_ints([dot_mod_r(k, kp)]), gen, ec2=g2)[0]
    for c in widths:
        msm_device(d_k, d_p, n, out, g2=g2, ctx=ctx, window_bits=c)
        assert np.array_equal(out.download((w,)), exp), c
    if not g2:  # the default table width from 3 * 2^22 G1 points: 20 bits, 13 windows
        t = MsmTable(d_p, n, ctx=ctx)
        assert t.info()["window_bits"] == 20 and t.info()["windows"] == 13
        t.run(d_k, out)
        assert np.array_equal(out.download((w,)), exp)
        t.close()
    m = 1 << 20  # a table over the first 2^20 points
    t = MsmTable(d_p, m, g2=g2, window_bits=18, ctx=ctx)
    t.run(d_k, out)
    exp_m = O.ec_scmul_aff(inputs.from_ints([dot_mod_r(k[:m], kp[:m])]), gen, ec2=g2)[0]
    assert np.array_equal(out.download((w,)), exp_m)
    t.close()
