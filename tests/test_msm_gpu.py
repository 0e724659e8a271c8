"""GPU multi-exponentiation parity: HIP Pippenger vs the reference fixtures and
the CPU oracle (restated ec_jacreduce_server_h), bit-exact affine outputs."""
import numpy as np
import pytest

import oracle as O
from cusnarks_amd import inputs
from cusnarks_amd.msm import msm_host
from oracle import pyoracle as P
from tests.gpu_util import g1_fixture, g2_fixture

pytestmark = pytest.mark.gpu

INF_G1 = np.array(P.g1_to_words(None), dtype=np.uint32)
INF_G2 = np.array(P.g2_to_words(None), dtype=np.uint32)


def _g1_gen():
    return np.array(P.g1_to_words(P.G1_GEN), dtype=np.uint32)


def _g2_gen():
    return np.array(P.g2_to_words(P.G2_GEN), dtype=np.uint32)


def rand_points(seed, n, g2=False):
    k = inputs.field_elements(seed, n, inputs.R, nonzero=True)
    return O.ec_scmul_aff(k, _g2_gen() if g2 else _g1_gen(), ec2=g2)


def test_g1_fixture():
    scl, pts, exp = g1_fixture()
    out = msm_host(scl, pts)
    assert np.array_equal(out, exp)


def test_g2_fixture():
    scl, pts, exp = g2_fixture()
    out = msm_host(scl, pts, g2=True)
    assert np.array_equal(out, exp)


@pytest.mark.parametrize("window_bits", [0, 2, 4, 7, 9, 10, 13, 16, 18, 20])
@pytest.mark.parametrize("chunk", [0, 1, 5])
def test_g1_fixture_configs(window_bits, chunk):
    scl, pts, exp = g1_fixture()
    out = msm_host(scl, pts, window_bits=window_bits, chunk=chunk)
    assert np.array_equal(out, exp)


@pytest.mark.parametrize("window_bits", [0, 3, 8, 11, 15, 16, 17, 18])
def test_g2_fixture_configs(window_bits):
    scl, pts, exp = g2_fixture()
    out = msm_host(scl, pts, g2=True, window_bits=window_bits)
    assert np.array_equal(out, exp)


@pytest.mark.parametrize("n", [1, 2, 3, 31, 257, 2048, 3 * 4096 + 77, 1 << 17])
def test_g1_random_vs_oracle(n):
    pts = rand_points(100 + n, n)
    scl = inputs.field_elements(200 + n, n)
    exp, _ = O.ec_jacreduce(scl, pts)
    assert np.array_equal(msm_host(scl, pts), exp)


@pytest.mark.parametrize("n", [1, 5, 300])
def test_g2_random_vs_oracle(n):
    pts = rand_points(300 + n, n, g2=True)
    scl = inputs.field_elements(400 + n, n)
    exp, _ = O.ec_jacreduce(scl, pts, ec2=True)
    assert np.array_equal(msm_host(scl, pts, g2=True), exp)


def test_empty_and_zero():
    pts = rand_points(7, 16)
    assert np.array_equal(msm_host(np.zeros((0, 8), np.uint32), np.zeros((0, 16), np.uint32)), INF_G1)
    assert np.array_equal(msm_host(np.zeros((16, 8), np.uint32), pts), INF_G1)
    g2p = rand_points(8, 4, g2=True)
    assert np.array_equal(msm_host(np.zeros((4, 8), np.uint32), g2p, g2=True), INF_G2)


def test_doubling_and_cancellation_in_bucket():
    # the same point twice with the same scalar lands twice in one bucket
    # (P + P inside the mixed addition), and P, -P with equal scalars cancel
    pts = rand_points(9, 4)
    neg = pts.copy()
    y = P.from_limbs(neg[0, 8:16])
    neg[0, 8:16] = P.to_limbs((P.Q - y) % P.Q)
    scl = inputs.field_elements(10, 4)
    s = np.stack([scl[1], scl[1], scl[2], scl[2], scl[3]])
    p = np.stack([pts[1], pts[1], pts[0], neg[0], pts[3]])
    exp, _ = O.ec_jacreduce(s, p)
    for c in (0, 3, 9):
        assert np.array_equal(msm_host(s, p, window_bits=c), exp)
    # only the cancelling pair: the result is infinity
    assert np.array_equal(msm_host(s[2:4], p[2:4]), INF_G1)


def test_infinity_inputs():
    n = 64
    pts = rand_points(11, n)
    scl = inputs.field_elements(12, n)
    inf_idx = [0, 5, 63]
    pts_inf = pts.copy()
    pts_inf[inf_idx] = INF_G1  # the reference's (0, 1*R) encoding
    keep = [i for i in range(n) if i not in inf_idx]
    exp, _ = O.ec_jacreduce(scl[keep], pts[keep])
    assert np.array_equal(msm_host(scl, pts_inf), exp)
    pts_zero = pts.copy()
    pts_zero[inf_idx] = 0  # zero padding is read as infinity too
    assert np.array_equal(msm_host(scl, pts_zero), exp)


def test_full_width_scalars():
    # 256-bit scalars >= r are reduced mod r (every G1 point has order r)
    n = 200
    pts = rand_points(13, n)
    rng = np.random.default_rng(14)
    scl = rng.integers(0, 2**32, size=(n, 8), dtype=np.uint64).astype(np.uint32)
    scl[:, 7] |= 0x80000000
    exp, _ = O.ec_jacreduce(scl, pts)
    assert np.array_equal(msm_host(scl, pts), exp)


def test_skewed_single_bucket():
    # every scalar equal: one bucket per window holds all points, and with a
    # chunk of 3 entries each bucket spans many chunks (long carry chains)
    n = 500
    pts = rand_points(15, n)
    k = inputs.field_elements(16, 1)
    scl = np.repeat(k, n, axis=0)
    exp, _ = O.ec_jacreduce(scl, pts)
    assert np.array_equal(msm_host(scl, pts, chunk=3), exp)
    assert np.array_equal(msm_host(scl, pts), exp)


def test_small_scalars():
    n = 1000
    pts = rand_points(17, n)
    scl = np.zeros((n, 8), np.uint32)
    scl[:, 0] = np.arange(n) % 5  # many zeros and tiny digits
    exp, _ = O.ec_jacreduce(scl, pts)
    assert np.array_equal(msm_host(scl, pts), exp)


def test_window_width_above_limit_is_an_error():
    from cusnarks_amd._capi import CskError
    scl, pts, _ = g1_fixture()
    with pytest.raises(CskError):
        msm_host(scl, pts, window_bits=21)


def test_doubling_and_cancellation_in_bucket_g2():
    # G2 twin of the G1 case: P + P inside the mixed addition, P + (-P) -> infinity
    pts = rand_points(19, 4, g2=True)
    neg = pts.copy()
    for k in range(2):  # y = (y0, y1) -> (q - y0, q - y1)
        y = P.from_limbs(neg[0, 16 + 8 * k:24 + 8 * k])
        neg[0, 16 + 8 * k:24 + 8 * k] = P.to_limbs((P.Q - y) % P.Q)
    scl = inputs.field_elements(20, 4)
    s = np.stack([scl[1], scl[1], scl[2], scl[2], scl[3]])
    p = np.stack([pts[1], pts[1], pts[0], neg[0], pts[3]])
    exp, _ = O.ec_jacreduce(s, p, ec2=True)
    for c in (0, 3, 9):
        assert np.array_equal(msm_host(s, p, g2=True, window_bits=c), exp)
    assert np.array_equal(msm_host(s[2:4], p[2:4], g2=True), INF_G2)


@pytest.mark.parametrize("g2", [False, True])
@pytest.mark.parametrize("chunk", [1, 2, 5])
def test_repeated_point_carries(g2, chunk):
    # one point under one scalar: every chunk's run of a bucket is a multiple
    # of P, so the carry rounds and the fix-up add equal points (the doubling
    # branch of the general addition), and runs of two or more double inside
    # the bucket loop
    n = 96
    pts = np.repeat(rand_points(31 + g2, 1, g2=g2), n, axis=0)
    scl = np.repeat(inputs.field_elements(33, 1), n, axis=0)
    exp, _ = O.ec_jacreduce(scl, pts, ec2=g2)
    assert np.array_equal(msm_host(scl, pts, g2=g2, chunk=chunk), exp)


def test_infinity_inputs_g2():
    n = 24
    pts = rand_points(35, n, g2=True)
    scl = inputs.field_elements(36, n)
    inf_idx = [0, 7, 23]
    keep = [i for i in range(n) if i not in inf_idx]
    exp, _ = O.ec_jacreduce(scl[keep], pts[keep], ec2=True)
    pts_inf = pts.copy()
    pts_inf[inf_idx] = INF_G2
    assert np.array_equal(msm_host(scl, pts_inf, g2=True), exp)
    pts_zero = pts.copy()
    pts_zero[inf_idx] = 0
    assert np.array_equal(msm_host(scl, pts_zero, g2=True), exp)
