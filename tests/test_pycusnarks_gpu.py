"""The drop-in module on the GPU, called the way the reference's callers call
pycusnarks: the kernel-launch dicts ec_mad_cuda2 builds over init_EC_P-packed
vectors (src/python/cuda_wrapper.py:216-284, src/python/groth_prover.py:1548-1688),
the host helpers, and the raw C entry points they reach (csk_field_op INVM /
MULM_EXT, csk_ec_jac2aff).  Every result is compared bit for bit with the
oracle."""
import math
import random

import numpy as np
import pytest

import oracle as O
from cusnarks_amd import _capi, inputs
from cusnarks_amd import pycusnarks as pc
from cusnarks_amd.poly import CSR
from oracle import pyoracle as P

pytestmark = pytest.mark.gpu


# ------------------------------------------------------------------ caller emulation
def _shfl_blockD(nsamples, max_block_size=8):
    """get_shfl_blockD (cuda_wrapper.py:1570-1587), restated for the test."""
    lg = max(math.ceil(math.log2(nsamples)), 5)
    nb = math.ceil(lg / max_block_size)
    lpb = math.ceil(lg / nb)
    blk, rem = [1 << lpb], lg - lpb
    while rem > lpb:
        blk.append(1 << lpb)
        rem -= lpb
    if rem:
        blk.append(1 << max(rem, 5))
    return sorted(blk)[::-1]


def _mad_call(obj, vector, ec2, separate_k, grouping=6, stream_id=0):
    """The kernelLaunch call of ec_mad_cuda2 (cuda_wrapper.py:216-284) with
    shamir_en = 1, as compute_proof_ecp makes it (groth_prover.py:1509-1522)."""
    outdims = pc.ECP2_JAC_OUTDIMS if ec2 else pc.ECP_JAC_OUTDIMS
    indims_e = (pc.ECP2_JAC_INDIMS if ec2 else pc.ECP_JAC_INDIMS) + pc.U256_NDIMS
    nsamples = len(vector) // indims_e
    blk = _shfl_blockD(math.ceil(nsamples / grouping))
    if separate_k:
        blk = [128] + blk
    nk = len(blk)
    params = {"stride": [outdims] * nk, "premul": [0] * nk, "premod": [0] * nk, "midx": [0] * nk,
              "in_length": [nsamples * indims_e] * nk, "out_length": outdims, "padding_idx": [1] * nk}
    params["stride"][0] = indims_e * grouping
    cfg = {"blockD": blk, "gridD": [0] * nk}
    cfg["gridD"][-1] = 1
    if separate_k:
        red = pc.CB_EC2_JAC_RED if ec2 else pc.CB_EC_JAC_RED
        mul = pc.CB_EC2_JAC_MUL_OPT if ec2 else pc.CB_EC_JAC_MUL_OPT
        cfg["kernel_idx"] = [mul] + [red] * (nk - 1)
        params["in_length"][1] = nsamples * outdims // grouping
    else:
        cfg["kernel_idx"] = [pc.CB_EC2_JAC_MAD_SHFL if ec2 else pc.CB_EC_JAC_MAD_SHFL] * nk
        params["premul"][0] = 1
    cfg["smemS"] = [0] * nk
    res, _ = obj.kernelLaunch(vector, cfg, params, 0, stream_id, n_kernels=nk)
    if stream_id:
        res, _ = obj.streamSync(0, stream_id)
    return res


def _init_ec_p(batch, ec2):
    """init_EC_P (groth_prover.py:1548-1561): zero vector, scalar 1 in the last slot."""
    step = 4 if ec2 else 2
    ns = int(np.prod(_shfl_blockD(batch)))
    v = np.zeros((ns * (step + 1), 8), np.uint32)
    v[ns - 1, 0] = 1
    return ns, v, ns - batch, ns + step * (ns - batch)


def _points(n, seed, ec2):
    k = inputs.field_elements(seed, n, inputs.R, nonzero=True)
    gen = np.array(P.g2_to_words(P.G2_GEN) if ec2 else P.g1_to_words(P.G1_GEN), np.uint32)
    return O.ec_scmul_aff(k, gen, ec2=ec2)


@pytest.mark.parametrize("ec2", [False, True])
@pytest.mark.parametrize("separate_k", [0, 1])
def test_kernel_launch_chunked_multiexp(ec2, separate_k):
    # two chunks of 700 and 300 points; the second chunk's last slot carries
    # (1, first chunk's affine result), as findECPointsDispatch chains them
    step = 4 if ec2 else 2
    n1, n2 = 700, 300
    pts = _points(n1 + n2, 60 + ec2, ec2)
    scl = inputs.field_elements(70 + separate_k, n1 + n2)
    obj = (pc.EC2BN128 if ec2 else pc.ECBN128)(1024, seed=0)
    prev = np.zeros((step, 8), np.uint32)
    for a, b, sid in ((0, n1, 0), (n1, n1 + n2, 2)):
        batch = b - a + 1
        ns, v, s0, p0 = _init_ec_p(batch, ec2)
        v[s0:ns - 1] = scl[a:b]
        v[p0:v.shape[0] - step] = pts[a:b].reshape(-1, 8)
        v[v.shape[0] - step:] = prev
        res = _mad_call(obj, v, ec2, separate_k, stream_id=sid)
        assert res.shape == (6 if ec2 else 3, 8)
        aff = (pc.ec2_jac2aff_h if ec2 else pc.ec_jac2aff_h)(res.reshape(-1), pc.MOD_GROUP, strip_last=1)
        prev = aff.reshape(step, 8)
    exp, _ = O.ec_jacreduce(scl, pts, ec2=ec2)
    assert np.array_equal(prev.reshape(-1), exp)


def test_kernel_launch_rejects_other_ec_kernels():
    obj = pc.ECBN128(64)
    v = np.zeros((3 * 32, 8), np.uint32)
    with pytest.raises(NotImplementedError):
        obj.kernelLaunch(v, {"kernel_idx": [pc.CB_EC_JAC_MUL_OPT]},
                         {"in_length": [96], "out_length": 3, "midx": [0]}, n_kernels=1)
    with pytest.raises(ValueError):  # more rows than the constructor reserved
        obj.kernelLaunch(np.zeros((3 * 64 + 3, 8), np.uint32), {"kernel_idx": [pc.CB_EC_JAC_MAD_SHFL]},
                         {"in_length": [3 * 65], "out_length": 3, "midx": [0]})


@pytest.mark.parametrize("ec2", [False, True])
def test_ec_jacreduce_strip_last(ec2):
    n = 333
    pts = _points(n, 80 + ec2, ec2)
    scl = inputs.field_elements(81, n)
    exp, _ = O.ec_jacreduce(scl, pts, ec2=ec2)
    nw = 16 if ec2 else 8
    got1 = pc.ec_jacreduce_h(scl.reshape(-1), pts.reshape(-1), b"", 0, 0, 6, int(ec2), 0, 1, 1, 1, 0)
    assert np.array_equal(got1.reshape(-1), exp)
    got0 = pc.ec_jacreduce_h(scl.reshape(-1), pts.reshape(-1), b"", 0, 0, 6, int(ec2), 0, 1, 1, 0, 0)
    one = np.array(P.to_limbs(P.to_mont(1)) + [0] * (nw - 8), np.uint32)
    assert np.array_equal(got0.reshape(-1), np.concatenate([exp, one]))
    # the empty sum: infinity (0, 1R) and Jacobian (0, 1R, 0)
    z = pc.ec_jacreduce_h(np.zeros(8 * 4, np.uint32), pts[:4].reshape(-1), b"", 0, 0, 6, int(ec2), 0, 1, 1, 0, 0)
    inf = np.zeros(3 * nw, np.uint32)
    inf[nw:nw + 8] = P.to_limbs(P.to_mont(1))
    assert np.array_equal(z.reshape(-1), inf)


def _jac_of(aff_words, zs, ec2):
    """Jacobian (X, Y, Z) = (x Z^2, y Z^3, Z) of affine points, Montgomery words."""
    out = []
    if not ec2:
        for w, z in zip(aff_words, zs):
            pt = P.g1_from_words(w)
            x, y = pt
            out.append(P.to_limbs(P.to_mont(x * z * z % P.Q)) + P.to_limbs(P.to_mont(y * z ** 3 % P.Q))
                       + P.to_limbs(P.to_mont(z)))
    else:
        for w, z in zip(aff_words, zs):
            (x, y) = P.g2_from_words(w)
            z2 = P.f2_mul(z, z)
            X, Y = P.f2_mul(x, z2), P.f2_mul(y, P.f2_mul(z2, z))
            out.append(sum((P.to_limbs(P.to_mont(v)) for v in (*X, *Y, *z)), []))
    return np.array(out, np.uint32)


@pytest.mark.parametrize("ec2", [False, True])
def test_jac2aff_random_points_and_infinity(ec2):
    n = 40
    pts = _points(n, 90 + ec2, ec2)
    rng = random.Random(5)
    zs = [(rng.randrange(1, P.Q), rng.randrange(P.Q)) if ec2 else rng.randrange(1, P.Q) for _ in range(n)]
    jac = _jac_of(pts, zs, ec2)
    nw = 16 if ec2 else 8
    jac[7, 2 * nw:] = 0  # Z = 0: infinity
    fn = pc.ec2_jac2aff_h if ec2 else pc.ec_jac2aff_h
    a1 = fn(jac.reshape(-1), pc.MOD_GROUP, strip_last=1).reshape(n, 2 * nw)
    a0 = fn(jac.reshape(-1), pc.MOD_GROUP, strip_last=0).reshape(n, 3 * nw)
    one = np.array(P.to_limbs(P.to_mont(1)) + [0] * (nw - 8), np.uint32)
    for i in range(n):
        if i == 7:
            exp = np.concatenate([np.zeros(nw, np.uint32), one])
            assert np.array_equal(a1[i], exp)
            assert np.array_equal(a0[i], np.concatenate([exp, np.zeros(nw, np.uint32)]))
        else:
            assert np.array_equal(a1[i], pts[i])
            assert np.array_equal(a0[i], np.concatenate([pts[i], one]))


@pytest.mark.parametrize("pidx", [0, 1])
def test_field_helpers(pidx):
    p = P.PRIMES[pidx]
    n = 64
    a = inputs.field_elements(100 + pidx, n, p)
    b = inputs.field_elements(110 + pidx, n, p)
    ai, bi = inputs.to_ints(a), inputs.to_ints(b)
    rinv = pow(1 << 256, -1, p)
    mm = pc.montmultN_h(a.reshape(-1), b.reshape(-1), pidx)
    assert inputs.to_ints(mm) == [x * y * rinv % p for x, y in zip(ai, bi)]
    assert np.array_equal(mm, O.montmultN(a, b, pidx))
    assert inputs.to_ints(pc.to_montgomeryN_h(a.reshape(-1), pidx)) == [x * (1 << 256) % p for x in ai]
    fm = pc.from_montgomeryN_h(a.reshape(-1), pidx, 0)
    assert inputs.to_ints(fm) == [x * rinv % p for x in ai]
    # strip_last 1: Jacobian G1 rows (X, Y, Z) -> packed (X, Y); 2: G2 (6 -> 4)
    s1 = pc.from_montgomeryN_h(a[:63].reshape(-1), pidx, 1)
    exp1 = [x * rinv % p for i, x in enumerate(ai[:63]) if i % 3 < 2]
    assert inputs.to_ints(s1[:42]) == exp1 and not s1[42:].any() and s1.shape == (63, 8)
    s2 = pc.from_montgomeryN_h(a[:60].reshape(-1), pidx, 2)
    exp2 = [x * rinv % p for i, x in enumerate(ai[:60]) if i % 6 < 4]
    assert inputs.to_ints(s2[:40]) == exp2 and not s2[40:].any()
    neg = pc.montmult_neg_h(a[:1].reshape(-1), b[:1].reshape(-1), pidx)
    assert inputs.to_ints(neg.reshape(-1, 8)) == [(-ai[0] * bi[0] * rinv) % p]
    sq = pc.montsquare_h(a[:1].reshape(-1), pidx)
    assert inputs.to_ints(sq) == [ai[0] * ai[0] * rinv % p]
    # Montgomery inverse: (aR)^-1 R, and 0 -> 0
    z = a.copy()
    z[3] = 0
    inv = inputs.to_ints(pc.montinv_h(z.reshape(-1), pidx))
    r = 1 << 256
    for i, x in enumerate(inputs.to_ints(z)):
        assert inv[i] == (0 if x == 0 else pow(x * rinv % p, -1, p) * r % p)


def test_fq2_product_entry_point():
    n = 50
    a = inputs.field_elements(120, 2 * n, P.Q).reshape(n, 16)
    b = inputs.field_elements(121, 2 * n, P.Q).reshape(n, 16)
    got = pc._field_op(_capi.OP_MULM_EXT, a, b, 0, words=16).reshape(n, 16)
    rinv = pow(1 << 256, -1, P.Q)
    for i in range(n):
        x = (P.from_limbs(a[i, :8]) * rinv % P.Q, P.from_limbs(a[i, 8:]) * rinv % P.Q)
        y = (P.from_limbs(b[i, :8]) * rinv % P.Q, P.from_limbs(b[i, 8:]) * rinv % P.Q)
        z = P.f2_mul(x, y)
        assert list(got[i]) == P.to_limbs(P.to_mont(z[0])) + P.to_limbs(P.to_mont(z[1]))
    # and the C oracle's Fq2 product (montmult_ext_h restated) on the first element
    one = np.ascontiguousarray(a[0])
    two = np.ascontiguousarray(b[0])
    out = np.zeros(16, np.uint32)
    O.lib().or_montmult_ext(O._p(out), O._p(one), O._p(two), 0)
    assert np.array_equal(out, got[0])


def test_ntt_interpolandmul_and_mpoly_eval():
    lg = 10
    m = 1 << lg
    pA = inputs.field_elements(130, m)
    pB = inputs.field_elements(131, m)
    roots = O.roots(lg + 1)
    got = pc.ntt_interpolandmul_h(pA.reshape(-1), pB.reshape(-1), roots.reshape(-1), 1, pc.MOD_FIELD)
    assert np.array_equal(got.reshape(-1, 8), O.interpolandmul(pA, pB, lg))
    # mpoly_eval_h over the reference's per-wire sparse layout
    rng = np.random.default_rng(3)
    nw, nrows = 20, 64
    polys = []
    for _ in range(nw):
        k = int(rng.integers(0, 6))
        polys.append((rng.choice(nrows, size=k, replace=False), inputs.field_elements(int(rng.integers(1 << 30)), k)))
    pdata = [nw] + [len(r) for r, _ in polys]
    for r, v in polys:
        pdata += list(r) + list(v.reshape(-1))
    pdata = np.array(pdata, np.uint32)
    w = inputs.field_elements(140, nw)
    got = pc.mpoly_eval_h(w, pdata, 1, nrows, 0, nw, 0, pc.MOD_FIELD)
    rows = np.concatenate([r for r, _ in polys])
    cols = np.concatenate([np.full(len(r), i) for i, (r, _) in enumerate(polys)])
    vals = np.concatenate([v for _, v in polys])
    csr = CSR.from_triples(rows, cols, vals, nrows)
    assert np.array_equal(got.reshape(-1, 8), O.mpoly_eval_csr(csr.rowptr, csr.col, csr.val, w, nrows))
