"""CPU tests that pin the oracle: the C restatement of the reference CPU path
is checked against the reference's own known answers (tests/golden) and
against closed forms / brute force in Python integers, chosen so that a
dropped term, a wrong sign or a transposed operand fails."""
import json
import os
import random

import numpy as np
import pytest

import oracle as O
from cusnarks_amd import inputs
from oracle import pyoracle as P
from tests.gpu_util import GOLDEN, g1_fixture, g2_fixture, load_tbin


def test_montmult_kat():
    t = json.load(open(os.path.join(GOLDEN, "kat_tables.json")))
    A = np.array(t["A_test"], np.uint32).reshape(-1, 8)
    B = np.array(t["B_test"], np.uint32).reshape(-1, 8)
    C = np.array(t["C_test"], np.uint32).reshape(-1, 8)
    assert np.array_equal(O.montmultN(A, B, 1), C)
    # the KAT itself is the closed form a*b*R^-1 mod r
    for a, b, c in zip(A, B, C):
        assert P.montmult(P.from_limbs(a), P.from_limbs(b), P.RR) == P.from_limbs(c)


@pytest.mark.parametrize("pidx", [0, 1])
def test_field_closed_forms(pidx):
    p = P.PRIMES[pidx]
    rng = random.Random(pidx)
    xs = [rng.randrange(p) for _ in range(200)] + [0, 1, p - 1]
    ys = [rng.randrange(p) for _ in range(200)] + [p - 1, 0, p - 1]
    a = inputs.from_ints(xs)
    b = inputs.from_ints(ys)
    got = inputs.to_ints(O.montmultN(a, b, pidx))
    assert got == [x * y * pow(P.RMONT, -1, p) % p for x, y in zip(xs, ys)]
    assert inputs.to_ints(O.to_montgomeryN(a, pidx)) == [x * P.RMONT % p for x in xs]
    assert inputs.to_ints(O.from_montgomeryN(a, pidx)) == [x * pow(P.RMONT, -1, p) % p for x in xs]


def test_roots_closed_form():
    r = O.roots(5)
    w = P.root_of_unity(5)
    assert [P.from_mont(v, P.RR) for v in inputs.to_ints(r)] == [pow(w, k, P.RR) for k in range(32)]
    # 5^((r-1)/2^28) is the reference's proots_field[28] (constants.cpp:352)
    assert P.to_limbs(P.to_mont(P.root_of_unity(28), P.RR)) == [
        2161196444, 1668182869, 608567254, 2720789364, 514982872, 1447373504, 694417895, 409005972]


def test_g1_msm_fixture():
    scl, pts, exp = g1_fixture()
    for order in (2, 6, 8):
        out, _ = O.ec_jacreduce(scl, pts, order=order)
        assert np.array_equal(out, exp)


def test_g2_msm_fixture():
    scl, pts, exp = g2_fixture()
    out, _ = O.ec_jacreduce(scl, pts, ec2=True)
    assert np.array_equal(out, exp)


def test_fixture_is_textbook_sum():
    # the reference's stored G1 result equals sum k_i P_i by textbook affine arithmetic
    scl, pts, exp = g1_fixture()
    ks = [int(v) for v in inputs.to_ints(scl)]
    ps = [P.g1_from_words(p) for p in pts]
    assert P.g1_to_words(P.G1.msm(ks, ps)) == list(exp)


def test_g1_scalar_mul_fixture():
    # per-point products k_i P_i of test_scmul_soln.tbin (Jacobian) vs the oracle
    scl, pts, _ = g1_fixture()
    jac = load_tbin("test_scmul_soln.tbin").reshape(-1, 24)
    idx = list(range(0, 1024, 97)) + [1023]
    for i in idx:
        mine = O.ec_scmul_aff(scl[i:i + 1], pts[i])
        ref = np.zeros(16, np.uint32)
        O.lib().or_ec_jac2aff(O._p(ref), O._p(np.ascontiguousarray(jac[i])), 1, 0, 1)
        assert np.array_equal(mine[0], ref), i


@pytest.mark.parametrize("ec2", [False, True])
def test_msm_bruteforce_small(ec2):
    rng = random.Random(11 + ec2)
    C = P.G2 if ec2 else P.G1
    gen = P.G2_GEN if ec2 else P.G1_GEN
    enc = P.g2_to_words if ec2 else P.g1_to_words
    n = 9
    ks = [rng.randrange(P.RR) for _ in range(n)]
    pts = [C.mul(rng.randrange(1, P.RR), gen) for _ in range(n)]
    pts[3] = None  # infinity input
    ks[5] = 0
    pts[7] = pts[6]  # same point twice
    ks[7] = ks[6]
    exp = enc(C.msm(ks, pts))
    out, _ = O.ec_jacreduce(inputs.from_ints(ks), np.array([enc(p) for p in pts], np.uint32), ec2=ec2, order=3)
    assert list(out) == exp


def test_msm_thread_partition_is_reference():
    # ec_jacreduce_server_h: threads get multiples of (order<<5)*order, last takes the rest
    n = 1 << 14
    k = inputs.field_elements(3, 64)
    pts = O.ec_scmul_aff(k, np.array(P.g1_to_words(P.G1_GEN), np.uint32))
    scl = inputs.field_elements(4, n)
    pts = np.tile(pts, (n // 64, 1))
    one, t1 = O.ec_jacreduce(scl, pts, max_threads=1)
    many, t8 = O.ec_jacreduce(scl, pts, max_threads=8)
    assert t1 == 1 and t8 >= 1
    assert np.array_equal(one, many)


@pytest.mark.parametrize("lg", [0, 1, 3, 6])
def test_ntt_oracle_vs_dft(lg):
    rng = random.Random(lg)
    a = [rng.randrange(P.RR) for _ in range(1 << lg)]
    am = inputs.from_ints([P.to_mont(x, P.RR) for x in a])
    X = [P.from_mont(v, P.RR) for v in inputs.to_ints(O.ntt(am, lg))]
    assert X == P.dft(a, lg)
    assert inputs.to_ints(O.intt(am, lg, fmt=0)) == P.dft(a, lg, inverse=True)
    Y = [P.from_mont(v, P.RR) for v in inputs.to_ints(O.intt(am, lg, fmt=1))]
    assert Y == P.dft(a, lg, inverse=True)


def test_python_ntt_vs_dft():
    rng = random.Random(3)
    a = [rng.randrange(P.RR) for _ in range(64)]
    assert P.ntt(a, 6) == P.dft(a, 6)
    assert P.ntt(P.ntt(a, 6), 6, inverse=True) == a


@pytest.mark.parametrize("lg", [1, 3, 5])
def test_h_oracle_vs_schoolbook(lg):
    rng = random.Random(100 + lg)
    m = 1 << lg
    pa = [rng.randrange(P.RR) for _ in range(m)]
    pb = [rng.randrange(P.RR) for _ in range(m)]
    got = O.interpolandmul(inputs.from_ints([P.to_mont(x, P.RR) for x in pa]),
                           inputs.from_ints([P.to_mont(x, P.RR) for x in pb]), lg)
    assert inputs.to_ints(got) == P.h_poly(pa, pb, lg)
    # and the interpolants really take the given values (A(w^i) = pA[i])
    a = P.ntt(pa, lg, inverse=True)
    w = P.root_of_unity(lg)
    assert all(sum(c * pow(w, i * j, P.RR) for j, c in enumerate(a)) % P.RR == pa[i] for i in range(m))
    # the threaded 4-step form (the timed CPU baseline) gives the same words
    got_mt = O.interpolandmul(inputs.from_ints([P.to_mont(x, P.RR) for x in pa]),
                              inputs.from_ints([P.to_mont(x, P.RR) for x in pb]), lg, threads=3)
    assert np.array_equal(got_mt, got)


@pytest.mark.parametrize("lg", [2, 7, 10, 13])
def test_h_oracle_threaded_4step_equals_radix2(lg):
    # odd and even transform sizes (the 2^lg-point transforms split 2^(lg//2) x 2^(lg - lg//2))
    a = inputs.field_elements(300 + lg, 1 << lg)
    b = inputs.field_elements(400 + lg, 1 << lg)
    assert np.array_equal(O.interpolandmul(a, b, lg, threads=4), O.interpolandmul(a, b, lg))


def test_mpoly_eval_oracle():
    rng = np.random.default_rng(0)
    nrows, nw = 50, 30
    rows = rng.integers(0, nrows, size=120)
    cols = rng.integers(0, nw, size=120)
    vals = [int(v) for v in rng.integers(0, 2**62, size=120)]
    from cusnarks_amd.poly import CSR
    vm = inputs.from_ints([P.to_mont(v, P.RR) for v in vals])
    csr = CSR.from_triples(rows, cols, vm, nrows)
    w = inputs.field_elements(1, nw)
    got = [P.from_mont(v, P.RR) for v in inputs.to_ints(O.mpoly_eval_csr(csr.rowptr, csr.col, csr.val, w, nrows))]
    exp = P.mpoly_eval(list(zip(rows, cols, vals)), inputs.to_ints(w), nrows)
    assert got == exp


def test_groth16_python_oracle_matches_c_oracle_composition():
    # the Python-integer composition and the C restatement agree on a tiny circuit
    from cusnarks_amd.synth import Circuit, KeyScalars
    circ = Circuit(20, 12, 1, seed=5)
    keys = KeyScalars(circ, seed=5)
    g1 = np.array(P.g1_to_words(P.G1_GEN), np.uint32)
    g2 = np.array(P.g2_to_words(P.G2_GEN), np.uint32)
    pts = {k: O.ec_scmul_aff(getattr(keys, k), g1) for k in ("A", "B1", "C", "hExps")}
    pts["B2"] = O.ec_scmul_aff(keys.B2, g2, ec2=True)
    r, s = P.reference_rs(5)
    nv, m, npub = circ.n_vars, circ.m, circ.n_public
    w = circ.witness
    pA = O.mpoly_eval_csr(circ.A.rowptr, circ.A.col, circ.A.val, w, m)
    pB = O.mpoly_eval_csr(circ.B.rowptr, circ.B.col, circ.B.val, w, m)
    polH = O.interpolandmul(pA, pB, circ.log2m)
    one = inputs.from_ints([1])
    rs_ = inputs.from_ints([r, s, (-r * s) % P.RR])
    pa, _ = O.ec_jacreduce(np.concatenate([w, one, rs_[0:1]]), pts["A"])
    pb, _ = O.ec_jacreduce(np.concatenate([w, one, rs_[1:2]]), pts["B2"], ec2=True)
    pb1, _ = O.ec_jacreduce(np.concatenate([w, one, rs_[1:2]]), pts["B1"])
    pcp, _ = O.ec_jacreduce(w[npub + 1:], pts["C"][npub + 1:])
    pc2, _ = O.ec_jacreduce(np.concatenate([polH[m:2 * m - 1], rs_[2:3]]), pts["hExps"])
    pc, _ = O.ec_jacreduce(np.concatenate([one, one, rs_[1:2], rs_[0:1]]), np.stack([pc2, pcp, pa, pb1]))

    def trip(csr):
        return [(r_, int(csr.col[e]), P.from_mont(inputs.to_ints(csr.val[e:e + 1])[0], P.RR))
                for r_ in range(csr.nrows) for e in range(int(csr.rowptr[r_]), int(csr.rowptr[r_ + 1]))]
    aff = {k: [P.g1_from_words(x) for x in v] for k, v in pts.items() if k != "B2"}
    aff["B2"] = [P.g2_from_words(x) for x in pts["B2"]]
    ea, eb, ec = P.groth16_prove(aff, trip(circ.A), trip(circ.B), inputs.to_ints(w), npub, circ.log2m, r, s)
    assert list(pa) == P.g1_to_words(ea)
    assert list(pb) == P.g2_to_words(eb)
    assert list(pc) == P.g1_to_words(ec)


def test_reference_rs_draw():
    # r, s follow the stdlib draw the reference makes after random.seed(seed)
    import random as _r
    g = _r.Random(42)
    assert P.reference_rs(42) == (g.randint(1, P.RR - 1), g.randint(1, P.RR - 1))


def test_closed_form_eval_helpers():
    # Horner and barycentric evaluation against the Python closed forms, and
    # the H-step identity polH(z) = A(z) B(z) they are used for at full size
    lg = 8
    m = 1 << lg
    rng = np.random.default_rng(4)
    z = int(rng.integers(1, 2**62)) * 977 + 3
    zm = inputs.from_ints([P.to_mont(z, P.RR)])[0]
    coeffs = inputs.field_elements(31, 300)
    got = inputs.to_ints(O.poly_eval(coeffs, zm).reshape(1, 8))[0]
    assert got == P.poly_at(inputs.to_ints(coeffs), z)
    ev = inputs.field_elements(32, m)  # Montgomery evaluations
    evn = [P.from_mont(v, P.RR) for v in inputs.to_ints(ev)]
    got = P.from_mont(inputs.to_ints(O.lagrange_eval(ev, lg, zm).reshape(1, 8))[0], P.RR)
    assert got == P.poly_at(P.ntt(evn, lg, inverse=True), z)
    ev2 = inputs.field_elements(33, m)
    polH = O.interpolandmul(ev, ev2, lg)
    a = P.from_mont(inputs.to_ints(O.lagrange_eval(ev, lg, zm).reshape(1, 8))[0], P.RR)
    b = P.from_mont(inputs.to_ints(O.lagrange_eval(ev2, lg, zm).reshape(1, 8))[0], P.RR)
    h = inputs.to_ints(O.poly_eval(polH, zm).reshape(1, 8))[0]
    assert h == a * b % P.RR
    # a point of the domain is refused
    with pytest.raises(ValueError):
        O.lagrange_eval(ev, lg, inputs.from_ints([P.to_mont(P.root_of_unity(lg), P.RR)])[0])
