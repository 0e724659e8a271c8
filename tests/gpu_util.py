"""Helpers shared by the GPU parity tests."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_tbin(name):
    return np.fromfile(os.path.join(GOLDEN, name), dtype=np.uint32)


def g1_fixture():
    s = load_tbin("test_scmul_samples.tbin").reshape(-1, 8)
    n = s.shape[0] // 3
    return s[:n].copy(), s[n:].reshape(n, 16).copy(), load_tbin("test_scmul_reduction_soln.tbin")


def g2_fixture():
    s = load_tbin("test2_scmul_samples.tbin").reshape(-1, 8)
    n = s.shape[0] // 5
    return s[:n].copy(), s[n:].reshape(n, 32).copy(), load_tbin("test2_scmul_reduction_soln.tbin")


RMONT = 1 << 256


def dot_mod_r(a: np.ndarray, b: np.ndarray) -> int:
    """sum a_i b_i mod r for (n, 8) uint32 normal-form vectors, from the C
    oracle's Montgomery products (a b R^-1 mod r, each < r) summed limb-wise."""
    import oracle as O
    from oracle import pyoracle as P
    a = np.ascontiguousarray(a, dtype=np.uint32).reshape(-1, 8)
    b = np.ascontiguousarray(b, dtype=np.uint32).reshape(-1, 8)
    tot = 0
    for s in range(0, a.shape[0], 1 << 24):  # 2^24 * 2^32 < 2^64 per limb sum
        prod = O.montmultN(a[s:s + (1 << 24)], b[s:s + (1 << 24)], 1)
        limbs = prod.astype(np.uint64).sum(axis=0)
        tot += sum(int(v) << (32 * i) for i, v in enumerate(limbs))
    return tot * RMONT % P.RR


def gen_mul(s: int, g2: bool = False) -> np.ndarray:
    """s * G (G1 generator (1, 2) or the standard G2 generator), affine Montgomery words."""
    import oracle as O
    from oracle import pyoracle as P
    from cusnarks_amd import inputs
    gen = np.array(P.g2_to_words(P.G2_GEN) if g2 else P.g1_to_words(P.G1_GEN), np.uint32)
    return O.ec_scmul_aff(inputs.from_ints([s % P.RR]), gen, ec2=g2)[0]


def check_h_at_point(circ, H: np.ndarray, z: int) -> None:
    """The H step's 2m-coefficient product (normal form; slot 2m-1 ignored)
    against polH(z) == A(z) B(z), A and B the interpolants of the QAP row
    sums the C oracle evaluates: pins every coefficient with probability
    1 - 2m/r by Schwartz-Zippel, in O(m)."""
    import oracle as O
    from oracle import pyoracle as P
    from cusnarks_amd import inputs
    m = circ.m
    zm = inputs.from_ints([P.to_mont(z, P.RR)])[0]
    pA = O.mpoly_eval_csr(circ.A.rowptr, circ.A.col, circ.A.val, circ.witness, m)
    pB = O.mpoly_eval_csr(circ.B.rowptr, circ.B.col, circ.B.val, circ.witness, m)
    a = P.from_mont(inputs.to_ints(O.lagrange_eval(pA, circ.log2m, zm).reshape(1, 8))[0], P.RR)
    b = P.from_mont(inputs.to_ints(O.lagrange_eval(pB, circ.log2m, zm).reshape(1, 8))[0], P.RR)
    h = H.reshape(-1, 8)[: 2 * m - 1]
    assert inputs.to_ints(O.poly_eval(h, zm).reshape(1, 8))[0] == a * b % P.RR


def expected_proof(circ, keys, H: np.ndarray, r: int, s: int):
    """(pi_a, pi_b, pi_c) words from the key's discrete logs: with
    a = <w|1|r, kA>, b = <w|1|s, kB2>, b1 = <w|1|s, kB1>,
    c' = <w_priv, kC_priv>, c2 = <H[m:2m-1] | -rs, kH>,
    pi_a = a G1, pi_b = b G2, pi_c = (c2 + c' + s a + r b1) G1."""
    from cusnarks_amd import inputs
    from oracle import pyoracle as P
    R = P.RR
    m, lo = circ.m, circ.n_public + 1
    w = circ.witness
    wr = np.concatenate([w, inputs.from_ints([1, r % R])])
    ws = np.concatenate([w, inputs.from_ints([1, s % R])])
    a = dot_mod_r(wr, keys.A)
    b = dot_mod_r(ws, keys.B2)
    b1 = dot_mod_r(ws, keys.B1)
    cp = dot_mod_r(w[lo:], keys.C[lo:])
    hs = np.concatenate([H.reshape(-1, 8)[m:2 * m - 1], inputs.from_ints([(-r * s) % R])])
    c2 = dot_mod_r(hs, keys.hExps)
    c = (c2 + cp + s * a + r * b1) % R
    return gen_mul(a), gen_mul(b, True), gen_mul(c)
