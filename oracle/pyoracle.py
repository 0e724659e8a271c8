"""Pure-Python big-integer oracle for the BN254 Groth16 hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package may import this
module; only tests/, the smoke check in __graft_entry__ and the cpu_baseline
leg of bench.py use it, and only as the checker.

Every routine here is a closed-form / textbook statement of the mathematics
the reference computes, written with Python integers so that it is easy to
read and hard to get subtly wrong.  Each function names the reference code
whose result it defines (paths are relative to the cusnarks source tree).

Conventions shared with the reference (src/cuda/types.h:228-233,
src/cuda/constants.cpp:49-67, 264-292):
  * Q  = BN254 base field prime   (reference "MOD_GROUP", pidx 0)
  * RR = BN254 scalar field prime (reference "MOD_FIELD", pidx 1)
  * elements are 8 x u32 little-endian limbs, Montgomery R = 2^256
  * affine infinity is encoded (0, 1*R mod Q); Jacobian infinity (0, R, 0)
  * Fq2 = Fq[u]/(u^2+1), G2 lives on y^2 = x^3 + 3/(9+u)
"""
from __future__ import annotations

import random

Q = 21888242871839275222246405745257275088696311157297823662689037894645226208583
RR = 21888242871839275222246405745257275088548364400416034343698204186575808495617
RMONT = 1 << 256
PRIMES = (Q, RR)  # indexed by pidx, as in the reference

# generators, src/python/zutils.py:117-136 (normal form)
G1_GEN = (1, 2)
G2_GEN = (
    (10857046999023057135944570762232829481370756359578518086990519993285655852781,
     11559732032986387107991004021392285783925812861821192530917403151452391805634),
    (8495653923123431417604973247489272438418190587263600148770280649306958101930,
     4082367875863433681332203403145435568316851327593401208105741076214120093531),
)
B1 = 3


# ----------------------------------------------------------------- limbs
def to_limbs(x: int, n: int = 8) -> list[int]:
    return [(x >> (32 * i)) & 0xFFFFFFFF for i in range(n)]


def from_limbs(ls) -> int:
    v = 0
    for i, w in enumerate(ls):
        v |= int(w) << (32 * i)
    return v


def to_mont(x: int, p: int = Q) -> int:
    return x * RMONT % p


def from_mont(x: int, p: int = Q) -> int:
    return x * pow(RMONT, -1, p) % p


def montmult(a: int, b: int, p: int) -> int:
    """a*b*R^-1 mod p: the result of montmult_h (src/cuda/ff.cpp:225-388)."""
    return a * b * pow(RMONT, -1, p) % p


# ----------------------------------------------------------------- Fq2
def f2_add(a, b):
    return ((a[0] + b[0]) % Q, (a[1] + b[1]) % Q)


def f2_sub(a, b):
    return ((a[0] - b[0]) % Q, (a[1] - b[1]) % Q)


def f2_mul(a, b):
    """(a0 + a1 u)(b0 + b1 u), u^2 = -1 (nonres = Q-1, constants.cpp:175)."""
    return ((a[0] * b[0] - a[1] * b[1]) % Q, (a[0] * b[1] + a[1] * b[0]) % Q)


def f2_inv(a):
    d = pow((a[0] * a[0] + a[1] * a[1]) % Q, -1, Q)
    return (a[0] * d % Q, (-a[1]) * d % Q)


B2 = f2_mul((3, 0), f2_inv((9, 1)))  # twist b' = 3/(9+u)


# ----------------------------------------------------------------- curve ops
class Curve:
    """Affine short-Weierstrass y^2 = x^3 + b over Fq (G1) or Fq2 (G2).

    None is the point at infinity.  Textbook chord-and-tangent formulas:
    they define what ec_jacadd_h / ec_jacaddmixed_h / ec_jacdouble_h
    (src/cuda/ec.cpp:131-293, 462-514) and the G2 twins compute, once
    converted back to affine.
    """

    def __init__(self, ext: bool):
        self.ext = ext
        if ext:
            self.add_f, self.sub_f, self.mul_f, self.inv_f = f2_add, f2_sub, f2_mul, f2_inv
            self.zero, self.one, self.b = (0, 0), (1, 0), B2
        else:
            self.add_f = lambda a, b: (a + b) % Q
            self.sub_f = lambda a, b: (a - b) % Q
            self.mul_f = lambda a, b: a * b % Q
            self.inv_f = lambda a: pow(a, -1, Q)
            self.zero, self.one, self.b = 0, 1, B1

    def on_curve(self, P) -> bool:
        if P is None:
            return True
        x, y = P
        m = self.mul_f
        return self.sub_f(m(y, y), self.add_f(m(m(x, x), x), self.b)) == self.zero

    def neg(self, P):
        if P is None:
            return None
        return (P[0], self.sub_f(self.zero, P[1]))

    def add(self, P, R_):
        if P is None:
            return R_
        if R_ is None:
            return P
        m, s, a = self.mul_f, self.sub_f, self.add_f
        x1, y1 = P
        x2, y2 = R_
        if x1 == x2:
            if a(y1, y2) == self.zero:
                return None
            three = (3, 0) if self.ext else 3
            two = (2, 0) if self.ext else 2
            lam = m(m(three, m(x1, x1)), self.inv_f(m(two, y1)))
        else:
            lam = m(s(y2, y1), self.inv_f(s(x2, x1)))
        x3 = s(s(m(lam, lam), x1), x2)
        y3 = s(m(lam, s(x1, x3)), y1)
        return (x3, y3)

    # Jacobian helpers keep brute-force MSMs fast (one inversion at the end)
    def _jdbl(self, P):
        X, Y, Z = P
        m, s, a = self.mul_f, self.sub_f, self.add_f
        if Z == self.zero or Y == self.zero:
            return (self.one, self.one, self.zero)
        A = m(X, X)
        B = m(Y, Y)
        C = m(B, B)
        t = a(X, B)
        D = s(s(m(t, t), A), C)
        D = a(D, D)
        E = a(a(A, A), A)
        F = m(E, E)
        X3 = s(F, a(D, D))
        C8 = a(C, C); C8 = a(C8, C8); C8 = a(C8, C8)
        Y3 = s(m(E, s(D, X3)), C8)
        Z3 = m(a(Y, Y), Z)
        return (X3, Y3, Z3)

    def _jadd(self, P, R_):
        m, s, a = self.mul_f, self.sub_f, self.add_f
        if P[2] == self.zero:
            return R_
        if R_[2] == self.zero:
            return P
        X1, Y1, Z1 = P
        X2, Y2, Z2 = R_
        Z1Z1 = m(Z1, Z1)
        Z2Z2 = m(Z2, Z2)
        U1 = m(X1, Z2Z2)
        U2 = m(X2, Z1Z1)
        S1 = m(m(Y1, Z2), Z2Z2)
        S2 = m(m(Y2, Z1), Z1Z1)
        if U1 == U2:
            if S1 == S2:
                return self._jdbl(P)
            return (self.one, self.one, self.zero)
        H = s(U2, U1)
        Rr = s(S2, S1)
        H2 = m(H, H)
        H3 = m(H2, H)
        V = m(U1, H2)
        X3 = s(s(m(Rr, Rr), H3), a(V, V))
        Y3 = s(m(Rr, s(V, X3)), m(S1, H3))
        Z3 = m(m(Z1, Z2), H)
        return (X3, Y3, Z3)

    def to_jac(self, P):
        if P is None:
            return (self.one, self.one, self.zero)
        return (P[0], P[1], self.one)

    def to_aff(self, J):
        if J[2] == self.zero:
            return None
        zi = self.inv_f(J[2])
        zi2 = self.mul_f(zi, zi)
        return (self.mul_f(J[0], zi2), self.mul_f(J[1], self.mul_f(zi2, zi)))

    def mul(self, k: int, P):
        """k*P by left-to-right double-and-add (k may be any non-negative int)."""
        acc = self.to_jac(None)
        Pj = self.to_jac(P)
        for bit in bin(k)[2:] if k else "":
            acc = self._jdbl(acc)
            if bit == "1":
                acc = self._jadd(acc, Pj)
        return self.to_aff(acc)

    def msm(self, scalars, points):
        """Brute-force sum k_i*P_i: the value ec_jacreduce_server_h
        (src/cuda/ec.cpp:1536-1704) and the GPU multiexp return."""
        acc = self.to_jac(None)
        for k, P in zip(scalars, points):
            if k == 0 or P is None:
                continue
            acc = self._jadd(acc, self.to_jac(self.mul(k, P)))
        return self.to_aff(acc)


G1 = Curve(False)
G2 = Curve(True)


# ----------------------------------------------------------------- encodings
def g1_to_words(P) -> list[int]:
    """Affine Montgomery (X, Y); infinity -> (0, R mod Q) (constants.cpp:264-292)."""
    if P is None:
        return to_limbs(0) + to_limbs(to_mont(1))
    return to_limbs(to_mont(P[0])) + to_limbs(to_mont(P[1]))


def g1_from_words(w):
    x = from_mont(from_limbs(w[0:8]))
    y = from_mont(from_limbs(w[8:16]))
    if x == 0 and y in (0, 1):
        return None
    return (x, y)


def g2_to_words(P) -> list[int]:
    """[X.c0, X.c1, Y.c0, Y.c1] Montgomery; infinity -> (0, 0, R, 0)."""
    if P is None:
        return to_limbs(0) * 2 + to_limbs(to_mont(1)) + to_limbs(0)
    (x0, x1), (y0, y1) = P
    return sum((to_limbs(to_mont(v)) for v in (x0, x1, y0, y1)), [])


def g2_from_words(w):
    v = [from_mont(from_limbs(w[8 * i:8 * i + 8])) for i in range(4)]
    if v[0] == 0 and v[1] == 0 and v[3] == 0 and v[2] in (0, 1):
        return None
    return ((v[0], v[1]), (v[2], v[3]))


# ----------------------------------------------------------------- NTT / H
def root_of_unity(log2n: int) -> int:
    """Primitive 2^log2n-th root: 5^((r-1)/2^k) (proots_field, constants.cpp:323-352)."""
    assert 0 <= log2n <= 28
    return pow(5, (RR - 1) >> log2n, RR)


def dft(a, log2n: int, inverse: bool = False):
    """O(N^2) definition of ntt_h / intt_h (src/cuda/ntt.cpp:453-603), normal form."""
    n = 1 << log2n
    w = root_of_unity(log2n)
    if inverse:
        w = pow(w, -1, RR)
    out = []
    for k in range(n):
        acc = 0
        wk = pow(w, k, RR)
        t = 1
        for i in range(n):
            acc += a[i] * t
            t = t * wk % RR
        out.append(acc % RR)
    if inverse:
        ninv = pow(n, -1, RR)
        out = [x * ninv % RR for x in out]
    return out


def ntt(a, log2n: int, inverse: bool = False):
    """Recursive radix-2 transform; same value as dft(), O(N log N)."""
    n = 1 << log2n
    assert len(a) == n
    w = root_of_unity(log2n)
    if inverse:
        w = pow(w, -1, RR)

    def rec(v, w):
        m = len(v)
        if m == 1:
            return list(v)
        ev = rec(v[0::2], w * w % RR)
        od = rec(v[1::2], w * w % RR)
        out = [0] * m
        t = 1
        for k in range(m // 2):
            x = od[k] * t % RR
            out[k] = (ev[k] + x) % RR
            out[k + m // 2] = (ev[k] - x) % RR
            t = t * w % RR
        return out

    out = rec([x % RR for x in a], w)
    if inverse:
        ninv = pow(n, -1, RR)
        out = [x * ninv % RR for x in out]
    return out


def poly_mul(a, b):
    """Schoolbook product mod RR."""
    out = [0] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        if x == 0:
            continue
        for j, y in enumerate(b):
            out[i + j] = (out[i + j] + x * y) % RR
    return out


def h_poly(pA, pB, log2m: int):
    """ntt_interpolandmul_server_h (src/cuda/ntt.cpp:672-1070): coefficients of
    A(x)*B(x) where A, B are the degree < m interpolants of the evaluations
    pA, pB on <w_m>.  Returns all 2m coefficients (the last is always 0), normal
    form -- exactly the buffer the reference returns; the prover keeps
    [m : 2m-1] (src/python/groth_prover.py:742)."""
    a = ntt(pA, log2m, inverse=True)
    b = ntt(pB, log2m, inverse=True)
    prod = poly_mul(a, b)
    m = 1 << log2m
    return prod + [0] * (2 * m - len(prod))


# ----------------------------------------------------------------- QAP eval
def mpoly_eval(rows, wit, nrows):
    """Sparse evaluation pX[row] = sum coeff * w[wire] (mpoly_eval_h,
    src/cuda/mpoly.cpp:125-168).  rows: list of (row, wire, coeff)."""
    out = [0] * nrows
    for row, wire, coeff in rows:
        out[row] = (out[row] + coeff * wit[wire]) % RR
    return out


# ----------------------------------------------------------------- Groth16
def groth16_prove(pk, qapA, qapB, witness, n_public, log2m, rnd, snd, swap_rs_in_c=False):
    """Groth16 proof exactly as the reference composes it
    (src/python/groth_prover.py:575-790 pysnarkP_CPU, 1414-1453 gen_proof):

      pA, pB = sum_wire coeff * w   (mpoly_eval_h, per QAP row)
      polH   = coefficients of A(x)B(x) (ntt_interpolandmul); H = polH[m:2m-1]
      pi_a   = sum w_i A_i + 1*A[nVars] + r*A[nVars+1]
      pi_b   = sum w_i B2_i + 1*B2[nVars] + s*B2[nVars+1]      (G2)
      pi_b1  = sum w_i B1_i + 1*B1[nVars] + s*B1[nVars+1]
      pi_c'  = sum_{nPublic < i < nVars} w_i C_i
      pi_c2  = sum_j H_j hExps_j + (-r s) hExps[m-1]            (hExps[m-1] = delta1)
      pi_c   = pi_c2 + pi_c' + s pi_a + r pi_b1

    pk: dict of affine point lists A, B1 (G1, nVars+2), B2 (G2, nVars+2),
    C (nVars), hExps (m); qapA/qapB: (row, wire, coeff) triples (normal form);
    witness: ints.  Returns affine (pi_a, pi_b, pi_c).  swap_rs_in_c plants a
    mistake (pi_c = ... + r pi_a + s pi_b1) for the tests that show the
    verifier below rejects it."""
    m = 1 << log2m
    w = [x % RR for x in witness]
    nv = len(w)
    pA = mpoly_eval(qapA, w, m)
    pB = mpoly_eval(qapB, w, m)
    H = h_poly(pA, pB, log2m)[m:2 * m - 1]
    A = G1.msm(w + [1, rnd], pk["A"])
    B2 = G2.msm(w + [1, snd], pk["B2"])
    B1 = G1.msm(w + [1, snd], pk["B1"])
    Cp = G1.msm(w[n_public + 1:nv], pk["C"][n_public + 1:nv])
    C2 = G1.msm(H + [(-rnd * snd) % RR], pk["hExps"])
    ka, kb = (rnd, snd) if swap_rs_in_c else (snd, rnd)
    C = G1.add(G1.add(C2, Cp), G1.add(G1.mul(ka, A), G1.mul(kb, B1)))
    return A, B2, C


# ----------------------------------------------------------------- Groth16 check
def poly_at(coeffs, x: int) -> int:
    """Horner evaluation mod RR."""
    acc = 0
    for cf in reversed(coeffs):
        acc = (acc * x + cf) % RR
    return acc


def qap_at(triples, w, log2m: int, tau: int) -> int:
    """sum_i w_i X_i(tau) for the QAP polynomials X_i of a matrix given as
    (row, wire, coeff) triples: the interpolant of the row sums (M w)_row on
    <w_m>, evaluated at tau by Horner.  (The setup derives the per-wire values
    from Lagrange basis values instead; this is a second, independent route.)"""
    m = 1 << log2m
    return poly_at(ntt(mpoly_eval(triples, w, m), log2m, inverse=True), tau)


def groth16_exponents(triples_a, triples_b, triples_c, witness, n_public, log2m, td, rnd, snd):
    """Discrete logs (a, b, c) of a Groth16 proof A = a G1, B = b G2, C = c G1
    from the definition (Groth, EUROCRYPT 2016, section 3.2), for toxic
    scalars td = (tau, alpha, beta, gamma, delta) and randomness r, s:

      a = alpha + A(tau) + r delta
      b = beta  + B(tau) + s delta
      c = (sum_{i > nPublic} w_i (beta A_i + alpha B_i + C_i)(tau) + h(tau) Z(tau)) / delta
          + s a + r b - r s delta,     h(tau) Z(tau) = A(tau) B(tau) - C(tau)

    with X(tau) = sum_i w_i X_i(tau) and Z = x^m - 1.  Also returns
    ic = sum_{i <= nPublic} w_i (beta A_i + alpha B_i + C_i)(tau), the public
    input term (gamma times the verifying key's IC combination).  The
    verification equation e(A, B) = e(alpha, beta) e(IC, gamma) e(C, delta)
    holds in the exponent iff a b = alpha beta + ic + c delta."""
    tau, alpha, beta, gamma, delta = td
    m = 1 << log2m
    w = [x % RR for x in witness]
    At, Bt, Ct = (qap_at(t, w, log2m, tau) for t in (triples_a, triples_b, triples_c))
    pub = [[e for e in t if e[1] <= n_public] for t in (triples_a, triples_b, triples_c)]
    Ap, Bp, Cp = (qap_at(t, w, log2m, tau) for t in pub)
    hz = (At * Bt - Ct) % RR
    priv = (beta * (At - Ap) + alpha * (Bt - Bp) + (Ct - Cp)) % RR
    a = (alpha + At + rnd * delta) % RR
    b = (beta + Bt + snd * delta) % RR
    dinv = pow(delta, -1, RR)
    c = ((priv + hz) * dinv + snd * a + rnd * b - rnd * snd * delta) % RR
    ic = (beta * Ap + alpha * Bp + Cp) % RR
    assert pow(tau, m, RR) != 1, "tau lies in the evaluation domain"
    return a, b, c, ic


def groth16_check(proof, triples_a, triples_b, triples_c, witness, n_public, log2m, td, rnd, snd) -> bool:
    """True iff the affine proof (A, B, C) is the Groth16 proof of the witness
    for randomness (r, s) under the trapdoor td, and that proof satisfies the
    verification equation (checked in the exponent, where the trapdoor makes
    pairings unnecessary)."""
    a, b, c, ic = groth16_exponents(triples_a, triples_b, triples_c, witness, n_public, log2m, td, rnd, snd)
    _, alpha, beta, _, delta = td
    if (a * b - (alpha * beta + ic + c * delta)) % RR:
        return False
    A, B, C = proof
    return A == G1.mul(a, G1_GEN) and B == G2.mul(b, G2_GEN) and C == G1.mul(c, G1_GEN)


def reference_rs(seed: int):
    """r, s as the reference draws them: random.seed(seed) then two
    random.randint(1, r-1) (src/python/groth_prover.py:117-123, 1249-1261)."""
    rng = random.Random(seed)
    return rng.randint(1, RR - 1), rng.randint(1, RR - 1)
