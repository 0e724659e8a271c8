"""Seeded synthetic Groth16 circuits and proving keys.

R1CS with nConstraints rows over nVars wires: A, B and C have `per_row`
nonzeros per row at uniform columns with coefficients uniform mod r; the
witness is uniform mod r with w[0] = 1.  The QAP polynomials follow the
reference's layout (cusnarks src/cuda/mpoly.cpp:201-263, setup
src/python/groth_setup.py:434-452): polsA gets the public-input rows
nConstraints + i with coefficient 1 on wire i (i = 0 .. nPublic), the domain
is m = 2^ceil(log2(nConstraints + nPublic))
(src/python/pysnarks_utils.py:398-403).  Coefficients are generated directly
as Montgomery representations (a uniform element's Montgomery form is
uniform).  Proving-key points are seeded random multiples of the G1 / G2
generators, made on the GPU by the fixed-base kernel.

With ``satisfied=True`` the C coefficients are solved so that every
constraint holds, (A w)_row (B w)_row = (C w)_row, and
``KeyScalars.from_trapdoor`` derives the key from known toxic scalars
(tau, alpha, beta, gamma, delta) as the reference's setup does
(src/python/groth_setup.py:488-787, evalLagrangePoly 454-485,
GrothSetupComputePS_h / GrothSetupComputeeT_h
src/cython/_cusnarks_kernel.pyx:756-809).  Proofs over such a key are valid
Groth16 proofs, which lets a checker that knows the trapdoor verify them.
That setup runs in Python integers and is meant for circuits up to about
2^16 constraints.
"""
from __future__ import annotations

import math

import numpy as np

from . import inputs
from .poly import CSR


class Circuit:
    def __init__(self, n_constraints: int, n_vars: int, n_public: int = 1, per_row: int = 3, seed: int = 0,
                 satisfied: bool = False):
        if n_vars < n_public + 2:
            raise ValueError("need at least nPublic + 2 wires")
        self.n_constraints = n_constraints
        self.n_vars = n_vars
        self.n_public = n_public
        self.seed = seed
        self.log2m = max(1, math.ceil(math.log2(n_constraints + n_public)))
        self.m = 1 << self.log2m
        if n_constraints + n_public >= self.m:
            # the reference's public-input rows would fall outside the domain
            raise ValueError("nConstraints + nPublic must not be a power of two")
        rng = np.random.default_rng(seed)
        nnz = n_constraints * per_row
        rows = np.repeat(np.arange(n_constraints, dtype=np.int64), per_row)
        self.satisfied = satisfied
        w = inputs.field_elements(seed * 16 + 4, n_vars)
        w[0] = 0
        w[0, 0] = 1
        self.witness = w
        self.A = self._matrix(rows, rng.integers(0, n_vars, size=nnz), seed * 16 + 1, extend=True)
        self.B = self._matrix(rows, rng.integers(0, n_vars, size=nnz), seed * 16 + 2, extend=False)
        ccols = rng.integers(0, n_vars, size=nnz)
        cvals = inputs.field_elements(seed * 16 + 3, nnz)
        if satisfied:
            cvals = self._solve_c(rows, ccols, cvals, per_row)
        self.C = CSR.from_triples(rows, ccols, cvals, self.m)

    def _solve_c(self, rows, ccols, cvals, per_row):
        """Replace the last C coefficient of every row so that the row holds:
        c_last = ((A w)_row (B w)_row - sum of the other C terms) / w[col_last]
        (col_last is moved to wire 0, w = 1, where its witness value is 0)."""
        n, R = self.n_constraints, inputs.R
        w = inputs.to_ints(self.witness)
        a = _row_sums(self.A, w, n)
        b = _row_sums(self.B, w, n)
        cv = [v * _RINV % R for v in inputs.to_ints(cvals)]
        out = list(cv)
        for r in range(n):
            base = r * per_row
            last = base + per_row - 1
            if w[int(ccols[last])] == 0:
                ccols[last] = 0
            rest = sum(cv[k] * w[int(ccols[k])] for k in range(base, last))
            x = (a[r] * b[r] - rest) * pow(w[int(ccols[last])], -1, R) % R
            out[last] = x
        return inputs.from_ints([x * _RMONT % R for x in out])

    def triples(self, name: str):
        """(row, wire, coefficient) of matrix A, B or C, coefficients in normal form."""
        csr = getattr(self, name)
        vals = [v * _RINV % inputs.R for v in inputs.to_ints(csr.val)]
        rows = np.repeat(np.arange(csr.nrows), np.diff(csr.rowptr.astype(np.int64)))
        return [(int(r), int(c), v) for r, c, v in zip(rows, csr.col, vals)]

    def _matrix(self, rows, cols, vseed, extend):
        vals = inputs.field_elements(vseed, rows.size)
        if extend:
            npub = self.n_public + 1
            one_mont = inputs.from_ints([(1 << 256) % inputs.R])
            rows = np.concatenate([rows, self.n_constraints + np.arange(npub)])
            cols = np.concatenate([cols, np.arange(npub)])
            vals = np.concatenate([vals, np.repeat(one_mont, npub, axis=0)])
        return CSR.from_triples(rows, cols, vals, self.m)


_RMONT = (1 << 256) % inputs.R
_RINV = pow(_RMONT, -1, inputs.R)


def _row_sums(csr: CSR, w, nrows):
    """(M w)_row for the first nrows rows, normal-form Python ints."""
    R = inputs.R
    vals = inputs.to_ints(csr.val)
    out = [0] * nrows
    rp = csr.rowptr
    for r in range(nrows):
        acc = 0
        for e in range(int(rp[r]), int(rp[r + 1])):
            acc += vals[e] * w[int(csr.col[e])]
        out[r] = acc * _RINV % R
    return out


def _root_of_unity(log2n: int) -> int:
    return pow(5, (inputs.R - 1) >> log2n, inputs.R)


def _batch_inv(xs):
    """Inverses of nonzero field elements with one modular inversion."""
    R = inputs.R
    pre = [1] * (len(xs) + 1)
    for i, x in enumerate(xs):
        pre[i + 1] = pre[i] * x % R
    inv = pow(pre[-1], -1, R)
    out = [0] * len(xs)
    for i in range(len(xs) - 1, -1, -1):
        out[i] = pre[i] * inv % R
        inv = inv * xs[i] % R
    return out


class Trapdoor:
    """Toxic scalars of a Groth16 setup: tau, alpha, beta, gamma, delta
    (random nonzero field elements, as groth_setup.py:520-525 draws them)."""

    def __init__(self, seed: int = 0):
        vals = inputs.to_ints(inputs.field_elements(seed * 64 + 33, 5, inputs.R, nonzero=True))
        self.tau, self.alpha, self.beta, self.gamma, self.delta = vals


def lagrange_at(tau: int, log2m: int):
    """L_j(tau), j < m, for the domain <w_m>: (tau^m - 1)/m * w^j / (tau - w^j)
    (evalLagrangePoly, groth_setup.py:454-485 and pyx:756-771)."""
    R, m = inputs.R, 1 << log2m
    w = _root_of_unity(log2m)
    zt = (pow(tau, m, R) - 1) % R
    if zt == 0:
        raise ValueError("tau lies in the evaluation domain")
    l = zt * pow(m, -1, R) % R
    wj, pw = [], 1
    for _ in range(m):
        wj.append(pw)
        pw = pw * w % R
    inv = _batch_inv([(tau - x) % R for x in wj])
    return [l * x % R * y % R for x, y in zip(wj, inv)]


def wire_evals(circuit: "Circuit", tau: int):
    """(A_i(tau), B_i(tau), C_i(tau)) per wire: sum over the wire's entries of
    coefficient * L_row(tau) (mpoly_madd_h over the Lagrange values,
    groth_setup.py:771-787)."""
    R = inputs.R
    L = lagrange_at(tau, circuit.log2m)
    out = []
    for name in ("A", "B", "C"):
        ev = [0] * circuit.n_vars
        for r, c, v in circuit.triples(name):
            ev[c] = (ev[c] + v * L[r]) % R
        out.append(ev)
    return out


class KeyScalars:
    """The discrete logs behind a synthetic proving key (k * generator)."""

    def __init__(self, circuit: Circuit, seed: int = 0):
        nv, m = circuit.n_vars, circuit.m
        f = lambda s, n: inputs.field_elements(seed * 64 + s, n, inputs.R, nonzero=True)  # noqa: E731
        # A, B1, B2 carry nVars points + the two appended (alpha|beta, delta)
        self.alpha, self.beta, self.delta = f(1, 1), f(2, 1), f(3, 1)
        self.A = np.concatenate([f(4, nv), self.alpha, self.delta])
        self.B1 = np.concatenate([f(5, nv), self.beta, self.delta])
        self.B2 = np.concatenate([f(6, nv), self.beta, self.delta])
        self.C = f(7, nv)
        self.hExps = np.concatenate([f(8, m - 1), self.delta])  # m-1 points + delta1

    @classmethod
    def from_trapdoor(cls, circuit: Circuit, td: Trapdoor) -> "KeyScalars":
        """A Groth16 key from known toxic scalars, with the reference's layout
        (groth_setup.py:575-743; alpha/beta/delta appended as the prover does,
        groth_prover.py:623, 648, 697):
          A_i = A_i(tau), B1_i = B2_i = B_i(tau)                  (+ alpha|beta, delta)
          C_i = (beta A_i + alpha B_i + C_i(tau)) / delta, i > nPublic (0 below)
          hExps_j = tau^j Z(tau) / delta, j < m-1                 (+ delta)
          IC_i = (beta A_i + alpha B_i + C_i(tau)) / gamma, i <= nPublic."""
        R = inputs.R
        if not circuit.satisfied:
            raise ValueError("a trapdoor key is meant for a satisfied circuit")
        self = cls.__new__(cls)
        at, bt, ct = wire_evals(circuit, td.tau)
        dinv, ginv = pow(td.delta, -1, R), pow(td.gamma, -1, R)
        npub = circuit.n_public
        ps = [(td.beta * a + td.alpha * b + c) % R for a, b, c in zip(at, bt, ct)]
        zt = (pow(td.tau, circuit.m, R) - 1) % R
        zd = zt * dinv % R
        hx, t = [], 1
        for _ in range(circuit.m - 1):
            hx.append(t * zd % R)
            t = t * td.tau % R
        f = inputs.from_ints
        self.alpha, self.beta, self.delta = f([td.alpha]), f([td.beta]), f([td.delta])
        self.A = f(at + [td.alpha, td.delta])
        self.B1 = f(bt + [td.beta, td.delta])
        self.B2 = self.B1.copy()
        self.C = f([0] * (npub + 1) + [x * dinv % R for x in ps[npub + 1:]])
        self.hExps = f(hx + [td.delta])
        self.IC = f([x * ginv % R for x in ps[:npub + 1]])
        self.trapdoor = td
        return self
