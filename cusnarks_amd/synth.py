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
"""
from __future__ import annotations

import math

import numpy as np

from . import inputs
from .poly import CSR


class Circuit:
    def __init__(self, n_constraints: int, n_vars: int, n_public: int = 1, per_row: int = 3, seed: int = 0):
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
        self.A = self._matrix(rows, rng.integers(0, n_vars, size=nnz), seed * 16 + 1, extend=True)
        self.B = self._matrix(rows, rng.integers(0, n_vars, size=nnz), seed * 16 + 2, extend=False)
        self.C = self._matrix(rows, rng.integers(0, n_vars, size=nnz), seed * 16 + 3, extend=False)
        w = inputs.field_elements(seed * 16 + 4, n_vars)
        w[0] = 0
        w[0, 0] = 1
        self.witness = w

    def _matrix(self, rows, cols, vseed, extend):
        vals = inputs.field_elements(vseed, rows.size)
        if extend:
            npub = self.n_public + 1
            one_mont = inputs.from_ints([(1 << 256) % inputs.R])
            rows = np.concatenate([rows, self.n_constraints + np.arange(npub)])
            cols = np.concatenate([cols, np.arange(npub)])
            vals = np.concatenate([vals, np.repeat(one_mont, npub, axis=0)])
        return CSR.from_triples(rows, cols, vals, self.m)


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
