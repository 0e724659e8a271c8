"""CPU oracle for the cusnarks_amd hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker / the timed CPU baseline.  The
product package (cusnarks_amd) never imports it.

Two layers:
  * ``oracle.pyoracle``  -- closed forms in Python integers (pins everything);
  * this module          -- numpy/ctypes front end of ``liboracle.so``, the C
    restatement of the reference CPU path (Straus multiexp server, radix-2
    NTT, H step, QAP evaluation).  Arrays are uint32 with 8 limbs per element.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    so = os.path.join(_HERE, "liboracle.so")
    cmd = ["make", "-C", _HERE] + (["-B"] if force else [])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        L = ctypes.CDLL(so)
        P = ctypes.c_void_p
        U = ctypes.c_uint32
        L.or_montmult.argtypes = [P, P, P, U]
        L.or_addm.argtypes = [P, P, P, U]
        L.or_subm.argtypes = [P, P, P, U]
        L.or_montinv.argtypes = [P, P, U]
        L.or_montmultN.argtypes = [P, P, P, U, U]
        L.or_to_montgomeryN.argtypes = [P, P, U, U]
        L.or_from_montgomeryN.argtypes = [P, P, U, U]
        L.or_montmult_ext.argtypes = [P, P, P, U]
        L.or_ec_jacadd.argtypes = [P, P, P, U]
        L.or_ec_jacaddmixed.argtypes = [P, P, P, U]
        L.or_ec_jacdouble.argtypes = [P, P, U]
        L.or_ec_jac2aff.argtypes = [P, P, U, U, U]
        L.or_ec_jacreduce.argtypes = [P, P, P, U, U, U, U, P]
        L.or_ec_jacreduce.restype = ctypes.c_int
        L.or_ec_scmul_aff.argtypes = [P, P, P, U, U]
        L.or_roots.argtypes = [P, U]
        L.or_ntt.argtypes = [P, P, U, U]
        L.or_intt.argtypes = [P, P, U, U, U]
        L.or_interpolandmul.argtypes = [P, P, P, P, U, U]
        L.or_interpolandmul_mt.argtypes = [P, P, P, P, U, U]
        L.or_mpoly_eval_csr.argtypes = [P, P, P, P, P, U]
        L.or_poly_eval.argtypes = [P, P, ctypes.c_uint64, P]
        L.or_lagrange_eval.argtypes = [P, P, U, P]
        L.or_lagrange_eval.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


# ------------------------------------------------------------------ field
def montmultN(a, b, pidx):
    a, b = _u32(a).reshape(-1, 8), _u32(b).reshape(-1, 8)
    out = np.empty_like(a)
    lib().or_montmultN(_p(out), _p(a), _p(b), a.shape[0], pidx)
    return out


def to_montgomeryN(a, pidx):
    a = _u32(a).reshape(-1, 8)
    out = np.empty_like(a)
    lib().or_to_montgomeryN(_p(out), _p(a), a.shape[0], pidx)
    return out


def from_montgomeryN(a, pidx):
    a = _u32(a).reshape(-1, 8)
    out = np.empty_like(a)
    lib().or_from_montgomeryN(_p(out), _p(a), a.shape[0], pidx)
    return out


# ------------------------------------------------------------------ curve
def ec_jacreduce(scl, pts, ec2=False, order=6, max_threads=0):
    """Straus multiexp server (ec_jacreduce_server_h). Returns (affine words,
    threads used).  scl: (n,8) normal form; pts: (n,16) or (n,32) affine mont."""
    scl = _u32(scl).reshape(-1, 8)
    nw = 32 if ec2 else 16
    pts = _u32(pts).reshape(-1, nw)
    assert pts.shape[0] == scl.shape[0]
    out = np.zeros(nw, dtype=np.uint32)
    used = ctypes.c_uint32(0)
    rc = lib().or_ec_jacreduce(_p(out), _p(scl), _p(pts), scl.shape[0], order,
                               1 if ec2 else 0, max_threads, ctypes.byref(used))
    if rc != 0:
        raise ValueError("or_ec_jacreduce failed")
    return out, used.value


def ec_scmul_aff(scl, base_aff, ec2=False):
    """k_i * base for every scalar (affine Montgomery words)."""
    scl = _u32(scl).reshape(-1, 8)
    nw = 32 if ec2 else 16
    base = _u32(base_aff).reshape(nw)
    out = np.zeros((scl.shape[0], nw), dtype=np.uint32)
    lib().or_ec_scmul_aff(_p(out), _p(scl), _p(base), scl.shape[0], 1 if ec2 else 0)
    return out


# ------------------------------------------------------------------ NTT / H
def roots(log2n):
    out = np.zeros(((1 << log2n), 8), dtype=np.uint32)
    lib().or_roots(_p(out), log2n)
    return out


def ntt(a, log2n, rstride=1, rts=None):
    a = _u32(a).reshape(-1, 8).copy()
    rts = roots(log2n) if rts is None else _u32(rts)
    lib().or_ntt(_p(a), _p(rts), log2n, rstride)
    return a


def intt(a, log2n, fmt=1, rstride=1, rts=None):
    a = _u32(a).reshape(-1, 8).copy()
    rts = roots(log2n) if rts is None else _u32(rts)
    lib().or_intt(_p(a), _p(rts), log2n, rstride, fmt)
    return a


def interpolandmul(pA, pB, log2m, threads=0, rts=None):
    """2m normal-form coefficients of A(x)B(x) (ntt_interpolandmul_h).
    threads > 0 computes them as the reference's threaded server does
    (4-step transforms over that many threads)."""
    pA, pB = _u32(pA).reshape(-1, 8), _u32(pB).reshape(-1, 8)
    m = 1 << log2m
    assert pA.shape[0] == m and pB.shape[0] == m
    rts = roots(log2m + 1) if rts is None else _u32(rts)
    out = np.zeros((2 * m, 8), dtype=np.uint32)
    fn = lib().or_interpolandmul_mt if threads else lib().or_interpolandmul
    fn(_p(out), _p(pA), _p(pB), _p(rts), log2m, threads)
    return out


def mpoly_eval_csr(rowptr, col, val, wit, nrows):
    rowptr = np.ascontiguousarray(rowptr, dtype=np.uint64)
    col = _u32(col)
    val = _u32(val).reshape(-1, 8)
    wit = _u32(wit).reshape(-1, 8)
    out = np.zeros((nrows, 8), dtype=np.uint32)
    lib().or_mpoly_eval_csr(_p(out), _p(rowptr), _p(col), _p(val), _p(wit), nrows)
    return out


def poly_eval(coeffs, x_mont):
    """Horner value (normal form) at x (Montgomery words) of normal-form coefficients."""
    c = _u32(coeffs).reshape(-1, 8)
    x = _u32(x_mont).reshape(8)
    out = np.zeros(8, dtype=np.uint32)
    lib().or_poly_eval(_p(out), _p(c), c.shape[0], _p(x))
    return out


def lagrange_eval(evals, log2m, z_mont):
    """Value at z (Montgomery words) of the interpolant of 2^log2m Montgomery
    evaluations on the 2^log2m-th roots of unity (Montgomery result)."""
    e = _u32(evals).reshape(-1, 8)
    assert e.shape[0] == 1 << log2m
    z = _u32(z_mont).reshape(8)
    out = np.zeros(8, dtype=np.uint32)
    if lib().or_lagrange_eval(_p(out), _p(e), log2m, _p(z)) != 0:
        raise ValueError("or_lagrange_eval failed (z in the domain or out of memory)")
    return out
