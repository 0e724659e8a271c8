"""Fr NTT / INTT, the QAP H step and QAP evaluation on the GPU.

Data are (n, 8) uint32 rows (8 little-endian u32 limbs per element).  NTT
inputs and outputs are Montgomery form; INTT output is Montgomery (fmt=1) or
normal (fmt=0), as the reference's intt_h (cusnarks src/cuda/ntt.cpp:583-603).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi
from ._capi import check, lib
from .device import DeviceBuffer, dptr


def _log2(n: int) -> int:
    lg = n.bit_length() - 1
    if n <= 0 or (1 << lg) != n:
        raise ValueError("length must be a power of two, got %d" % n)
    return lg


def ntt_device(d_in, d_out, log2n: int, inverse: bool = False, fmt: int = 1, ctx=None, stream=None) -> None:
    ctx = ctx or _capi.default_context()
    check(lib().csk_ntt(ctx.ptr, dptr(d_in), dptr(d_out), log2n, 1 if inverse else 0, fmt, stream), "csk_ntt")


def ntt_host(a: np.ndarray, inverse: bool = False, fmt: int = 1, ctx=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint32).reshape(-1, 8)
    lg = _log2(a.shape[0])
    ctx = ctx or _capi.default_context()
    d = DeviceBuffer.from_array(ctx, a)
    ntt_device(d, d, lg, inverse, fmt, ctx=ctx)
    return d.download(a.shape)


def interpolandmul_device(d_pA, d_pB, log2m: int, d_out, ctx=None, stream=None) -> None:
    ctx = ctx or _capi.default_context()
    check(lib().csk_interpolandmul(ctx.ptr, dptr(d_pA), dptr(d_pB), log2m, dptr(d_out), stream),
          "csk_interpolandmul")


def interpolandmul_batch_device(d_pAs, d_pBs, log2m: int, d_outs, ctx=None, stream=None) -> None:
    """H steps of several same-size problems at once (csk_interpolandmul_batch):
    d_outs[t] = the 2m coefficients of A_t(x) B_t(x), as interpolandmul_device."""
    ctx = ctx or _capi.default_context()
    k = len(d_pAs)
    if len(d_pBs) != k or len(d_outs) != k:
        raise ValueError("member lists differ in length")
    arr = lambda xs: (ctypes.c_void_p * max(1, k))(*[dptr(x) for x in xs])  # noqa: E731
    check(lib().csk_interpolandmul_batch(ctx.ptr, k, arr(d_pAs), arr(d_pBs), log2m, arr(d_outs), stream),
          "csk_interpolandmul_batch")


def interpolandmul_host(pA: np.ndarray, pB: np.ndarray, ctx=None) -> np.ndarray:
    """2m normal-form coefficients of A(x)B(x) (the reference's
    ntt_interpolandmul_h result; the prover keeps rows m .. 2m-2)."""
    pA = np.ascontiguousarray(pA, dtype=np.uint32).reshape(-1, 8)
    pB = np.ascontiguousarray(pB, dtype=np.uint32).reshape(-1, 8)
    if pA.shape != pB.shape:
        raise ValueError("pA and pB differ in length")
    lg = _log2(pA.shape[0])
    ctx = ctx or _capi.default_context()
    da, db = DeviceBuffer.from_array(ctx, pA), DeviceBuffer.from_array(ctx, pB)
    dout = DeviceBuffer(ctx, 2 * pA.nbytes)
    interpolandmul_device(da, db, lg, dout, ctx=ctx)
    return dout.download((2 * pA.shape[0], 8))


# ------------------------------------------------------------------ QAP evaluation
class CSR:
    """Row-major sparse matrix over Fr: rowptr (nrows+1) u64, col u32, val (nnz, 8)
    Montgomery coefficients."""

    def __init__(self, rowptr, col, val, nrows):
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.uint64)
        self.col = np.ascontiguousarray(col, dtype=np.uint32)
        self.val = np.ascontiguousarray(val, dtype=np.uint32).reshape(-1, 8)
        self.nrows = int(nrows)

    @classmethod
    def from_triples(cls, rows, cols, vals, nrows):
        rows = np.asarray(rows, dtype=np.int64)
        order = np.argsort(rows, kind="stable")
        counts = np.bincount(rows, minlength=nrows)
        rowptr = np.zeros(nrows + 1, dtype=np.uint64)
        np.cumsum(counts, out=rowptr[1:])
        return cls(rowptr, np.asarray(cols)[order], np.asarray(vals).reshape(-1, 8)[order], nrows)


def mpoly_to_csr(pdata: np.ndarray, ncoeff: int, start_idx: int, last_idx: int) -> tuple[CSR, np.ndarray]:
    """Parse the reference's per-wire sparse format
    [npoly, ncoeff_0 .. ncoeff_{npoly-1}, then per poly: rows (ncoeff_i u32),
    values (ncoeff_i x 8 u32)] (cusnarks src/cuda/mpoly.cpp:125-168, 195-200)
    for polys start_idx .. last_idx-1 into a row-major CSR over `ncoeff` rows.
    Returns the CSR and the wire index of each entry's scalar."""
    pdata = np.ascontiguousarray(pdata, dtype=np.uint32)
    npoly = int(pdata[0])
    nc = pdata[1:1 + npoly].astype(np.int64)
    # start of poly i's block: 1 + npoly + 9 * sum_{j<i} nc_j
    starts = 1 + npoly + 9 * np.concatenate([[0], np.cumsum(nc)[:-1]])
    sel = np.arange(start_idx, last_idx)
    cnt = nc[sel]
    total = int(cnt.sum())
    wire = np.repeat(sel, cnt)
    within = np.arange(total) - np.repeat(np.concatenate([[0], np.cumsum(cnt)[:-1]]), cnt)
    base = np.repeat(starts[sel], cnt)
    rows = pdata[base + within].astype(np.int64)
    if total and int(rows.max()) >= ncoeff:
        raise ValueError("a coefficient's row index is outside the %d rows" % ncoeff)
    vbase = base + np.repeat(cnt, cnt) + 8 * within
    vals = pdata[vbase[:, None] + np.arange(8)[None, :]]
    csr = CSR.from_triples(rows, wire, vals, ncoeff)
    return csr, wire


def mpoly_eval_device(csr_dev: dict, d_wit, d_out, ctx=None, stream=None) -> None:
    ctx = ctx or _capi.default_context()
    check(lib().csk_mpoly_eval_csr(ctx.ptr, dptr(csr_dev["rowptr"]), dptr(csr_dev["col"]), dptr(csr_dev["val"]),
                                   dptr(d_wit), csr_dev["nrows"], dptr(d_out), stream), "csk_mpoly_eval_csr")


def upload_csr(csr: CSR, ctx=None) -> dict:
    ctx = ctx or _capi.default_context()
    return {
        "rowptr": DeviceBuffer.from_array(ctx, csr.rowptr),
        "col": DeviceBuffer.from_array(ctx, csr.col if csr.col.size else np.zeros(1, np.uint32)),
        "val": DeviceBuffer.from_array(ctx, csr.val if csr.val.size else np.zeros((1, 8), np.uint32)),
        "nrows": csr.nrows,
    }


def mpoly_eval_host(scl: np.ndarray, pdata: np.ndarray, ncoeff: int, start_idx: int, last_idx: int,
                    ctx=None) -> np.ndarray:
    """mpoly_eval_h: out[row] = sum over wires w in [start_idx, last_idx) of
    coeff * toMont(scl[w]); (ncoeff, 8) Montgomery rows."""
    ctx = ctx or _capi.default_context()
    csr, _ = mpoly_to_csr(pdata, ncoeff, start_idx, last_idx)
    scl = np.ascontiguousarray(scl, dtype=np.uint32).reshape(-1, 8)
    dev = upload_csr(csr, ctx)
    d_w = DeviceBuffer.from_array(ctx, scl if scl.size else np.zeros((1, 8), np.uint32))
    d_out = DeviceBuffer(ctx, max(1, ncoeff) * 32)
    mpoly_eval_device(dev, d_w, d_out, ctx=ctx)
    return d_out.download((ncoeff, 8))
