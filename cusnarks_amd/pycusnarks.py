"""Drop-in for cusnarks' ``pycusnarks`` binding, backed by the MI355X library.

Mirrors the names, argument meaning and return shapes of the reference's
Cython module (cusnarks src/cython/_cusnarks_kernel.pyx) for the GPU
multi-exponentiation and polynomial path, so reference-style callers such as
``ec_mad_cuda2`` (src/python/cuda_wrapper.py:216-284) and the prover's host
calls (src/python/groth_prover.py) run unchanged:

  classes    U256, ECBN128, EC2BN128, ZCUPoly with kernelLaunch / streamSync /
             streamDel (pyx:56-304)
  functions  ec_jacreduce_h, ec_jac2aff_h, ec2_jac2aff_h, ec_jacaddreduce_h,
             ec2_jacaddreduce_h, ntt_interpolandmul_h, mpoly_eval_h,
             to_montgomeryN_h, from_montgomeryN_h, montmultN_h, montmult_neg_h,
             montsquare_h, init_h, release_h, get_nprocs_h (pyx:306-1133)

Differences a caller can observe, all on error paths: the reference calls
exit(1) on CUDA errors and ``assert False`` on bad lengths; here every failure
raises (CskError / ValueError / NotImplementedError) and the process lives.
Every arithmetic call runs on the GPU; there is no CPU fallback.
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import _capi
from ._capi import CskError, P, check, lib
from .device import DeviceBuffer

# ---------------------------------------------------------------- constants
# (src/cuda/types.h:36-131, 228-424; src/python/constants.py)
NWORDS_256BIT = 8
MOD_GROUP, MOD_FIELD = 0, 1
FMT_EXT, FMT_MONT = 0, 1
U256_NDIMS = 1
ECP_JAC_INDIMS, ECP_JAC_OUTDIMS = 2, 3
ECP2_JAC_INDIMS, ECP2_JAC_OUTDIMS = 4, 6
DEFAULT_U256_BSELM = 6
DEFAULT_U256_BSELM_CUDA = 8
MAX_NCORES_OMP = 32
N_STREAMS_PER_GPU = 1 + 4

(CB_U256_ADDM, CB_U256_SUBM, CB_U256_MOD, CB_U256_MULM, CB_U256_ADDM_REDUCE,
 CB_U256_ADDM_REDUCE_SHFL, CB_U256_SHR1, CB_U256_SHL1, CB_U256_SHL, CB_U256_ALMINV,
 CB_U256_MULM2) = range(11)
(CB_EC_JACAFF_ADD, CB_EC_JAC_ADD, CB_EC_JACAFF_DOUBLE, CB_EC_JAC_DOUBLE, CB_EC_JAC_MUL,
 CB_EC_JAC_MUL1, CB_EC_JAC_MAD, CB_EC_JAC_MAD_SHFL, CB_EC_JAC_MUL_OPT, CB_EC_JAC_RED,
 CB_EC_JAC_MUL_PRECOMP) = range(11)
(CB_EC2_JACAFF_ADD, CB_EC2_JAC_ADD, CB_EC2_JACAFF_DOUBLE, CB_EC2_JAC_DOUBLE, CB_EC2_JAC_MUL,
 CB_EC2_JAC_MUL1, CB_EC2_JAC_MAD, CB_EC2_JAC_MAD_SHFL, CB_EC2_JAC_MUL_OPT, CB_EC2_JAC_RED) = range(10)

_MONT_ONE_Q = np.array([3314486685, 3546104717, 4123462461, 175696680,
                        2021213740, 1718526831, 2584207151, 235567041], dtype=np.uint32)


def _ctx(gpu_id: int = 0):
    return _capi.default_context(gpu_id)


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


# ---------------------------------------------------------------- field helpers
def _field_op(op: int, a: np.ndarray, b: np.ndarray | None, pidx: int, words: int = 8) -> np.ndarray:
    a = _u32(a).reshape(-1)
    n = a.size // words
    if n == 0:
        return np.zeros((0, NWORDS_256BIT), dtype=np.uint32)
    ctx = _ctx()
    da = DeviceBuffer.from_array(ctx, a)
    db = DeviceBuffer.from_array(ctx, _u32(b).reshape(-1)) if b is not None else None
    dout = DeviceBuffer(ctx, a.nbytes)
    check(lib().csk_field_op(ctx.ptr, op, pidx, da.ptr, db.ptr if db else None, dout.ptr, n, None),
          "csk_field_op")
    return dout.download((n * words // 8, 8))


def to_montgomeryN_h(in_v, pidx):
    """pyx:1087-1094 -- x -> x*R mod p, elementwise."""
    return _field_op(_capi.OP_TOMONT, in_v, None, pidx)


def from_montgomeryN_h(in_v, pidx, strip_last=0):
    """pyx:1096-1103 (ff.cpp:503-535) -- x*R -> x, elementwise.  strip_last 1
    reads the input as Jacobian G1 points (3 elements each) and packs their
    (X, Y), strip_last 2 as G2 points (6 -> 4); the output keeps the input's
    length, zero after the packed rows, as the reference returns it."""
    a = _u32(in_v).reshape(-1, NWORDS_256BIT)
    conv = _field_op(_capi.OP_FROMMONT, a, None, pidx)
    if not strip_last:
        return conv
    per, keep = (3, 2) if int(strip_last) == 1 else (6, 4)
    sel = (np.arange(a.shape[0]) % per) < keep
    out = np.zeros_like(a)
    out[:int(sel.sum())] = conv[sel]
    return out


def montmultN_h(in_veca, in_vecb, pidx):
    """pyx:317-327 -- elementwise Montgomery products."""
    return _field_op(_capi.OP_MULM, in_veca, in_vecb, pidx)


def montsquare_h(in_veca, pidx):
    """pyx:310-315."""
    return _field_op(_capi.OP_MULM, in_veca, in_veca, pidx)


def montmult_neg_h(in_veca, in_vecb, pidx):
    """pyx:324-331 -- -(a*b*R^-1) mod p (the prover's -r*s scalar)."""
    m = _field_op(_capi.OP_MULM, in_veca, in_vecb, pidx)
    return _field_op(_capi.OP_SUBM, np.zeros_like(m), m, pidx).reshape(-1)


def montinv_h(in_veca, pidx):
    return _field_op(_capi.OP_INVM, in_veca, None, pidx)


# ---------------------------------------------------------------- curve helpers
def _jac2aff(in_v, ec2: bool, strip_last: int) -> np.ndarray:
    nw = 16 if ec2 else 8
    a = _u32(in_v).reshape(-1)
    n = a.size // (3 * nw)
    if n == 0:
        return np.zeros((0, NWORDS_256BIT), dtype=np.uint32)
    ctx = _ctx()
    din = DeviceBuffer.from_array(ctx, a[:n * 3 * nw])
    od = 2 if strip_last else 3
    dout = DeviceBuffer(ctx, n * od * nw * 4)
    check(lib().csk_ec_jac2aff(ctx.ptr, 1 if ec2 else 0, din.ptr, n, dout.ptr, strip_last, None),
          "csk_ec_jac2aff")
    return dout.download((n * od * nw // 8, 8))


def ec_jac2aff_h(in_v, pidx=MOD_GROUP, strip_last=0):
    """pyx:811-824 (ec.cpp:963-1012): Jacobian -> affine; Z == 0 -> (0, 1*R[, 0])."""
    return _jac2aff(in_v, False, strip_last)


def ec2_jac2aff_h(in_v, pidx=MOD_GROUP, strip_last=0):
    """pyx:826-838 (ec.cpp:1014-1062)."""
    return _jac2aff(in_v, True, strip_last)


def _msm(scl: np.ndarray, pts: np.ndarray, ec2: bool) -> np.ndarray:
    from .msm import msm_host
    return msm_host(scl, pts, g2=ec2, ctx=_ctx())


def _affine_to_output(aff: np.ndarray, ec2: bool, strip_last: int) -> np.ndarray:
    """Affine Montgomery result -> the reference's output rows: affine (X, Y)
    when strip_last, else Jacobian (X, Y, 1) or infinity (0, 1, 0)."""
    nw = 16 if ec2 else 8
    aff = _u32(aff).reshape(-1)
    if strip_last:
        return aff.reshape(-1, NWORDS_256BIT)
    z = np.zeros(nw, dtype=np.uint32)
    if aff[:nw].any() or not _is_one(aff[nw:2 * nw]):
        z[:8] = _MONT_ONE_Q
    return np.concatenate([aff, z]).reshape(-1, NWORDS_256BIT)


def _is_one(v: np.ndarray) -> bool:
    return bool(np.array_equal(v[:8], _MONT_ONE_Q) and not v[8:].any())


def ec_jacreduce_h(inscl, intbl, fname=b"", offset=0, total_words=0, order=DEFAULT_U256_BSELM,
                   ec2=0, pidx=MOD_GROUP, to_aff=1, add_in=1, strip_last=1, pippen_conf=0):
    """pyx:992-1053 (ec.cpp:1536-1704): sum k_i P_i of normal-form scalars and
    affine Montgomery points.  Computed by the GPU Pippenger multiexp; order
    and pippen_conf (CPU Straus tuning knobs) do not change the result and are
    accepted for compatibility.  Precomputed-table files (fname) are not
    supported."""
    if fname:
        raise NotImplementedError("precomputed EC table files are not supported by this backend")
    nw = 32 if ec2 else 16
    scl = _u32(inscl).reshape(-1, 8)
    pts = _u32(intbl).reshape(-1)
    n = scl.shape[0]
    if pts.size < n * nw:
        raise ValueError("intbl holds fewer points than scalars")
    aff = _msm(scl, pts[:n * nw].reshape(n, nw), bool(ec2))
    return _affine_to_output(aff, bool(ec2), strip_last)


def _jacaddreduce(inv, ec2: bool, to_aff, add_in, strip_last):
    nw = 16 if ec2 else 8
    a = _u32(inv).reshape(-1)
    if add_in:
        aff = a.reshape(-1, 2 * nw)
    else:
        aff = _jac2aff(a, ec2, 1).reshape(-1, 2 * nw)
    ones = np.zeros((aff.shape[0], 8), dtype=np.uint32)
    ones[:, 0] = 1
    res = _msm(ones, aff, ec2)
    return _affine_to_output(res, ec2, strip_last)


def ec_jacaddreduce_h(inv, pidx=MOD_GROUP, to_aff=1, add_in=0, strip_last=0):
    """pyx:949-969 (ec.cpp:1210-1264): sum of n points (affine if add_in,
    else Jacobian)."""
    return _jacaddreduce(inv, False, to_aff, add_in, strip_last)


def ec2_jacaddreduce_h(inv, pidx=MOD_GROUP, to_aff=1, add_in=0, strip_last=0):
    """pyx:971-989."""
    return _jacaddreduce(inv, True, to_aff, add_in, strip_last)


# ---------------------------------------------------------------- polynomials
def ntt_interpolandmul_h(inva, invb, invroots, rstride, pidx=MOD_FIELD):
    """pyx:479-504 (ntt.cpp:672-1070): the 2m coefficients (normal form) of
    A(x)B(x), A and B interpolated from their m evaluations (Montgomery) on
    the m-th roots of unity.  invroots/rstride describe the reference's root
    table; the GPU derives the same roots from the field constants."""
    from .poly import interpolandmul_host
    return interpolandmul_host(_u32(inva).reshape(-1, 8), _u32(invb).reshape(-1, 8), ctx=_ctx())


def mpoly_eval_h(scldata, pdata, reduce_coeff, ncoeff, start_idx, last_idx, max_threads, pidx=MOD_FIELD):
    """pyx:452-477 (mpoly.cpp:125-168): per-constraint evaluation
    out[row] = sum_wire coeff * toMont(w[wire]) of the per-wire sparse polys
    in pdata ([npoly, ncoeff_0.., per poly: rows.., values (8 words)..])."""
    from .poly import mpoly_eval_host
    return mpoly_eval_host(_u32(scldata).reshape(-1, 8), _u32(pdata).reshape(-1), int(ncoeff),
                           int(start_idx), int(last_idx), ctx=_ctx())


# ---------------------------------------------------------------- misc host API
def init_h():
    """pyx:1129 -- binds device 0 (scratch is allocated lazily)."""
    _ctx(0)


def release_h():
    """pyx:1132."""
    for ctx in list(_capi._default_ctx.values()):
        ctx.close()
    _capi._default_ctx.clear()


def get_nprocs_h():
    return os.cpu_count() or 1


# ---------------------------------------------------------------- GPU classes
class CUSnarks:
    """pyx:56-238: base of the kernel-launch classes.  The constructor's
    in_len bounds the accepted input and output rows (the reference
    preallocates device buffers of that size per GPU and stream:
    in_len * (indims + 1) input and in_len * outdims output rows for the EC
    classes, src/cuda/ecbn128.cu:64-81)."""

    _IN_MULT = 1
    _OUT_MULT = 1

    def __init__(self, in_len, out_len=0, in_size=0, out_size=0, seed=0):
        self.in_dim = int(in_len) * self._IN_MULT
        self.out_dim = (int(out_len) if out_len else int(in_len)) * self._OUT_MULT
        self._t0 = time.time()
        self._pending: dict[tuple[int, int], np.ndarray] = {}

    def elapsedTime(self):
        return time.time() - self._t0

    def kernelLaunch(self, in_vec, config, params, gpu_id=0, stream_id=0, n_kernels=1):
        """Run the launch sequence described by config/params on in_vec (rows
        of 8 u32).  Returns (out rows, seconds).  With stream_id > 0 the
        result is also kept for streamSync(gpu_id, stream_id), as the
        reference's asynchronous streams deliver it."""
        in_vec = _u32(in_vec)
        if in_vec.ndim != 2 or in_vec.shape[1] != NWORDS_256BIT:
            raise ValueError("in_vec must be a (rows, 8) uint32 array")
        out_len = int(params["out_length"])
        if in_vec.shape[0] > self.in_dim or out_len > self.out_dim:
            raise ValueError("Incorrect arguments: %d rows > %d or out %d > %d"
                             % (in_vec.shape[0], self.in_dim, out_len, self.out_dim))
        t0 = time.time()
        out = self._launch(in_vec, config, params, int(gpu_id), int(n_kernels))
        out = np.ascontiguousarray(out[:out_len]) if out.shape[0] >= out_len else out
        if stream_id > 0:
            self._pending[(int(gpu_id), int(stream_id))] = out
        return np.copy(out), time.time() - t0

    def streamSync(self, gpu_id, stream_id):
        """pyx:197-212: the output of the last launch on (gpu, stream)."""
        out = self._pending.pop((int(gpu_id), int(stream_id)), None)
        if out is None:
            out = np.zeros((0, NWORDS_256BIT), dtype=np.uint32)
        return out, 0.0

    def streamDel(self, gpu_id, stream_id):
        self._pending.pop((int(gpu_id), int(stream_id)), None)

    def getDeviceInfo(self):
        n = _capi.ctypes.c_int(0)
        check(lib().csk_device_count(_capi.ctypes.byref(n)), "csk_device_count")
        return {"n_devices": n.value}

    def _launch(self, in_vec, config, params, gpu_id, n_kernels):
        raise NotImplementedError


class U256(CUSnarks):
    """pyx:240-251 over the U256 kernels (u256_device.cu): elementwise modular
    add/sub/Montgomery-multiply of interleaved pairs [x0, y0, x1, y1, ...]."""

    _OPS = {CB_U256_ADDM: _capi.OP_ADDM, CB_U256_SUBM: _capi.OP_SUBM, CB_U256_MULM: _capi.OP_MULM}

    def _launch(self, in_vec, config, params, gpu_id, n_kernels):
        kidx = int(config["kernel_idx"][0])
        if kidx not in self._OPS:
            raise NotImplementedError("U256 kernel %d is not provided by this backend" % kidx)
        pidx = int(params["midx"][0])
        pairs = in_vec[: (in_vec.shape[0] // 2) * 2].reshape(-1, 2, 8)
        return _field_op(self._OPS[kidx], pairs[:, 0], pairs[:, 1], pidx)


class _ECBase(CUSnarks):
    _EC2 = False

    def _launch(self, in_vec, config, params, gpu_id, n_kernels):
        ec2 = self._EC2
        kseq = [int(k) for k in np.asarray(config["kernel_idx"]).reshape(-1)[:max(1, n_kernels)]]
        kidx = kseq[0]
        indims = ECP2_JAC_INDIMS if ec2 else ECP_JAC_INDIMS
        mad = (CB_EC2_JAC_MAD_SHFL, CB_EC2_JAC_MAD) if ec2 else (CB_EC_JAC_MAD_SHFL, CB_EC_JAC_MAD)
        red = CB_EC2_JAC_RED if ec2 else CB_EC_JAC_RED
        mul_opt = CB_EC2_JAC_MUL_OPT if ec2 else CB_EC_JAC_MUL_OPT
        # ec_mad_cuda2(separate_k=True), the prover's own call (groth_prover.py:1514,
        # cuda_wrapper.py:251-259): per-point products, then reduction kernels
        # down to one point -- the same multiexp over the same packed vector
        if kidx == mul_opt and len(kseq) > 1 and all(k == red for k in kseq[1:]):
            kidx = mad[0]
        if kidx in mad:
            # [n scalars | n affine points], n = in_length / (indims + 1) (cuda_wrapper.py:216-284)
            n = int(params["in_length"][0]) // (indims + U256_NDIMS)
            scl = in_vec[:n]
            pts = in_vec[n:n + n * indims].reshape(n, indims * 8)
            aff = _msm(scl, pts, ec2)
            return _affine_to_output(aff, ec2, 0)
        if kidx == red:
            outd = ECP2_JAC_OUTDIMS if ec2 else ECP_JAC_OUTDIMS
            n = int(params["in_length"][0]) // outd
            return _jacaddreduce(in_vec[:n * outd], ec2, 1, 0, 0)
        raise NotImplementedError("EC kernel %d is not provided by this backend" % kidx)


class ECBN128(_ECBase):
    """pyx:253-268: G1 kernels; 3 input and 3 output rows per constructor length."""
    _IN_MULT = ECP_JAC_INDIMS + U256_NDIMS
    _OUT_MULT = ECP_JAC_OUTDIMS


class EC2BN128(_ECBase):
    """pyx:271-286: G2 kernels; 5 input and 6 output rows per constructor length."""
    _IN_MULT = ECP2_JAC_INDIMS + U256_NDIMS
    _OUT_MULT = ECP2_JAC_OUTDIMS
    _EC2 = True


class ZCUPoly(CUSnarks):
    """pyx:290-304.  The reference's GPU polynomial kernels are disabled in its
    prover (groth_prover.py:130); the polynomial path is served by
    ntt_interpolandmul_h and the NTT entry points of cusnarks_amd.poly."""

    def _launch(self, in_vec, config, params, gpu_id, n_kernels):
        raise NotImplementedError("use ntt_interpolandmul_h / cusnarks_amd.poly for the polynomial path")


__all__ = [n for n in dir() if not n.startswith("_")] + ["CskError", "P"]
