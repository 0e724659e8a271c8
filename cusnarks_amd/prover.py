"""Device-resident Groth16 prover (BN254) on the MI355X kernels.

Composition exactly as the reference prover builds a proof
(cusnarks src/python/groth_prover.py:575-790 pysnarkP_CPU, 1245-1459 gen_proof):

  pA, pB = QAP evaluation of the witness (mpoly_eval_h)          [GPU SpMV]
  polH   = ntt_interpolandmul(pA, pB); H = polH[m .. 2m-2]         [GPU NTT]
  pi_a   = sum w_i A_i + 1*A[nVars] + r*A[nVars+1]                  [G1 multiexp]
  pi_b   = sum w_i B2_i + 1*B2[nVars] + s*B2[nVars+1]               [G2 multiexp]
  pi_b1  = sum w_i B1_i + 1*B1[nVars] + s*B1[nVars+1]               [G1 multiexp]
  pi_c'  = sum_{nPublic < i < nVars} w_i C_i                        [G1 multiexp]
  pi_c2  = sum_j H_j hExps_j + (-r s) delta1                        [G1 multiexp]
  pi_c   = pi_c2 + pi_c' + s pi_a + r pi_b1                         [G1 multiexp, 4 terms]

(A, B1, B2 hold nVars+2 points: the appended (alpha1|beta, delta); hExps
holds m-1 points followed by delta1.)  r and s are passed explicitly; the
reference draws them with random.randint(1, r-1) after random.seed(seed)
(groth_prover.py:117-123, 1249-1261).  Everything stays in HBM; the host
only passes r, s and reads back the three affine points.
"""
from __future__ import annotations

import time

import numpy as np

from . import _capi, inputs
from ._capi import check, lib
from .device import DeviceBuffer
from .msm import fixed_base, msm_device
from .poly import interpolandmul_device, mpoly_eval_device, upload_csr
from .synth import Circuit, KeyScalars

G1W, G2W = 16, 32


class ProvingKey:
    """Proving-key points and QAP matrices resident on one GPU."""

    def __init__(self, circuit: Circuit, keys: KeyScalars, ctx=None, stream=None):
        self.ctx = ctx or _capi.default_context()
        self.n_vars, self.n_public = circuit.n_vars, circuit.n_public
        self.m, self.log2m = circuit.m, circuit.log2m
        self.A = self._points(keys.A, False, stream)
        self.B1 = self._points(keys.B1, False, stream)
        self.B2 = self._points(keys.B2, True, stream)
        self.C = self._points(keys.C, False, stream)
        self.hExps = self._points(keys.hExps, False, stream)
        self.qapA = upload_csr(circuit.A, self.ctx)
        self.qapB = upload_csr(circuit.B, self.ctx)
        check(lib().csk_stream_sync(self.ctx.ptr, stream), "sync")

    def _points(self, scalars, g2, stream):
        n = scalars.shape[0]
        d_k = DeviceBuffer.from_array(self.ctx, scalars)
        out = DeviceBuffer(self.ctx, n * (G2W if g2 else G1W) * 4)
        fixed_base(d_k, n, out, g2=g2, ctx=self.ctx, stream=stream)
        check(lib().csk_stream_sync(self.ctx.ptr, stream), "sync")
        d_k.close()
        return out

    def nbytes(self) -> int:
        pts = (3 * (self.n_vars + 2) + self.n_vars + self.m) * G1W * 4 + (self.n_vars + 2) * G2W * 4
        csr = sum(b.nbytes for q in (self.qapA, self.qapB) for k, b in q.items() if k != "nrows")
        return pts + csr


class GrothProver:
    """Proves on one context/stream; buffers are sized once for the key."""

    def __init__(self, pk: ProvingKey, ctx=None, stream=None):
        self.pk = pk
        self.ctx = ctx or pk.ctx
        self.stream = stream
        nv, m = pk.n_vars, pk.m
        c = self.ctx
        self.wA = DeviceBuffer(c, (nv + 2) * 32)
        self.wB = DeviceBuffer(c, (nv + 2) * 32)
        self.pA = DeviceBuffer(c, m * 32)
        self.pB = DeviceBuffer(c, m * 32)
        self.H = DeviceBuffer(c, 2 * m * 32)
        self.rs = DeviceBuffer(c, 3 * 32)  # r, s, scratch
        self.outs = DeviceBuffer(c, 8 * G2W * 4)
        self.fin_s = DeviceBuffer(c, 4 * 32)
        self.fin_p = DeviceBuffer(c, 4 * G1W * 4)
        self.times: dict[str, float] = {}

    def _sync(self):
        check(lib().csk_stream_sync(self.ctx.ptr, self.stream), "sync")

    def prove(self, witness: np.ndarray, r: int, s: int, timed: bool = False):
        pk, c, st = self.pk, self.ctx, self.stream
        nv, m = pk.n_vars, pk.m
        w = np.ascontiguousarray(witness, dtype=np.uint32).reshape(-1, 8)
        if w.shape[0] != nv:
            raise ValueError("witness has %d entries, the key %d" % (w.shape[0], nv))
        t = time.perf_counter()
        one = inputs.from_ints([1])
        rs = inputs.from_ints([r % inputs.R, s % inputs.R])
        self.wA.upload(np.concatenate([w, one, rs[0:1]]), stream=st)
        self.wB.upload(np.concatenate([w, one, rs[1:2]]), stream=st)
        self.rs.upload(np.concatenate([rs, np.zeros((1, 8), np.uint32)]), stream=st)
        mark = self._marker(timed)
        mark("upload", t)
        # QAP evaluation and H
        mpoly_eval_device(pk.qapA, self.wA, self.pA, ctx=c, stream=st)
        mpoly_eval_device(pk.qapB, self.wB, self.pB, ctx=c, stream=st)
        mark("qap")
        interpolandmul_device(self.pA, self.pB, pk.log2m, self.H, ctx=c, stream=st)
        # H[2m-1] (always zero) becomes the scalar of the appended delta1: -(r s) mod r
        rsp = self.rs.ptr
        check(lib().csk_field_op(c.ptr, _capi.OP_TOMONT, 1, rsp, None, rsp + 64, 1, st), "tomont")
        check(lib().csk_field_op(c.ptr, _capi.OP_MULM, 1, rsp + 64, rsp + 32, rsp + 64, 1, st), "mulm")
        hlast = self.H.ptr + (2 * m - 1) * 32
        zero = self.H.ptr + (2 * m - 1) * 32  # H[2m-1] == 0 here
        check(lib().csk_field_op(c.ptr, _capi.OP_SUBM, 1, zero, rsp + 64, hlast, 1, st), "subm")
        mark("h")
        o = self.outs.ptr
        oA, oB1, oCp, oC2, oB2, oC = o, o + 64, o + 128, o + 192, o + 256, o + 384
        msm_device(self.wA, pk.A, nv + 2, oA, ctx=c, stream=st)
        mark("msm_a")
        msm_device(self.wB, pk.B2, nv + 2, oB2, g2=True, ctx=c, stream=st)
        mark("msm_b2")
        msm_device(self.wB, pk.B1, nv + 2, oB1, ctx=c, stream=st)
        mark("msm_b1")
        lo = pk.n_public + 1
        msm_device(self.wA.ptr + lo * 32, pk.C.ptr + lo * G1W * 4, nv - lo, oCp, ctx=c, stream=st)
        mark("msm_c")
        msm_device(self.H.ptr + m * 32, pk.hExps, m, oC2, ctx=c, stream=st)
        mark("msm_h")
        # pi_c = 1*pi_c2 + 1*pi_c' + s*pi_a + r*pi_b1
        self.fin_s.upload(np.concatenate([one, one, rs[1:2], rs[0:1]]), stream=st)
        for i, src in enumerate((oC2, oCp, oA, oB1)):
            check(lib().csk_memcpy_d2d(c.ptr, self.fin_p.ptr + i * 64, src, 64, st), "d2d")
        msm_device(self.fin_s, self.fin_p, 4, oC, ctx=c, stream=st)
        out = self.outs.download((8 * G2W,), stream=st)
        mark("final")
        # words: A 0..15, B1 16..31, C' 32..47, C2 48..63, B2 64..95, C 96..111
        return out[0:16].copy(), out[64:96].copy(), out[96:112].copy()

    def _marker(self, timed):
        last = [time.perf_counter()]

        def mark(name, t0=None):
            if not timed:
                return
            self._sync()
            now = time.perf_counter()
            self.times[name] = now - (t0 if t0 is not None else last[0])
            last[0] = now
        return mark
