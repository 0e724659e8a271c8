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
from ._capi import P, check, lib
from .device import DeviceBuffer
from .msm import fixed_base, msm_device
from .poly import interpolandmul_batch_device, interpolandmul_device, mpoly_eval_device, upload_csr
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

    @classmethod
    def from_device(cls, ctx, n_vars: int, n_public: int, log2m: int, A, B1, B2, C, hExps, qapA: dict,
                    qapB: dict) -> "ProvingKey":
        """A key from device buffers already in the prover's layout (A, B1, B2:
        nVars + 2 points; C: nVars; hExps: m - 1 points + delta1; QAP A / B as
        uploaded CSR), e.g. read from a proving-key file (keyio.load_proving_key)."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.n_vars, self.n_public = int(n_vars), int(n_public)
        self.log2m = int(log2m)
        self.m = 1 << self.log2m
        self.A, self.B1, self.B2, self.C, self.hExps = A, B1, B2, C, hExps
        self.qapA, self.qapB = qapA, qapB
        return self

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
        return pts + csr + self.table_bytes()

    # --- fixed-base tables of the key's point arrays
    def build_tables(self, window_bits: int = 0, g2_window_bits: int = 0, stream=None) -> int:
        """Precompute fixed-base tables (csk_msm_table_*) for the five
        multiexps of a proof: A, B1, B2 (nVars + 2 points), the C slice past
        the public wires and hExps (m points).  Afterwards GrothProver and
        OverlappedProver run those multiexps from the tables: no window
        combination, one bucket reduction, no doublings.  The reference keeps
        such tables for its CPU multiexp (groth_setup.py:802-945).

        The provers run the four G1 tables as one pipeline in their own
        context's workspace (run_g1_tables), so they share one window width:
        window_bits, or else the library's default for the largest of them.
        The G2 table runs in a context of its own (one proof at a time per
        key), so it can run beside them on another stream.  Returns the table
        bytes held in HBM."""
        from .msm import MsmTable
        self.close_tables()
        nv, lo = self.n_vars, self.n_public + 1
        if nv <= lo:
            raise ValueError("the key has no private wires: no C multiexp to tabulate")
        from .tune import apply
        self._g2ctx = _capi.Context(self.ctx.device)
        apply(self._g2ctx)
        g1 = {"A": (self.A, nv + 2), "B1": (self.B1, nv + 2), "C": (self.C.ptr + lo * G1W * 4, nv - lo),
              "H": (self.hExps, self.m)}
        first = max(ProofSlot.G1_TABLES, key=lambda k: g1[k][1])
        tabs = {first: MsmTable(*g1[first], window_bits=window_bits, ctx=self.ctx, stream=stream)}
        wb = tabs[first].info()["window_bits"]
        for k in ProofSlot.G1_TABLES:
            if k not in tabs:
                tabs[k] = MsmTable(*g1[k], window_bits=wb, ctx=self.ctx, stream=stream)
        self.tables = {k: tabs[k] for k in ProofSlot.G1_TABLES}
        self.tables["B2"] = MsmTable(self.B2, nv + 2, g2=True, window_bits=g2_window_bits, ctx=self._g2ctx,
                                     stream=stream)
        check(lib().csk_stream_sync(self.ctx.ptr, stream), "sync")
        return self.table_bytes()

    def table_bytes(self) -> int:
        return sum(t.info()["table_bytes"] for t in getattr(self, "tables", {}).values())

    def run_g1_tables(self, members, ctx=None, stream=None) -> None:
        """The four G1 multiexps of a proof (ProofSlot.g1_members order) from
        this key's tables: one pipeline for the tables that share a window
        width (all four as build_tables makes them), one per width otherwise."""
        from .msm import MsmTable
        tabs = [self.tables[k] for k in ProofSlot.G1_TABLES]
        by_c: dict = {}
        for tab, (ds, _, _, do) in zip(tabs, members):
            by_c.setdefault(tab.info()["window_bits"], []).append((tab, ds, do))
        for runs in by_c.values():
            MsmTable.run_multi([r[0] for r in runs], [r[1] for r in runs], [r[2] for r in runs],
                               ctx=ctx or self.ctx, stream=stream)

    def close_tables(self) -> None:
        for t in getattr(self, "tables", {}).values():
            t.close()
        self.tables = {}


class ProofSlot:
    """Device buffers of one proof in flight for a key, and its stages.

    outs holds the affine results: G1 A | B1 | C' | C2 at words 0..63 (four
    consecutive points, which the final combination reads as its points), G2
    B at 64..95 and G1 C at 96..111.
    """

    def __init__(self, pk: ProvingKey, ctx=None):
        self.pk = pk
        c = self.ctx = ctx or pk.ctx
        nv, m = pk.n_vars, pk.m
        self.wA = DeviceBuffer(c, (nv + 2) * 32)
        self.wB = DeviceBuffer(c, (nv + 2) * 32)
        self.pA = DeviceBuffer(c, m * 32)
        self.pB = DeviceBuffer(c, m * 32)
        self.H = DeviceBuffer(c, 2 * m * 32)
        self.rs = DeviceBuffer(c, 3 * 32)  # r, s, scratch
        self.fin_s = DeviceBuffer(c, 4 * 32)
        self.outs = DeviceBuffer(c, 8 * G2W * 4)

    @staticmethod
    def bytes_for(n_vars: int, m: int) -> int:
        return 2 * (n_vars + 2) * 32 + 4 * m * 32 + 3 * 32 + 4 * 32 + 8 * G2W * 4

    @property
    def nbytes(self) -> int:
        return self.bytes_for(self.pk.n_vars, self.pk.m)

    def load(self, witness: np.ndarray, r: int, s: int, stream=None) -> None:
        """Upload w | 1 | r and w | 1 | s, (r, s) and the final scalars (s, r, 1, 1).

        The witness crosses PCIe once; wB's copy of it is made on the device.
        Returns after the host buffers may be reused (the copies are done)."""
        nv = self.pk.n_vars
        w = np.ascontiguousarray(witness, dtype=np.uint32).reshape(-1, 8)
        if w.shape[0] != nv:
            raise ValueError("witness has %d entries, the key %d" % (w.shape[0], nv))
        one = inputs.from_ints([1])
        rs = inputs.from_ints([r % inputs.R, s % inputs.R])
        L, c = lib(), self.ctx.ptr
        tail_a = np.concatenate([one, rs[0:1]])
        tail_b = np.concatenate([one, rs[1:2]])
        rs3 = np.concatenate([rs, np.zeros((1, 8), np.uint32)])
        fin = np.concatenate([rs[1:2], rs[0:1], one, one])
        for dst, arr in ((self.wA.ptr, w), (self.wA.ptr + nv * 32, tail_a), (self.wB.ptr + nv * 32, tail_b),
                         (self.rs.ptr, rs3), (self.fin_s.ptr, fin)):
            check(L.csk_memcpy_h2d(c, dst, arr.ctypes.data_as(P), arr.nbytes, stream), "csk_memcpy_h2d")
        check(L.csk_memcpy_d2d(c, self.wB.ptr, self.wA.ptr, nv * 32, stream), "csk_memcpy_d2d")
        check(L.csk_stream_sync(c, stream), "csk_stream_sync")

    def stage_qap(self, ctx=None, stream=None) -> None:
        pk, c = self.pk, ctx or self.ctx
        mpoly_eval_device(pk.qapA, self.wA, self.pA, ctx=c, stream=stream)
        mpoly_eval_device(pk.qapB, self.wB, self.pB, ctx=c, stream=stream)

    def stage_h(self, ctx=None, stream=None) -> None:
        """polH = interpolandmul(pA, pB); H[2m-1] := -(r s), the scalar of the appended delta1."""
        pk, c = self.pk, ctx or self.ctx
        interpolandmul_device(self.pA, self.pB, pk.log2m, self.H, ctx=c, stream=stream)
        self._h_tail(c, stream)

    @staticmethod
    def stage_h_batch(slots, ctx=None, stream=None) -> None:
        """stage_h of several slots of one key, their transforms batched
        (every NTT pass one launch for up to 32 proofs)."""
        if not slots:
            return
        pk = slots[0].pk
        if any(sl.pk.m != pk.m for sl in slots):
            raise ValueError("batched H steps need one domain size")
        c = ctx or slots[0].ctx
        interpolandmul_batch_device([sl.pA for sl in slots], [sl.pB for sl in slots], pk.log2m,
                                    [sl.H for sl in slots], ctx=c, stream=stream)
        for sl in slots:
            sl._h_tail(c, stream)

    def _h_tail(self, c, stream) -> None:
        m = self.pk.m
        rsp = self.rs.ptr
        check(lib().csk_field_op(c.ptr, _capi.OP_TOMONT, 1, rsp, None, rsp + 64, 1, stream), "tomont")
        check(lib().csk_field_op(c.ptr, _capi.OP_MULM, 1, rsp + 64, rsp + 32, rsp + 64, 1, stream), "mulm")
        hlast = self.H.ptr + (2 * m - 1) * 32  # polH has degree <= 2m-2: this slot is 0
        check(lib().csk_field_op(c.ptr, _capi.OP_SUBM, 1, hlast, rsp + 64, hlast, 1, stream), "subm")

    G1_TABLES = ("A", "B1", "C", "H")

    def g1_members(self):
        """(scalars, points, n, out) of the G1 multiexps A, B1, C', C2 (the
        key's tables for them are G1_TABLES, in this order)."""
        pk, o = self.pk, self.outs.ptr
        nv, m, lo = pk.n_vars, pk.m, pk.n_public + 1
        return [(self.wA.ptr, pk.A.ptr, nv + 2, o),
                (self.wB.ptr, pk.B1.ptr, nv + 2, o + 64),
                (self.wA.ptr + lo * 32, pk.C.ptr + lo * G1W * 4, nv - lo, o + 128),
                (self.H.ptr + m * 32, pk.hExps.ptr, m, o + 192)]

    def g2_member(self):
        pk = self.pk
        return (self.wB.ptr, pk.B2.ptr, pk.n_vars + 2, self.outs.ptr + 256)

    def final_member(self):
        """pi_c = s*pi_a + r*pi_b1 + 1*pi_c' + 1*pi_c2 over the four G1 results."""
        return (self.fin_s.ptr, self.outs.ptr, 4, self.outs.ptr + 384)

    def result(self, stream=None):
        out = self.outs.download((8 * G2W,), stream=stream)
        return out[0:16].copy(), out[64:96].copy(), out[96:112].copy()

    def close(self) -> None:
        for b in (self.wA, self.wB, self.pA, self.pB, self.H, self.rs, self.fin_s, self.outs):
            b.close()


class GrothProver:
    """Plain per-request prover: every stage through the single-call entry
    points, one after the other on one context and stream.  With the key's
    fixed-base tables built (ProvingKey.build_tables) the five multiexps run
    from them; the proof is the same."""

    def __init__(self, pk: ProvingKey, ctx=None, stream=None):
        self.pk = pk
        self.ctx = ctx or pk.ctx
        self.stream = stream
        self.slot = ProofSlot(pk, self.ctx)
        self.times: dict[str, float] = {}

    def _sync(self):
        check(lib().csk_stream_sync(self.ctx.ptr, self.stream), "sync")

    def prove(self, witness: np.ndarray, r: int, s: int, timed: bool = False):
        c, st, sl = self.ctx, self.stream, self.slot
        t = time.perf_counter()
        sl.load(witness, r, s, stream=st)
        mark = self._marker(timed)
        mark("upload", t)
        sl.stage_qap(c, st)
        mark("qap")
        sl.stage_h(c, st)
        mark("h")
        names = ("msm_a", "msm_b1", "msm_c", "msm_h")
        tabs = getattr(self.pk, "tables", None) or {}
        if tabs:
            self.pk.run_g1_tables(sl.g1_members(), ctx=c, stream=st)
            mark("msm_g1")
        else:
            for name, (ds, dp, n, do) in zip(names, sl.g1_members()):
                msm_device(ds, dp, n, do, ctx=c, stream=st)
                mark(name)
        ds, dp, n, do = sl.g2_member()
        if tabs:
            tabs["B2"].run(ds, do, stream=st)
        else:
            msm_device(ds, dp, n, do, g2=True, ctx=c, stream=st)
        mark("msm_b2")
        ds, dp, n, do = sl.final_member()
        msm_device(ds, dp, n, do, ctx=c, stream=st)
        res = sl.result(stream=st)
        mark("final")
        return res

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


class OverlappedProver:
    """One proof at a time with its stages overlapped: the G2 multiexp needs
    only the uploaded witness, so it runs on a second stream (and context)
    from the upload on, beside QAP evaluation, H and the G1 work; the four
    G1 multiexps run as one batch (one sort, one window combination per
    stage, graph replay), or from the key's fixed-base tables when it has
    them.  Results are those of GrothProver."""

    def __init__(self, pk: ProvingKey, device: int = 0):
        from .batch import MsmBatch
        from .device import Event, Stream
        from .tune import apply
        self.pk = pk
        self.ctx, self.ctx2 = _capi.Context(device), _capi.Context(device)
        for c in (self.ctx, self.ctx2):
            apply(c)
        self.st, self.st2 = Stream(self.ctx), Stream(self.ctx2)
        self.g2_done = Event(self.ctx2)
        self.slot = ProofSlot(pk, self.ctx)
        self.g1 = MsmBatch([m[2] for m in self.slot.g1_members()], ctx=self.ctx)

    def prove(self, witness: np.ndarray, r: int, s: int):
        sl, c = self.slot, self.ctx
        tabs = getattr(self.pk, "tables", None) or {}
        sl.load(witness, r, s, stream=None)  # returns when the copies are done
        ds, dp, n, do = sl.g2_member()
        if tabs:
            tabs["B2"].run(ds, do, stream=self.st2.ptr)
        else:
            msm_device(ds, dp, n, do, g2=True, ctx=self.ctx2, stream=self.st2.ptr)
        self.g2_done.record(self.st2)
        sl.stage_qap(c, self.st.ptr)
        sl.stage_h(c, self.st.ptr)
        mem = sl.g1_members()
        if tabs:
            self.pk.run_g1_tables(mem, ctx=c, stream=self.st.ptr)
        else:
            self.g1.run([m[0] for m in mem], [m[1] for m in mem], [m[3] for m in mem], use_graph=True,
                        stream=self.st.ptr)
        ds, dp, n, do = sl.final_member()
        msm_device(ds, dp, n, do, ctx=c, stream=self.st.ptr)
        self.st.wait(self.g2_done)
        self.st.sync()
        return sl.result()

