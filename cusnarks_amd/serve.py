"""Groth16 proof serving with continuous batching on one GPU.

Extends the role of the reference's proof server (startGPServer / startServer,
cusnarks src/python/groth_prover.py:792-891, started by launch_server.sh:70),
which keeps one proving key in host memory, proves one request at a time and
turns requests away while busy.  Here:

* the proving keys of several circuits stay resident in HBM (ProvingKey);
* requests queue in arrival order per lane (circuits below / above a domain
  size) and are admitted while the lane's share of an HBM budget for
  per-request state (witness / QAP / H buffers of a ProofSlot plus the
  request's share of the multiexp workspace) and a cap on proofs in flight
  allow; a request that could never fit is rejected;
* on admission the request's witness is uploaded and its QAP evaluation and
  H polynomial are enqueued at once on its lane's H stream, so they overlap
  the multiexps already on the device;
* each lane keeps up to two multiexp cohorts in flight: every proof of the
  lane whose H is ready joins the next one, as one G1 batch (A, B1, C', C2 of
  all its proofs) followed by one batch of the final combinations, and one
  G2 batch (B2) on a second stream; the second cohort queues behind the
  first on the same streams, so the device does not idle while the host
  completes one cohort and forms the next;
* with use_tables, a circuit whose key carries fixed-base tables
  (ProvingKey.build_tables, before the server is made) runs each of its five
  multiexps as one batched table run over the cohort's proofs of that
  circuit (off by default: slower than the batched plain multiexps);
* every proof is bit-identical to the proof of the same request alone.

`PlainServer` is the baseline: the same requests proved one at a time with a
plain per-request prover (single-call entry points, one stream), with the
keys resident in HBM or, as in the reference's flow, in host memory.
"""
from __future__ import annotations

import collections
import dataclasses
import random
import time

import numpy as np

from . import _capi, inputs
from .batch import MsmBatch
from .device import Event, Stream
from .prover import GrothProver, ProofSlot, ProvingKey

G1W, G2W = 16, 32


@dataclasses.dataclass
class Request:
    rid: int
    circuit: str
    witness: np.ndarray
    r: int
    s: int
    t_arrival: float = 0.0
    t_admit: float | None = None
    t_done: float | None = None
    proof: tuple | None = None
    error: str | None = None
    slot: object = None
    need: int = 0
    lane: int = 0
    ev_h: object = None

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrival


class Admission:
    """Per-lane admission by an HBM budget share and a cap on proofs in flight.

    Each lane owns a fixed share of the budget, so requests of the small lane
    never wait behind a large request that does not fit yet.  A lane accounts
    its slot buffers (busy and idle), its multiexp workspace high-water mark
    and the multiexp need of its proofs in flight (a lane's workspace grows to
    the need of its largest cohort, at most the sum over its proofs in flight).
    The share is soft in one case: a lane with nothing in flight admits a
    request that fits the share alone, whatever its high-water mark.
    """

    def __init__(self, budget_bytes: int, max_inflight: int, shares=(0.25, 0.75)):
        self.budget = [int(budget_bytes * f) for f in shares]
        self.max_inflight = int(max_inflight)
        n = len(shares)
        self.slot_bytes = [0] * n
        self.inflight = [0] * n
        self.need = [0] * n
        self.high = [0] * n

    def projected(self, lane: int, new_slot: int = 0, msm: int = 0) -> int:
        return self.slot_bytes[lane] + new_slot + max(self.high[lane], self.need[lane] + msm)

    def fits(self, lane: int, new_slot: int, msm: int) -> bool:
        if self.inflight[lane] >= self.max_inflight:
            return False
        if self.inflight[lane] == 0:
            # an idle lane takes any request that fits its share alone (the
            # caller rejects the others): its workspace high-water mark and
            # idle slots may exceed the share, and waiting on them would stall
            # the lane for good, since only a completion changes them
            return True
        return self.projected(lane, new_slot, msm) <= self.budget[lane]

    def never_fits(self, lane: int, slot: int, msm: int) -> bool:
        return slot + msm > self.budget[lane]

    def take(self, lane: int, new_slot: int, msm: int) -> None:
        self.slot_bytes[lane] += new_slot
        self.need[lane] += msm
        self.inflight[lane] += 1

    def release(self, lane: int, msm: int) -> None:
        self.need[lane] -= msm
        self.inflight[lane] -= 1

    def cohort(self, lane: int, workspace: int) -> None:
        self.high[lane] = max(self.high[lane], workspace)

    def free_slot(self, lane: int, nbytes: int) -> None:
        self.slot_bytes[lane] -= nbytes


class _Circuit:
    """A registered circuit: its resident key, lane and per-request estimates."""

    def __init__(self, name, pk: ProvingKey, lane: int, ctx, use_tables: bool = False):
        self.name, self.pk, self.lane = name, pk, lane
        self.slot_bytes = ProofSlot.bytes_for(pk.n_vars, pk.m)
        nv, m, lo = pk.n_vars, pk.m, pk.n_public + 1
        self.g1_sizes = [nv + 2, nv + 2, nv - lo, m]
        self.tables = (getattr(pk, "tables", None) or {}) if use_tables else {}
        if self.tables:  # a table run's workspace grows about linearly with its scalar vectors
            self.msm_bytes = (max(self.tables[k].info()["workspace_bytes"] for k in ProofSlot.G1_TABLES)
                              + self.tables["B2"].info()["workspace_bytes"])
        else:
            self.msm_bytes = (MsmBatch.workspace_bytes(self.g1_sizes, ctx=ctx)
                              + MsmBatch.workspace_bytes([nv + 2], g2=True, ctx=ctx))
        self.idle: list[ProofSlot] = []


class _Cohort:
    """The proofs of one launched cohort and the events that mark its end."""

    def __init__(self, reqs, ev_g1, ev_g2):
        self.reqs, self.ev_g1, self.ev_g2 = reqs, ev_g1, ev_g2

    def done(self) -> bool:
        return self.ev_g1.done() and self.ev_g2.done()


class _Lane:
    H_STREAMS = 4  # QAP + H of admitted proofs, dealt round-robin: a small proof's NTTs do not fill the GPU
    COHORTS = 2    # cohorts in flight: the next one queues behind the running one on the same streams
    T_STREAMS = 4  # table runs of a cohort, dealt round-robin: the runs of small proofs fill the GPU together

    def __init__(self, device: int, tables: bool = False):
        # separate contexts: each owns the workspace one stream uses at a time
        self.ctx_hs = [_capi.Context(device) for _ in range(self.H_STREAMS)]
        self.ctx_g1 = _capi.Context(device)
        self.ctx_g2 = _capi.Context(device)
        from .tune import apply
        for c in (*self.ctx_hs, self.ctx_g1, self.ctx_g2):
            apply(c)
        self.st_hs = [Stream(c) for c in self.ctx_hs]
        self.next_h = 0
        self.ctx_h = self.ctx_hs[0]  # owns the proof slots' buffers
        self.st_g1, self.st_g2 = Stream(self.ctx_g1), Stream(self.ctx_g2)
        # table runs: a context (workspace) and stream each, joined into the
        # G1 and G2 streams before the cohort's final combinations and its end
        self.ctx_ts = [_capi.Context(device) for _ in range(self.T_STREAMS if tables else 0)]
        self.st_ts = [Stream(c) for c in self.ctx_ts]
        self.ev_ts = [Event(c) for c in self.ctx_ts]
        self.pending: collections.deque[Request] = collections.deque()  # H enqueued, not yet in a cohort
        self.inflight: collections.deque[_Cohort] = collections.deque()  # launched, oldest first
        # one event pair per cohort position: a pair is reused only after its cohort completed
        self.events = [(Event(self.ctx_g1), Event(self.ctx_g2)) for _ in range(self.COHORTS)]
        self.batches: collections.OrderedDict = collections.OrderedDict()
        self.cohorts = 0
        self.cohort_sizes: list[int] = []

    def batch(self, sizes, g2: bool, pos: int) -> MsmBatch:
        # keyed by cohort position too: a batch rewrites its descriptor tables
        # only after its previous run ended, and the cohort in flight at the
        # other position may be that run
        key = (g2, tuple(sizes), pos)
        b = self.batches.get(key)
        if b is None:
            b = MsmBatch(sizes, g2=g2, ctx=self.ctx_g2 if g2 else self.ctx_g1, shared_workspace=True)
            b.uses = 0
            self.batches[key] = b
            while len(self.batches) > 32:  # each batch owns its streams and tables
                _, old = self.batches.popitem(last=False)
                old.close()
        else:
            self.batches.move_to_end(key)
        return b


class ProofServer:
    """Continuous-batching Groth16 server over resident proving keys.

    keys: {name: ProvingKey}; budget_bytes: HBM for per-request state, of
    which small_share is reserved for the small lane; large_log2m: circuits
    with log2(m) >= this go to the large lane; use_tables: run the multiexps
    of keys that carry fixed-base tables from them (batched table runs).  Off
    by default: measured on bursts of one circuit (2^14 .. 2^18 constraints)
    it serves a half to two thirds of the proofs/s of the batched plain
    multiexps, whose capped window widths keep a cohort's bucket sets small,
    while a table's wide windows give every proof Wt * 2^(c-1) buckets.
    """

    def __init__(self, keys: dict, budget_bytes: int, max_inflight: int = 64, large_log2m: int = 20,
                 device: int = 0, use_graph: bool = True, small_share: float = 0.25, use_tables: bool = False):
        self.lanes = [_Lane(device, use_tables), _Lane(device, use_tables)]
        ctx = self.lanes[0].ctx_g1
        self.circuits = {name: _Circuit(name, pk, int(pk.log2m >= large_log2m), ctx, use_tables)
                         for name, pk in keys.items()}
        self.adm = Admission(budget_bytes, max_inflight, shares=(small_share, 1.0 - small_share))
        self.queues = [collections.deque(), collections.deque()]
        self.use_graph = use_graph
        self.done: list[Request] = []
        self.rejected: list[Request] = []

    # ---------------------------------------------------------------- intake
    def submit(self, req: Request) -> None:
        if req.circuit not in self.circuits:
            req.error = "unknown circuit %r" % req.circuit
            self.rejected.append(req)
            return
        self.queues[self.circuits[req.circuit].lane].append(req)

    def _admit(self, now: float) -> bool:
        progressed = False
        for queue in self.queues:
            progressed |= self._admit_lane(queue, now)
        return progressed

    def _admit_lane(self, queue, now: float) -> bool:
        admitted: list[Request] = []
        while queue:
            req = queue[0]
            cc = self.circuits[req.circuit]
            new_slot = 0 if cc.idle else cc.slot_bytes
            if self.adm.never_fits(cc.lane, cc.slot_bytes, cc.msm_bytes):
                queue.popleft()
                req.error = "request needs %d bytes, lane budget %d" % (cc.slot_bytes + cc.msm_bytes,
                                                                       self.adm.budget[cc.lane])
                self.rejected.append(req)
                continue
            if not self.adm.fits(cc.lane, new_slot, cc.msm_bytes):
                if new_slot and self._drop_idle_slots(cc):
                    continue
                break
            queue.popleft()
            self.adm.take(cc.lane, new_slot, cc.msm_bytes)
            lane = self.lanes[cc.lane]
            req.slot = cc.idle.pop() if cc.idle else ProofSlot(cc.pk, lane.ctx_h)
            req.need, req.lane, req.t_admit = cc.msm_bytes, cc.lane, now
            # witness upload on the null stream (synchronous); QAP + H below
            req.slot.load(req.witness, req.r, req.s, stream=None)
            admitted.append(req)
        self._enqueue_h(admitted)
        return bool(admitted)

    def _enqueue_h(self, admitted) -> None:
        """QAP evaluations and H steps of the requests just admitted: the
        proofs of one circuit go to one H stream together and their H steps
        run batched (each NTT pass one launch over up to 32 proofs); circuits
        are dealt round-robin over the lane's H streams."""
        groups: dict = {}
        for req in admitted:
            groups.setdefault(req.circuit, []).append(req)
        for name, reqs in groups.items():
            lane = self.lanes[self.circuits[name].lane]
            k = lane.next_h
            lane.next_h = (k + 1) % len(lane.st_hs)
            ctx, st = lane.ctx_hs[k], lane.st_hs[k]
            for req in reqs:
                req.slot.stage_qap(ctx, st.ptr)
            ProofSlot.stage_h_batch([req.slot for req in reqs], ctx, st.ptr)
            for req in reqs:
                req.ev_h = Event(ctx)
                req.ev_h.record(st)
                lane.pending.append(req)

    def _drop_idle_slots(self, keep: _Circuit) -> bool:
        """Free the idle slots of the other circuits of keep's lane."""
        freed = False
        for cc in self.circuits.values():
            if cc is keep or cc.lane != keep.lane:
                continue
            while cc.idle:
                cc.idle.pop().close()
                self.adm.free_slot(cc.lane, cc.slot_bytes)
                freed = True
        return freed

    # --------------------------------------------------------------- cohorts
    def _launch(self, li: int) -> bool:
        lane = self.lanes[li]
        if len(lane.inflight) >= lane.COHORTS or not lane.pending:
            return False
        if lane.inflight and not all(req.ev_h.done() for req in lane.pending):
            # behind a running cohort, wait until every admitted proof is ready
            # rather than split them over small cohorts
            return False
        ready = []
        while lane.pending and lane.pending[0].ev_h.done():
            ready.append(lane.pending.popleft())
        if not ready:
            return False
        pos = lane.cohorts % lane.COHORTS  # the other position may be in flight; this one is free
        # proofs of circuits whose keys carry fixed-base tables run their five
        # multiexps from them, one table run per table over the circuit's
        # proofs; the others join the cohort's plain G1 / G2 batches
        tabbed: dict = {}
        plain = []
        for req in ready:
            if self.circuits[req.circuit].tables:
                tabbed.setdefault(req.circuit, []).append(req)
            else:
                plain.append(req)
        ws = 0
        nt = 0  # table runs dealt so far
        for name, reqs in tabbed.items():
            tabs = self.circuits[name].tables
            for c0 in range(0, len(reqs), 64):
                chunk = reqs[c0:c0 + 64]
                mem = [req.slot.g1_members() for req in chunk]
                g2m = [req.slot.g2_member() for req in chunk]
                runs_t = [(tabs[key], [m[k][0] for m in mem], [m[k][3] for m in mem])
                          for k, key in enumerate(ProofSlot.G1_TABLES)]
                runs_t.append((tabs["B2"], [m[0] for m in g2m], [m[3] for m in g2m]))
                for tab, sc, outs in runs_t:
                    k = nt % lane.T_STREAMS
                    tab.run_batch(sc, outs, ctx=lane.ctx_ts[k], stream=lane.st_ts[k].ptr)
                    nt += 1
                ws = max(ws, len(chunk) * self.circuits[name].msm_bytes)
        ws *= min(nt, lane.T_STREAMS)  # one workspace per table stream in use
        for k in range(min(nt, lane.T_STREAMS)):  # the cohort's G1 and G2 streams wait for its table runs
            lane.ev_ts[k].record(lane.st_ts[k])
            lane.st_g1.wait(lane.ev_ts[k])
            lane.st_g2.wait(lane.ev_ts[k])
        runs = []
        if plain:
            g1 = [m for req in plain for m in req.slot.g1_members()]
            g2 = [req.slot.g2_member() for req in plain]
            b1 = lane.batch([m[2] for m in g1], False, pos)
            b2 = lane.batch([m[2] for m in g2], True, pos)
            runs += [(b1, g1, lane.st_g1), (b2, g2, lane.st_g2)]
            ws += b1.info()["bytes"] + b2.info()["bytes"]
        fin = [req.slot.final_member() for req in ready]
        runs.append((lane.batch([4] * len(ready), False, pos), fin, lane.st_g1))
        self.adm.cohort(li, ws)
        for b, mem, st in runs:
            b.run([m[0] for m in mem], [m[1] for m in mem], [m[3] for m in mem],
                  use_graph=self.use_graph and b.uses > 0, stream=st.ptr)
            b.uses += 1
        ev_g1, ev_g2 = lane.events[pos]
        ev_g1.record(lane.st_g1)
        ev_g2.record(lane.st_g2)
        lane.inflight.append(_Cohort(ready, ev_g1, ev_g2))
        lane.cohorts += 1
        lane.cohort_sizes.append(len(ready))
        return True

    def _complete(self, li: int, now_fn) -> bool:
        lane = self.lanes[li]
        if not lane.inflight or not lane.inflight[0].done():
            return False
        cohort = lane.inflight.popleft()  # cohorts end in launch order (same streams)
        t = now_fn()
        for req in cohort.reqs:
            req.proof = req.slot.result()
            req.t_done = t
            cc = self.circuits[req.circuit]
            cc.idle.append(req.slot)
            req.slot = None
            req.ev_h.close()
            req.ev_h = None
            self.adm.release(req.lane, req.need)
            self.done.append(req)
        return True

    def step(self, now_fn) -> bool:
        """One pass of the loop: completions, admissions, cohort launches."""
        progressed = False
        for li in range(len(self.lanes)):
            progressed |= self._complete(li, now_fn)
        progressed |= self._admit(now_fn())
        for li in range(len(self.lanes)):
            progressed |= self._launch(li)
        return progressed

    def busy(self) -> bool:
        return any(self.queues) or any(l.inflight or l.pending for l in self.lanes)

    def run(self, requests, poll_s: float = 2e-4) -> list[Request]:
        """Serve requests whose t_arrival are offsets (s) from now; returns them
        once all are done or rejected."""
        reqs = sorted(requests, key=lambda q: q.t_arrival)
        t0 = time.perf_counter()
        now_fn = lambda: time.perf_counter() - t0  # noqa: E731
        i = 0
        last = now_fn()
        while i < len(reqs) or self.busy():
            now = now_fn()
            while i < len(reqs) and reqs[i].t_arrival <= now:
                self.submit(reqs[i])
                i += 1
            if self.step(now_fn):
                last = now_fn()
            else:
                idle = not any(l.inflight or l.pending for l in self.lanes)
                if idle and any(self.queues) and now_fn() - last > 30.0:
                    raise RuntimeError("proof server stalled: %d requests queued, none admitted for 30 s"
                                       % sum(len(q) for q in self.queues))
                wait = reqs[i].t_arrival - now_fn() if i < len(reqs) else poll_s
                time.sleep(min(max(wait, 0.0), poll_s))
        return reqs

    def stats(self) -> dict:
        return {"cohorts": [l.cohorts for l in self.lanes],
                "mean_cohort": [float(np.mean(l.cohort_sizes)) if l.cohort_sizes else 0.0 for l in self.lanes],
                "rejected": len(self.rejected)}


class HostKeyProver:
    """Plain per-request prover with the proving-key points in host memory,
    copied to the device for every multiexp call -- the reference's flow,
    whose GPU multiexp copies its input vector (scalars and points) in on each
    kernelLaunch (cusnarks src/cuda/cusnarks_kernel.cu:402-573) from a key
    held in host memory (src/python/groth_prover.py:1548-1688).  QAP and H
    run on the device as in GrothProver."""

    def __init__(self, pk: ProvingKey, ctx):
        from .device import DeviceBuffer
        self.pk, self.ctx = pk, ctx
        nv, m = pk.n_vars, pk.m
        self.host = {"A": pk.A.download((nv + 2, G1W)), "B1": pk.B1.download((nv + 2, G1W)),
                     "C": pk.C.download((nv, G1W)), "hExps": pk.hExps.download((m, G1W)),
                     "B2": pk.B2.download((nv + 2, G2W))}
        self.stage1 = DeviceBuffer(ctx, max(nv + 2, m) * G1W * 4)
        self.stage2 = DeviceBuffer(ctx, (nv + 2) * G2W * 4)
        self.slot = ProofSlot(pk, ctx)

    def _msm(self, scal, pts_host, n, out, g2):
        from ._capi import P, check, lib
        from .msm import msm_device
        st = self.stage2 if g2 else self.stage1
        arr = np.ascontiguousarray(pts_host[:n])
        check(lib().csk_memcpy_h2d(self.ctx.ptr, st.ptr, arr.ctypes.data_as(P), arr.nbytes, None), "h2d")
        msm_device(scal, st, n, out, g2=g2, ctx=self.ctx)

    def prove(self, witness, r, s):
        sl, pk, c = self.slot, self.pk, self.ctx
        lo = pk.n_public + 1
        sl.load(witness, r, s)
        sl.stage_qap(c)
        sl.stage_h(c)
        mem = sl.g1_members()
        for (ds, _, n, do), key, first in zip(mem, ("A", "B1", "C", "hExps"), (0, 0, lo, 0)):
            self._msm(ds, self.host[key][first:], n, do, False)
        ds, _, n, do = sl.g2_member()
        self._msm(ds, self.host["B2"], n, do, True)
        ds, dp, n, do = sl.final_member()
        from .msm import msm_device
        msm_device(ds, dp, n, do, ctx=c)
        return sl.result()


class PlainServer:
    """Baseline: one request at a time through a plain per-request prover,
    with the keys resident in HBM (GrothProver) or, with host_keys, in host
    memory and copied in per multiexp call (HostKeyProver)."""

    def __init__(self, keys: dict, device: int = 0, host_keys: bool = False):
        self.ctx = _capi.Context(device)
        from .tune import apply
        apply(self.ctx)
        mk = (lambda pk: HostKeyProver(pk, self.ctx)) if host_keys else (lambda pk: GrothProver(pk, ctx=self.ctx))
        self.provers = {name: mk(pk) for name, pk in keys.items()}

    def run(self, requests) -> list[Request]:
        reqs = sorted(requests, key=lambda q: q.t_arrival)
        t0 = time.perf_counter()
        for req in reqs:
            now = time.perf_counter() - t0
            if req.t_arrival > now:
                time.sleep(req.t_arrival - now)
            req.t_admit = time.perf_counter() - t0
            req.proof = self.provers[req.circuit].prove(req.witness, req.r, req.s)
            req.t_done = time.perf_counter() - t0
        return reqs


# ------------------------------------------------------------------ load
def poisson_requests(rate: float, count: int, witnesses: dict, seed: int = 0, weights=None) -> list[Request]:
    """count requests with exponential inter-arrival times (mean 1/rate s),
    circuits drawn from witnesses' keys (uniform, or by weights); r and s are
    the stdlib draw randint(1, r-1) twice after seeding with the request id."""
    rng = random.Random(seed)
    names = sorted(witnesses)
    t = 0.0
    out = []
    for rid in range(count):
        t += rng.expovariate(rate) if rate > 0 else 0.0
        name = rng.choices(names, weights=[weights[n] for n in names] if weights else None)[0]
        g = random.Random(rid)
        r, s = g.randint(1, inputs.R - 1), g.randint(1, inputs.R - 1)
        out.append(Request(rid, name, witnesses[name], r, s, t_arrival=t))
    return out


def summarize(reqs) -> dict:
    done = [q for q in reqs if q.t_done is not None]
    if not done:
        return {"proofs": 0}
    lat = np.array([q.latency for q in done]) * 1e3
    t_first = min(q.t_arrival for q in done)
    t_last = max(q.t_done for q in done)
    return {"proofs": len(done), "proofs_per_s": len(done) / (t_last - t_first) if t_last > t_first else None,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max())}
