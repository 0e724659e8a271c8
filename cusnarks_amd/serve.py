"""Groth16 proof serving with continuous batching on one GPU.

Extends the role of the reference's proof server (startGPServer / startServer,
cusnarks src/python/groth_prover.py:792-891, started by launch_server.sh:70),
which keeps one proving key in host memory, proves one request at a time and
turns requests away while busy.  Here:

* the proving keys of several circuits stay resident in HBM (ProvingKey);
* requests queue in arrival order and are admitted while an HBM budget for
  per-request state (witness / QAP / H buffers of a ProofSlot plus the
  request's share of the multiexp workspace) and a cap on proofs in flight
  allow; a request that could never fit is rejected;
* on admission the request's witness is uploaded and its QAP evaluation and
  H polynomial are enqueued at once on its lane's H stream, so they overlap
  the multiexps already on the device;
* each lane (circuits below / above a domain size) keeps one multiexp cohort
  in flight: every proof of the lane whose H is ready joins it, as one G1
  batch (A, B1, C', C2 of all its proofs) followed by one batch of the final
  combinations, and one G2 batch (B2) on a second stream; a lane is picked
  up again as soon as its cohort completes;
* every proof is bit-identical to the proof of the same request alone.

`PlainServer` is the baseline: the same requests proved one at a time with
the plain per-request prover (single-call entry points, one stream).
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
    """Arrival-order admission by an HBM budget and a cap on proofs in flight.

    Accounts three kinds of device memory: slot buffers (busy and idle ones),
    each lane's multiexp workspace high-water mark, and the multiexp need of
    the proofs in flight per lane (a lane's workspace grows to the need of
    its largest cohort, which is at most the sum over its proofs in flight).
    """

    def __init__(self, budget_bytes: int, max_inflight: int, lanes: int = 2):
        self.budget = int(budget_bytes)
        self.max_inflight = int(max_inflight)
        self.slot_bytes = 0
        self.inflight = 0
        self.lane_need = [0] * lanes
        self.lane_high = [0] * lanes

    def projected(self, new_slot: int = 0, lane: int = 0, msm: int = 0) -> int:
        need = list(self.lane_need)
        need[lane] += msm
        return self.slot_bytes + new_slot + sum(max(h, n) for h, n in zip(self.lane_high, need))

    def fits(self, new_slot: int, lane: int, msm: int) -> bool:
        return self.inflight < self.max_inflight and self.projected(new_slot, lane, msm) <= self.budget

    def never_fits(self, slot: int, msm: int) -> bool:
        return slot + msm > self.budget

    def take(self, new_slot: int, lane: int, msm: int) -> None:
        self.slot_bytes += new_slot
        self.lane_need[lane] += msm
        self.inflight += 1

    def release(self, lane: int, msm: int) -> None:
        self.lane_need[lane] -= msm
        self.inflight -= 1

    def cohort(self, lane: int, workspace: int) -> None:
        self.lane_high[lane] = max(self.lane_high[lane], workspace)

    def free_slot(self, nbytes: int) -> None:
        self.slot_bytes -= nbytes


class _Circuit:
    """A registered circuit: its resident key, lane and per-request estimates."""

    def __init__(self, name, pk: ProvingKey, lane: int, ctx):
        self.name, self.pk, self.lane = name, pk, lane
        self.slot_bytes = ProofSlot.bytes_for(pk.n_vars, pk.m)
        nv, m, lo = pk.n_vars, pk.m, pk.n_public + 1
        self.g1_sizes = [nv + 2, nv + 2, nv - lo, m]
        self.msm_bytes = (MsmBatch.workspace_bytes(self.g1_sizes, ctx=ctx)
                          + MsmBatch.workspace_bytes([nv + 2], g2=True, ctx=ctx))
        self.idle: list[ProofSlot] = []


class _Lane:
    def __init__(self, device: int):
        # separate contexts: each owns the workspace one stream uses at a time
        self.ctx_h = _capi.Context(device)
        self.ctx_g1 = _capi.Context(device)
        self.ctx_g2 = _capi.Context(device)
        from .tune import apply
        for c in (self.ctx_h, self.ctx_g1, self.ctx_g2):
            apply(c)
        self.st_h, self.st_g1, self.st_g2 = Stream(self.ctx_h), Stream(self.ctx_g1), Stream(self.ctx_g2)
        self.pending: collections.deque[Request] = collections.deque()  # H enqueued, not yet in a cohort
        self.cohort: list[Request] | None = None
        self.ev_g1 = Event(self.ctx_g1)
        self.ev_g2 = Event(self.ctx_g2)
        self.batches: collections.OrderedDict = collections.OrderedDict()
        self.cohorts = 0
        self.cohort_sizes: list[int] = []

    def batch(self, sizes, g2: bool) -> MsmBatch:
        key = (g2, tuple(sizes))
        b = self.batches.get(key)
        if b is None:
            b = MsmBatch(sizes, g2=g2, ctx=self.ctx_g2 if g2 else self.ctx_g1, shared_workspace=True)
            b.uses = 0
            self.batches[key] = b
            while len(self.batches) > 64:
                _, old = self.batches.popitem(last=False)
                old.close()
        else:
            self.batches.move_to_end(key)
        return b


class ProofServer:
    """Continuous-batching Groth16 server over resident proving keys.

    keys: {name: ProvingKey}; budget_bytes: HBM for per-request state;
    large_log2m: circuits with log2(m) >= this go to the large lane.
    """

    def __init__(self, keys: dict, budget_bytes: int, max_inflight: int = 64, large_log2m: int = 20,
                 device: int = 0, use_graph: bool = True):
        self.lanes = [_Lane(device), _Lane(device)]
        ctx = self.lanes[0].ctx_g1
        self.circuits = {name: _Circuit(name, pk, int(pk.log2m >= large_log2m), ctx) for name, pk in keys.items()}
        self.adm = Admission(budget_bytes, max_inflight, lanes=2)
        self.queue: collections.deque[Request] = collections.deque()
        self.use_graph = use_graph
        self.done: list[Request] = []
        self.rejected: list[Request] = []

    # ---------------------------------------------------------------- intake
    def submit(self, req: Request) -> None:
        if req.circuit not in self.circuits:
            req.error = "unknown circuit %r" % req.circuit
            self.rejected.append(req)
            return
        self.queue.append(req)

    def _admit(self, now: float) -> bool:
        progressed = False
        while self.queue:
            req = self.queue[0]
            cc = self.circuits[req.circuit]
            new_slot = 0 if cc.idle else cc.slot_bytes
            if self.adm.never_fits(cc.slot_bytes, cc.msm_bytes):
                self.queue.popleft()
                req.error = "request needs %d bytes, budget %d" % (cc.slot_bytes + cc.msm_bytes, self.adm.budget)
                self.rejected.append(req)
                continue
            if not self.adm.fits(new_slot, cc.lane, cc.msm_bytes):
                if new_slot and self._drop_idle_slots(except_circuit=cc.name):
                    continue
                break
            self.queue.popleft()
            self.adm.take(new_slot, cc.lane, cc.msm_bytes)
            lane = self.lanes[cc.lane]
            req.slot = cc.idle.pop() if cc.idle else ProofSlot(cc.pk, lane.ctx_h)
            req.need, req.lane, req.t_admit = cc.msm_bytes, cc.lane, now
            # witness upload on the null stream (synchronous), then QAP + H on the lane's H stream
            req.slot.load(req.witness, req.r, req.s, stream=None)
            req.slot.stage_qap(lane.ctx_h, lane.st_h.ptr)
            req.slot.stage_h(lane.ctx_h, lane.st_h.ptr)
            req.ev_h = Event(lane.ctx_h)
            req.ev_h.record(lane.st_h)
            lane.pending.append(req)
            progressed = True
        return progressed

    def _drop_idle_slots(self, except_circuit: str) -> bool:
        freed = False
        for name, cc in self.circuits.items():
            if name == except_circuit:
                continue
            while cc.idle:
                sl = cc.idle.pop()
                sl.close()
                self.adm.free_slot(cc.slot_bytes)
                freed = True
        return freed

    # --------------------------------------------------------------- cohorts
    def _launch(self, li: int) -> bool:
        lane = self.lanes[li]
        if lane.cohort is not None or not lane.pending:
            return False
        ready = []
        while lane.pending and lane.pending[0].ev_h.done():
            ready.append(lane.pending.popleft())
        if not ready:
            return False
        g1 = [m for req in ready for m in req.slot.g1_members()]
        g2 = [req.slot.g2_member() for req in ready]
        fin = [req.slot.final_member() for req in ready]
        b1 = lane.batch([m[2] for m in g1], g2=False)
        bf = lane.batch([4] * len(ready), g2=False)
        b2 = lane.batch([m[2] for m in g2], g2=True)
        self.adm.cohort(li, b1.info()["bytes"] + b2.info()["bytes"])
        for b, mem, st in ((b1, g1, lane.st_g1), (bf, fin, lane.st_g1), (b2, g2, lane.st_g2)):
            b.run([m[0] for m in mem], [m[1] for m in mem], [m[3] for m in mem],
                  use_graph=self.use_graph and b.uses > 0, stream=st.ptr)
            b.uses += 1
        lane.ev_g1.record(lane.st_g1)
        lane.ev_g2.record(lane.st_g2)
        lane.cohort = ready
        lane.cohorts += 1
        lane.cohort_sizes.append(len(ready))
        return True

    def _complete(self, li: int, now_fn) -> bool:
        lane = self.lanes[li]
        if lane.cohort is None or not (lane.ev_g1.done() and lane.ev_g2.done()):
            return False
        t = now_fn()
        for req in lane.cohort:
            req.proof = req.slot.result()
            req.t_done = t
            cc = self.circuits[req.circuit]
            cc.idle.append(req.slot)
            req.slot = None
            req.ev_h.close()
            req.ev_h = None
            self.adm.release(req.lane, req.need)
            self.done.append(req)
        lane.cohort = None
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
        return bool(self.queue) or any(l.cohort is not None or l.pending for l in self.lanes)

    def run(self, requests, poll_s: float = 2e-4) -> list[Request]:
        """Serve requests whose t_arrival are offsets (s) from now; returns them
        once all are done or rejected."""
        reqs = sorted(requests, key=lambda q: q.t_arrival)
        t0 = time.perf_counter()
        now_fn = lambda: time.perf_counter() - t0  # noqa: E731
        i = 0
        while i < len(reqs) or self.busy():
            now = now_fn()
            while i < len(reqs) and reqs[i].t_arrival <= now:
                self.submit(reqs[i])
                i += 1
            if not self.step(now_fn):
                wait = reqs[i].t_arrival - now_fn() if i < len(reqs) else poll_s
                time.sleep(min(max(wait, 0.0), poll_s))
        return reqs

    def stats(self) -> dict:
        return {"cohorts": [l.cohorts for l in self.lanes],
                "mean_cohort": [float(np.mean(l.cohort_sizes)) if l.cohort_sizes else 0.0 for l in self.lanes],
                "rejected": len(self.rejected)}


class PlainServer:
    """Baseline: one request at a time through the plain per-request prover."""

    def __init__(self, keys: dict, device: int = 0):
        self.ctx = _capi.Context(device)
        from .tune import apply
        apply(self.ctx)
        self.provers = {name: GrothProver(pk, ctx=self.ctx) for name, pk in keys.items()}

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
