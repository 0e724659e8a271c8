"""Proof serving: every proof served in the mix (continuous batching across
in-flight proofs, two lanes) equals the proof of the same request proved
alone by the plain prover, bit for bit."""
import numpy as np
import pytest

from cusnarks_amd import inputs
from cusnarks_amd.prover import GrothProver, ProvingKey
from cusnarks_amd.serve import PlainServer, ProofServer, Request, poisson_requests, summarize
from cusnarks_amd.synth import Circuit, KeyScalars
from oracle import pyoracle as P

pytestmark = pytest.mark.gpu

SHAPES = {"c60": (60, 50, 2), "c300": (300, 260, 1), "c1500": (1500, 1400, 3), "c5000": (5000, 4500, 1)}


@pytest.fixture(scope="module")
def keys():
    out, wit = {}, {}
    for i, (name, (nc, nv, npub)) in enumerate(sorted(SHAPES.items())):
        circ = Circuit(nc, nv, npub, seed=40 + i)
        out[name] = ProvingKey(circ, KeyScalars(circ, seed=40 + i))
        wit[name] = circ.witness
    return out, wit


def test_served_equals_alone(keys):
    pks, wit = keys
    reqs = poisson_requests(2000.0, 40, wit, seed=3)
    srv = ProofServer(pks, budget_bytes=1 << 30, max_inflight=16, large_log2m=11)
    assert {c.lane for c in srv.circuits.values()} == {0, 1}  # both lanes in use
    srv.run(reqs)
    assert not srv.rejected and all(q.proof is not None for q in reqs)
    assert max(srv.stats()["mean_cohort"]) > 1  # proofs were batched together
    alone = {name: GrothProver(pk) for name, pk in pks.items()}
    for q in reqs:
        exp = alone[q.circuit].prove(q.witness, q.r, q.s)
        for got, e in zip(q.proof, exp):
            assert np.array_equal(got, e), (q.rid, q.circuit)
    st = summarize(reqs)
    assert st["proofs"] == 40 and st["p99_ms"] >= st["p50_ms"] > 0


def test_served_from_tables_equals_alone():
    """Circuits whose keys carry fixed-base tables are served through batched
    table runs (several proofs per run, mixed with a circuit without tables in
    the same cohorts); every proof equals the plain prover's without tables."""
    pks, wit = {}, {}
    for i, (nc, nv, npub) in enumerate([(300, 260, 1), (1500, 1400, 3), (5000, 4500, 1)]):
        circ = Circuit(nc, nv, npub, seed=70 + i)
        name = "t%d" % nc
        pks[name] = ProvingKey(circ, KeyScalars(circ, seed=70 + i))
        wit[name] = circ.witness
    reqs = poisson_requests(0, 30, wit, seed=5)  # a burst: cohorts of several proofs
    alone = {name: GrothProver(pk) for name, pk in pks.items()}
    exp = {q.rid: alone[q.circuit].prove(q.witness, q.r, q.s) for q in reqs}
    for name in ("t300", "t5000"):
        pks[name].build_tables()
    assert not any(c.tables for c in ProofServer(pks, budget_bytes=1 << 30).circuits.values())  # opt-in
    srv = ProofServer(pks, budget_bytes=1 << 30, max_inflight=16, use_tables=True)
    assert sum(1 for c in srv.circuits.values() if c.tables) == 2
    srv.run(reqs)
    assert not srv.rejected and max(srv.stats()["mean_cohort"]) > 1
    for q in reqs:
        assert all(np.array_equal(a, b) for a, b in zip(q.proof, exp[q.rid])), (q.rid, q.circuit)
    for pk in pks.values():
        pk.close_tables()


def test_plain_server_matches_and_r_s_draw(keys):
    pks, wit = keys
    reqs = poisson_requests(0, 6, wit, seed=9)  # all at t=0
    import random
    g = random.Random(reqs[2].rid)
    assert (reqs[2].r, reqs[2].s) == (g.randint(1, inputs.R - 1), g.randint(1, inputs.R - 1))
    PlainServer(pks).run(reqs)
    alone = {name: GrothProver(pk) for name, pk in pks.items()}
    for q in reqs:
        assert all(np.array_equal(a, b) for a, b in zip(q.proof, alone[q.circuit].prove(q.witness, q.r, q.s)))
    # the host-resident-key baseline gives the same proofs
    hk = [Request(q.rid, q.circuit, q.witness, q.r, q.s) for q in reqs]
    PlainServer(pks, host_keys=True).run(hk)
    for q, h in zip(reqs, hk):
        assert all(np.array_equal(a, b) for a, b in zip(q.proof, h.proof))


def test_budget_rejects_oversized_and_admits_rest(keys):
    pks, wit = keys
    small_only = ProofServer(pks, budget_bytes=1, max_inflight=4, small_share=1.0)
    big = small_only.circuits["c5000"]
    small = small_only.circuits["c60"]
    budget = small.slot_bytes + small.msm_bytes + 4096
    assert budget < big.slot_bytes + big.msm_bytes
    srv = ProofServer(pks, budget_bytes=budget, max_inflight=4, small_share=1.0)
    reqs = [Request(i, name, wit[name], 5 + i, 7 + i) for i, name in enumerate(["c60", "c5000", "c60", "c60"])]
    srv.run(reqs)
    assert [q.rid for q in srv.rejected] == [1] and "budget" in reqs[1].error
    done = [q for q in reqs if q.proof is not None]
    assert len(done) == 3
    pr = GrothProver(pks["c60"])
    for q in done:
        assert all(np.array_equal(a, b) for a, b in zip(q.proof, pr.prove(q.witness, q.r, q.s)))


def test_served_tiny_circuit_vs_python_oracle(keys):
    pks, wit = keys
    pk = pks["c60"]
    circ = Circuit(*SHAPES["c60"], seed=40 + sorted(SHAPES).index("c60"))
    reqs = poisson_requests(0, 3, {"c60": wit["c60"]}, seed=1)
    ProofServer({"c60": pk}, budget_bytes=1 << 28).run(reqs)
    g1 = lambda buf, n: [P.g1_from_words(w) for w in buf.download((n, 16))]  # noqa: E731
    nv, m = pk.n_vars, pk.m
    aff = {"A": g1(pk.A, nv + 2), "B1": g1(pk.B1, nv + 2), "C": g1(pk.C, nv), "hExps": g1(pk.hExps, m),
           "B2": [P.g2_from_words(w) for w in pk.B2.download((nv + 2, 32))]}

    def trip(csr):
        return [(r, int(csr.col[e]), P.from_mont(inputs.to_ints(csr.val[e:e + 1])[0], P.RR))
                for r in range(csr.nrows) for e in range(int(csr.rowptr[r]), int(csr.rowptr[r + 1]))]
    ta, tb = trip(circ.A), trip(circ.B)
    for q in reqs:
        ea, eb, ec = P.groth16_prove(aff, ta, tb, inputs.to_ints(q.witness), pk.n_public, pk.log2m, q.r, q.s)
        assert list(q.proof[0]) == P.g1_to_words(ea)
        assert list(q.proof[1]) == P.g2_to_words(eb)
        assert list(q.proof[2]) == P.g1_to_words(ec)
