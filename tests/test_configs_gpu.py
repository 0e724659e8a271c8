"""The benchmark workloads at their full sizes, pinned by closed forms.

Every key point and multiexp input is k G with a known k, so each result is
(a known scalar) * G, computed from the oracle's field products; the H step
is pinned at a random point (polH(z) = A(z) B(z), tests.gpu_util):
  * the full Groth16 prove of a 2^20-constraint, 2^20-wire circuit (m = 2^21);
  * 64 batched G1 multiexps of sizes round(2^u), u ~ U[10, 22] (seed 7), as the
    bench draws them, eager and replayed from the captured graph;
  * the proof server over circuits of 2^14 .. 2^24 constraints (12 requests):
    every served proof equals the plain prover's, which equals the closed form.
"""
import numpy as np
import pytest

from cusnarks_amd import _capi, inputs
from cusnarks_amd.batch import MsmBatch
from cusnarks_amd.device import DeviceBuffer
from cusnarks_amd.msm import fixed_base
from cusnarks_amd.prover import GrothProver, ProvingKey
from cusnarks_amd.serve import ProofServer, poisson_requests
from cusnarks_amd.synth import Circuit, KeyScalars
from oracle import pyoracle as P
from tests.gpu_util import check_h_at_point, dot_mod_r, expected_proof, gen_mul

pytestmark = pytest.mark.gpu


def _prove_and_pin(pr: GrothProver, circ, keys, r, s, z):
    got = pr.prove(circ.witness, r, s)
    H = pr.slot.H.download((2 * circ.m, 8))
    check_h_at_point(circ, H, z)
    exp = expected_proof(circ, keys, H, r, s)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)
    return got


def test_prove_2e20_constraints_closed_form():
    circ = Circuit(1 << 20, 1 << 20, 1, seed=0)
    assert circ.m == 1 << 21
    keys = KeyScalars(circ, seed=0)
    pk = ProvingKey(circ, keys)
    r, s = P.reference_rs(2020)
    got = _prove_and_pin(GrothProver(pk), circ, keys, r, s, z=0x5EED2020)
    # the same proof from the key's fixed-base tables, plain and overlapped
    from cusnarks_amd.prover import OverlappedProver
    pk.build_tables()
    for pr in (GrothProver(pk), OverlappedProver(pk)):
        assert all(np.array_equal(g, e) for g, e in zip(pr.prove(circ.witness, r, s), got))
    pk.close_tables()


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_64_multiexps_closed_form(use_graph):
    ctx = _capi.default_context()
    seed, count = 7, 64
    sizes = inputs.log_uniform_sizes(seed, count, 10, 22)
    pts, scl, exp = [], [], []
    for j, n in enumerate(sizes):
        kp = inputs.field_elements(1000 * seed + 2 * j, n, nonzero=True)
        k = inputs.field_elements(1000 * seed + 2 * j + 1, n)
        d = DeviceBuffer.from_array(ctx, kp)
        p = DeviceBuffer(ctx, n * 64)
        fixed_base(d, n, p, ctx=ctx)
        _capi.check(_capi.lib().csk_stream_sync(ctx.ptr, None), "sync")
        d.close()
        pts.append(p)
        scl.append(DeviceBuffer.from_array(ctx, k))
        exp.append(gen_mul(dot_mod_r(k, kp)))
    out = DeviceBuffer(ctx, count * 64)
    batch = MsmBatch(sizes, ctx=ctx)
    for rep in range(2 if use_graph else 1):  # the second run replays the graph
        batch.run(scl, pts, [out.ptr + 64 * j for j in range(count)], use_graph=use_graph)
        got = out.download((count, 16))
        for j in range(count):
            assert np.array_equal(got[j], exp[j]), (rep, j, sizes[j])
    batch.close()


_SERVE_LGS = (14, 16, 18, 20, 22, 24)
_serve = {}


@pytest.mark.parametrize("lg", _SERVE_LGS)
def test_serving_keys_ready(lg):
    # one circuit per test, so a long setup reports progress as it goes
    circ = Circuit(1 << lg, 1 << lg, 1, seed=100 + lg)
    keys = KeyScalars(circ, seed=100 + lg)
    _serve[lg] = (circ, keys, ProvingKey(circ, keys))
    assert _serve[lg][2].m == 1 << (lg + 1)


def test_served_mix_equals_plain_equals_closed_form():
    if len(_serve) != len(_SERVE_LGS):
        pytest.skip("the key-building tests did not all run")
    keys = {"2^%d" % lg: v[2] for lg, v in _serve.items()}
    wit = {"2^%d" % lg: v[0].witness for lg, v in _serve.items()}
    reqs = poisson_requests(50.0, 12, wit, seed=21)
    assert len({q.circuit for q in reqs}) >= 4
    srv = ProofServer(keys, budget_bytes=96 << 30, max_inflight=64)
    srv.run(reqs)
    assert not srv.rejected and all(q.proof is not None for q in reqs)
    plain = {name: GrothProver(pk) for name, pk in keys.items()}
    for q in reqs:
        lg = int(q.circuit[2:])
        circ, ks, _ = _serve[lg]
        alone = _prove_and_pin(plain[q.circuit], circ, ks, q.r, q.s, z=0x5EED0000 + q.rid)
        for g, e in zip(q.proof, alone):
            assert np.array_equal(g, e), (q.rid, q.circuit)
    for v in plain.values():
        v.slot.close()
    _serve.clear()
