"""GPU Groth16 proofs are valid proofs: over a satisfied R1CS and a key from
known toxic scalars, each proof's points equal a G1, b G2, c G1 for the
exponents the Groth16 definition gives, and those satisfy the verification
equation (oracle.pyoracle.groth16_check, pinned on the CPU by
tests/test_groth16_valid.py).  The plain prover, the overlapped prover and
the proof server are all checked."""
import numpy as np
import pytest

from cusnarks_amd import inputs
from cusnarks_amd.prover import GrothProver, OverlappedProver, ProvingKey
from cusnarks_amd.serve import ProofServer, Request
from cusnarks_amd.synth import Circuit, KeyScalars, Trapdoor
from oracle import pyoracle as P

pytestmark = pytest.mark.gpu


def _valid(proof, circ, td, r, s):
    aff = (P.g1_from_words(proof[0]), P.g2_from_words(proof[1]), P.g1_from_words(proof[2]))
    return P.groth16_check(aff, circ.triples("A"), circ.triples("B"), circ.triples("C"),
                           inputs.to_ints(circ.witness), circ.n_public, circ.log2m,
                           (td.tau, td.alpha, td.beta, td.gamma, td.delta), r, s)


def _setup(nc, nv, npub, seed):
    circ = Circuit(nc, nv, npub, seed=seed, satisfied=True)
    td = Trapdoor(seed)
    return circ, td, ProvingKey(circ, KeyScalars.from_trapdoor(circ, td))


@pytest.mark.parametrize("nc,nv,npub,seed", [(1000, 900, 1, 21), (2000, 2100, 3, 22), (4000, 3500, 2, 23)])
def test_gpu_proof_is_valid(nc, nv, npub, seed):
    # m = 2^10, 2^11, 2^12
    circ, td, pk = _setup(nc, nv, npub, seed)
    r, s = P.reference_rs(seed)
    proof = GrothProver(pk).prove(circ.witness, r, s)
    assert _valid(proof, circ, td, r, s)
    # the checker is not vacuous on GPU output: swapped randomness fails
    assert not _valid(proof, circ, td, s, r)


def test_table_prover_proof_is_valid():
    circ, td, pk = _setup(2000, 2100, 3, 24)
    r, s = P.reference_rs(24)
    plain = GrothProver(pk).prove(circ.witness, r, s)
    pk.build_tables()
    proof = GrothProver(pk).prove(circ.witness, r, s)
    assert _valid(proof, circ, td, r, s)
    assert all(np.array_equal(a, b) for a, b in zip(proof, plain))
    pk.close_tables()


def test_overlapped_and_served_proofs_are_valid():
    circ, td, pk = _setup(1500, 1200, 2, 31)
    r, s = P.reference_rs(31)
    assert _valid(OverlappedProver(pk).prove(circ.witness, r, s), circ, td, r, s)
    reqs = []
    for i in range(4):
        rr, ss = P.reference_rs(100 + i)
        reqs.append(Request(100 + i, "c", circ.witness, rr, ss, t_arrival=0.0))
    srv = ProofServer({"c": pk}, budget_bytes=1 << 30, max_inflight=8)
    srv.run(reqs)
    assert not srv.rejected
    for q in reqs:
        assert _valid(q.proof, circ, td, q.r, q.s), q.rid
    assert np.array_equal(reqs[0].proof[1], GrothProver(pk).prove(circ.witness, reqs[0].r, reqs[0].s)[1])
