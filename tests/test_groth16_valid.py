"""Groth16 validity on the CPU: a satisfied synthetic R1CS, a key from known
toxic scalars, and the oracle's exponent-level verifier.

These tests pin the verifier itself (it accepts the composition the
reference's formulas give and rejects planted mistakes), so the GPU proofs
checked with it in test_groth16_valid_gpu.py are shown to be valid proofs.
"""
import numpy as np
import pytest

import oracle as O
from cusnarks_amd import inputs
from cusnarks_amd.synth import Circuit, KeyScalars, Trapdoor, wire_evals
from oracle import pyoracle as P


def _points(keys):
    g1 = np.array(P.g1_to_words(P.G1_GEN), np.uint32)
    g2 = np.array(P.g2_to_words(P.G2_GEN), np.uint32)
    pts = {k: [P.g1_from_words(x) for x in O.ec_scmul_aff(getattr(keys, k), g1)]
           for k in ("A", "B1", "C", "hExps")}
    pts["B2"] = [P.g2_from_words(x) for x in O.ec_scmul_aff(keys.B2, g2, ec2=True)]
    return pts


def _setup(nc=24, nv=14, npub=2, seed=3):
    circ = Circuit(nc, nv, npub, seed=seed, satisfied=True)
    td = Trapdoor(seed)
    keys = KeyScalars.from_trapdoor(circ, td)
    return circ, td, keys


def _tdtuple(td):
    return (td.tau, td.alpha, td.beta, td.gamma, td.delta)


def test_satisfied_circuit_holds_row_by_row():
    circ, _, _ = _setup()
    w = inputs.to_ints(circ.witness)
    m = circ.m
    ev = [P.mpoly_eval(circ.triples(k), w, m) for k in "ABC"]
    for a, b, c in zip(*ev):
        assert a * b % P.RR == c
    # the unsatisfied generator is unchanged: some row fails there
    raw = Circuit(24, 14, 2, seed=3)
    ev = [P.mpoly_eval(raw.triples(k), w, m) for k in "ABC"]
    assert any(a * b % P.RR != c for a, b, c in zip(*ev))


def test_setup_wire_values_match_interpolation():
    # the setup's Lagrange route and the oracle's interpolate-then-Horner route
    circ, td, _ = _setup()
    at, bt, ct = wire_evals(circ, td.tau)
    for wire in (0, 1, 5, circ.n_vars - 1):
        unit = [0] * circ.n_vars
        unit[wire] = 1
        for name, ev in zip("ABC", (at, bt, ct)):
            assert ev[wire] == P.qap_at(circ.triples(name), unit, circ.log2m, td.tau)


def test_verifying_key_ic_matches_public_term():
    circ, td, keys = _setup()
    w = inputs.to_ints(circ.witness)
    _, _, _, ic = P.groth16_exponents(circ.triples("A"), circ.triples("B"), circ.triples("C"), w,
                                      circ.n_public, circ.log2m, _tdtuple(td), 1, 1)
    ics = inputs.to_ints(keys.IC)
    assert sum(x * y for x, y in zip(w, ics)) * td.gamma % P.RR == ic


def _prove_py(circ, td, keys, r, s, witness=None, **kw):
    pts = _points(keys)
    w = inputs.to_ints(circ.witness) if witness is None else witness
    return P.groth16_prove(pts, circ.triples("A"), circ.triples("B"), w, circ.n_public, circ.log2m, r, s, **kw)


def _check(proof, circ, td, r, s, witness=None):
    w = inputs.to_ints(circ.witness) if witness is None else witness
    return P.groth16_check(proof, circ.triples("A"), circ.triples("B"), circ.triples("C"), w,
                           circ.n_public, circ.log2m, _tdtuple(td), r, s)


def test_reference_composition_is_a_valid_proof():
    circ, td, keys = _setup()
    r, s = P.reference_rs(11)
    assert _check(_prove_py(circ, td, keys, r, s), circ, td, r, s)


def test_c_oracle_composition_is_a_valid_proof():
    # the same composition from the C restatements of the reference routines
    circ, td, keys = _setup(nc=40, nv=30, npub=1, seed=8)
    g1 = np.array(P.g1_to_words(P.G1_GEN), np.uint32)
    g2 = np.array(P.g2_to_words(P.G2_GEN), np.uint32)
    pts = {k: O.ec_scmul_aff(getattr(keys, k), g1) for k in ("A", "B1", "C", "hExps")}
    pts["B2"] = O.ec_scmul_aff(keys.B2, g2, ec2=True)
    r, s = P.reference_rs(8)
    m, npub = circ.m, circ.n_public
    w = circ.witness
    pA = O.mpoly_eval_csr(circ.A.rowptr, circ.A.col, circ.A.val, w, m)
    pB = O.mpoly_eval_csr(circ.B.rowptr, circ.B.col, circ.B.val, w, m)
    polH = O.interpolandmul(pA, pB, circ.log2m)
    one = inputs.from_ints([1])
    rs_ = inputs.from_ints([r, s, (-r * s) % P.RR])
    pa, _ = O.ec_jacreduce(np.concatenate([w, one, rs_[0:1]]), pts["A"])
    pb, _ = O.ec_jacreduce(np.concatenate([w, one, rs_[1:2]]), pts["B2"], ec2=True)
    pb1, _ = O.ec_jacreduce(np.concatenate([w, one, rs_[1:2]]), pts["B1"])
    pcp, _ = O.ec_jacreduce(w[npub + 1:], pts["C"][npub + 1:])
    pc2, _ = O.ec_jacreduce(np.concatenate([polH[m:2 * m - 1], rs_[2:3]]), pts["hExps"])
    pc, _ = O.ec_jacreduce(np.concatenate([one, one, rs_[1:2], rs_[0:1]]), np.stack([pc2, pcp, pa, pb1]))
    proof = (P.g1_from_words(pa), P.g2_from_words(pb), P.g1_from_words(pc))
    assert _check(proof, circ, td, r, s)


def test_verifier_rejects_swapped_r_s_in_pi_c():
    circ, td, keys = _setup()
    r, s = P.reference_rs(11)
    bad = _prove_py(circ, td, keys, r, s, swap_rs_in_c=True)
    assert not _check(bad, circ, td, r, s)


def test_verifier_rejects_wrong_randomness_and_witness():
    circ, td, keys = _setup()
    r, s = P.reference_rs(12)
    proof = _prove_py(circ, td, keys, r, s)
    assert not _check(proof, circ, td, s, r)
    w = inputs.to_ints(circ.witness)
    w2 = list(w)
    w2[-1] = (w2[-1] + 1) % P.RR
    assert not _check(proof, circ, td, r, s, witness=w2)


def test_verifier_rejects_proof_of_unsatisfying_witness():
    # an honest composition over a witness that breaks a constraint: the high
    # half of A(x)B(x) is then not the quotient (AB - C)/Z, and C fails
    circ, td, keys = _setup()
    r, s = P.reference_rs(13)
    w2 = inputs.to_ints(circ.witness)
    w2[-1] = (w2[-1] + 5) % P.RR
    m = circ.m
    ev = [P.mpoly_eval(circ.triples(k), w2, m) for k in "ABC"]
    assert any(a * b % P.RR != c for a, b, c in zip(*ev))
    proof = _prove_py(circ, td, keys, r, s, witness=w2)
    assert not _check(proof, circ, td, r, s, witness=w2)
    # the satisfying witness passes
    good = _prove_py(circ, td, keys, r, s)
    assert _check(good, circ, td, r, s)


def test_from_trapdoor_refuses_unsatisfied_circuit():
    with pytest.raises(ValueError):
        KeyScalars.from_trapdoor(Circuit(20, 12, 1, seed=1), Trapdoor(1))
