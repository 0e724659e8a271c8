"""Groth16 prove on the GPU vs the oracle composition (bit-exact proofs)."""
import numpy as np
import pytest

import oracle as O
from cusnarks_amd import inputs
from cusnarks_amd.prover import GrothProver, ProvingKey
from cusnarks_amd.synth import Circuit, KeyScalars
from oracle import pyoracle as P

pytestmark = pytest.mark.gpu


def _affine_lists(pk: ProvingKey):
    g1 = lambda buf, n: [P.g1_from_words(w) for w in buf.download((n, 16))]  # noqa: E731
    nv, m = pk.n_vars, pk.m
    return {
        "A": g1(pk.A, nv + 2), "B1": g1(pk.B1, nv + 2), "C": g1(pk.C, nv), "hExps": g1(pk.hExps, m),
        "B2": [P.g2_from_words(w) for w in pk.B2.download((nv + 2, 32))],
    }


def _triples(csr):
    out = []
    for r in range(csr.nrows):
        for e in range(int(csr.rowptr[r]), int(csr.rowptr[r + 1])):
            out.append((r, int(csr.col[e]), P.from_mont(inputs.to_ints(csr.val[e:e + 1])[0], P.RR)))
    return out


@pytest.mark.parametrize("nc,nv,npub,seed", [(60, 50, 2, 1), (100, 120, 1, 2)])
def test_prove_small_vs_python_oracle(nc, nv, npub, seed):
    circ = Circuit(nc, nv, npub, seed=seed)
    keys = KeyScalars(circ, seed=seed)
    pk = ProvingKey(circ, keys)
    r, s = P.reference_rs(seed)
    a, b, c = GrothProver(pk).prove(circ.witness, r, s)
    ea, eb, ec = P.groth16_prove(_affine_lists(pk), _triples(circ.A), _triples(circ.B),
                                 inputs.to_ints(circ.witness), npub, circ.log2m, r, s)
    assert list(a) == P.g1_to_words(ea)
    assert list(b) == P.g2_to_words(eb)
    assert list(c) == P.g1_to_words(ec)


def _c_oracle_prove(circ, pk, r, s):
    """The same composition from the C oracle's restated reference routines."""
    nv, m, npub = circ.n_vars, circ.m, circ.n_public
    w = circ.witness
    pA = O.mpoly_eval_csr(circ.A.rowptr, circ.A.col, circ.A.val, w, m)
    pB = O.mpoly_eval_csr(circ.B.rowptr, circ.B.col, circ.B.val, w, m)
    polH = O.interpolandmul(pA, pB, circ.log2m)
    one = inputs.from_ints([1])
    rs_ = inputs.from_ints([r, s, (-r * s) % P.RR])
    A = pk.A.download((nv + 2, 16))
    B1 = pk.B1.download((nv + 2, 16))
    B2 = pk.B2.download((nv + 2, 32))
    C = pk.C.download((nv, 16))
    hE = pk.hExps.download((m, 16))
    pa, _ = O.ec_jacreduce(np.concatenate([w, one, rs_[0:1]]), A)
    pb, _ = O.ec_jacreduce(np.concatenate([w, one, rs_[1:2]]), B2, ec2=True)
    pb1, _ = O.ec_jacreduce(np.concatenate([w, one, rs_[1:2]]), B1)
    pcp, _ = O.ec_jacreduce(w[npub + 1:], C[npub + 1:])
    pc2, _ = O.ec_jacreduce(np.concatenate([polH[m:2 * m - 1], rs_[2:3]]), hE)
    pc, _ = O.ec_jacreduce(np.concatenate([one, one, rs_[1:2], rs_[0:1]]), np.stack([pc2, pcp, pa, pb1]))
    return pa, pb, pc


def test_prove_4k_constraints_vs_c_oracle():
    circ = Circuit(4000, 3000, 3, seed=7)
    pk = ProvingKey(circ, KeyScalars(circ, seed=7))
    r, s = P.reference_rs(7)
    got = GrothProver(pk).prove(circ.witness, r, s)
    exp = _c_oracle_prove(circ, pk, r, s)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


def test_prove_repeatable_and_randomised():
    circ = Circuit(200, 150, 1, seed=3)
    pk = ProvingKey(circ, KeyScalars(circ, seed=3))
    pr = GrothProver(pk)
    p1 = pr.prove(circ.witness, 5, 7)
    p2 = pr.prove(circ.witness, 5, 7)
    p3 = pr.prove(circ.witness, 6, 7)
    assert all(np.array_equal(x, y) for x, y in zip(p1, p2))
    assert not np.array_equal(p1[0], p3[0])  # pi_a depends on r
    assert np.array_equal(p1[1], p3[1])      # pi_b does not


def test_overlapped_prover_matches_plain():
    from cusnarks_amd.prover import OverlappedProver
    for nc, nv, npub, seed in [(60, 50, 2, 1), (3000, 2800, 1, 4)]:
        circ = Circuit(nc, nv, npub, seed=seed)
        pk = ProvingKey(circ, KeyScalars(circ, seed=seed))
        plain, ov = GrothProver(pk), OverlappedProver(pk)
        for r, s in [(5, 7), P.reference_rs(seed)]:
            exp = plain.prove(circ.witness, r, s)
            for _ in range(2):  # the second run replays the captured G1 batch
                got = ov.prove(circ.witness, r, s)
                assert all(np.array_equal(g, e) for g, e in zip(got, exp))


def test_table_prover_matches_plain():
    """Provers running the five multiexps from the key's fixed-base tables
    give the plain prover's proofs (default and explicit window widths);
    closing the tables returns the key to the plain path."""
    from cusnarks_amd.prover import OverlappedProver
    for nc, nv, npub, seed, widths in [(60, 50, 2, 1, (0, 0)), (3000, 2800, 1, 4, (0, 0)),
                                       (3000, 2800, 1, 4, (9, 6))]:
        circ = Circuit(nc, nv, npub, seed=seed)
        pk = ProvingKey(circ, KeyScalars(circ, seed=seed))
        rs = [(5, 7), P.reference_rs(seed)]
        exp = [GrothProver(pk).prove(circ.witness, r, s) for r, s in rs]
        nbytes = pk.build_tables(window_bits=widths[0], g2_window_bits=widths[1])
        assert nbytes == pk.table_bytes() > 0
        if widths[0]:
            assert all(pk.tables[k].info()["window_bits"] == widths[0] for k in ("A", "B1", "C", "H"))
            assert pk.tables["B2"].info()["window_bits"] == widths[1]
        provers = (GrothProver(pk), OverlappedProver(pk))
        for (r, s), e in zip(rs, exp):
            for pr in provers:
                got = pr.prove(circ.witness, r, s)
                assert all(np.array_equal(g, x) for g, x in zip(got, e)), (nc, widths, type(pr).__name__)
        pk.close_tables()
        assert pk.table_bytes() == 0
        got = GrothProver(pk).prove(circ.witness, *rs[1])
        assert all(np.array_equal(g, x) for g, x in zip(got, exp[1]))


def test_pinned_witness_proves_like_an_ordinary_array(tmp_path):
    """A witness in page-locked host memory (device.pinned_empty, and
    keyio.read_witness(..., pinned=True)) gives the same proofs, plain and
    overlapped; the buffer is freed with its last view."""
    import gc
    import weakref
    from cusnarks_amd import keyio
    from cusnarks_amd.device import pinned_empty
    from cusnarks_amd.prover import OverlappedProver
    circ = Circuit(3000, 2800, 2, seed=23)
    pk = ProvingKey(circ, KeyScalars(circ, seed=23))
    r, s = P.reference_rs(23)
    want = GrothProver(pk).prove(circ.witness, r, s)
    wp = pinned_empty(circ.witness.shape, np.uint32, ctx=pk.ctx)
    assert wp.shape == circ.witness.shape and wp.dtype == np.uint32
    wp[...] = circ.witness
    for prover in (GrothProver(pk), OverlappedProver(pk)):
        got = prover.prove(wp, r, s)
        assert all(np.array_equal(a, b) for a, b in zip(got, want))
    path = str(tmp_path / "w.bin")
    keyio.write_witness(path, circ.witness)
    wf = keyio.read_witness(path, circ.n_vars, pinned=True, ctx=pk.ctx)
    assert np.array_equal(wf, circ.witness)
    assert all(np.array_equal(a, b) for a, b in zip(GrothProver(pk).prove(wf, r, s), want))
    owner = wp.base
    while getattr(owner, "base", None) is not None:
        owner = owner.base
    probe = weakref.ref(owner)
    del wp, wf, owner
    gc.collect()
    assert probe() is None
