"""Proof / public-input files in the reference prover's formats, pinned by
the sample files the reference ships (its write_proof / write_pdata
output): the text round-trips byte for byte, and the sample's coordinates,
read in the documented order, are points of G1 and of the G2 twist."""
import os

import numpy as np

from cusnarks_amd import inputs, proofio
from oracle import pyoracle as P
from tests.gpu_util import GOLDEN

PROOF = os.path.join(GOLDEN, "sample_proof.json")
PUBLIC = os.path.join(GOLDEN, "sample_public.json")


def test_sample_proof_round_trips_byte_for_byte():
    a, b, c = proofio.read_proof_json(PROOF)
    with open(PROOF) as fh:
        assert proofio.proof_json(a, b, c) == fh.read()


def test_sample_public_round_trips_byte_for_byte():
    s = proofio.read_public_json(PUBLIC)
    assert s.shape == (8, 8)
    with open(PUBLIC) as fh:
        assert proofio.public_json(s) == fh.read()


def test_sample_points_lie_on_the_curves():
    # pins the coordinate order: pi_b is [[x.c0, x.c1], [y.c0, y.c1]] of the twist
    a, b, c = proofio.read_proof_json(PROOF)
    assert P.G1.on_curve(P.g1_from_words(a)) and P.G1.on_curve(P.g1_from_words(c))
    assert P.G2.on_curve(P.g2_from_words(b))
    swapped = np.concatenate([b[8:16], b[0:8], b[24:32], b[16:24]])
    assert not P.G2.on_curve(P.g2_from_words(swapped))


def test_binary_formats(tmp_path):
    rng = np.random.default_rng(3)
    pts = [P.G1.mul(int(rng.integers(1, 2**62)), P.G1_GEN) for _ in range(2)]
    q2 = P.G2.mul(12345, P.G2_GEN)
    a, c = (np.array(P.g1_to_words(p), np.uint32) for p in pts)
    b = np.array(P.g2_to_words(q2), np.uint32)
    f = str(tmp_path / "proof.bin")
    proofio.write_proof(f, a, b, c)
    raw = np.fromfile(f, dtype="<u4")
    assert raw.size == 64
    vals = [int(sum(int(x) << (32 * i) for i, x in enumerate(raw[8 * k:8 * k + 8]))) for k in range(8)]
    assert vals[0:2] == list(pts[0]) and vals[6:8] == list(pts[1])
    assert vals[2:6] == [q2[0][0], q2[0][1], q2[1][0], q2[1][1]]
    pub = inputs.field_elements(4, 5)
    g = str(tmp_path / "public.bin")
    proofio.write_public(g, pub)
    raw = np.fromfile(g, dtype="<u4")
    assert raw[0] == 5 and np.array_equal(raw[1:].reshape(5, 8), pub)


def test_infinity_is_written_projectively():
    inf1 = np.array(P.g1_to_words(None), np.uint32)
    inf2 = np.array(P.g2_to_words(None), np.uint32)
    assert proofio.g1_coords(inf1) == ["0", "1", "0"]
    assert proofio.g2_coords(inf2)[2] == ["0", "0"]
