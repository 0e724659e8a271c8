"""Proof and public-input files in the reference prover's output formats.

write_proof / write_public follow GrothProver.write_proof / write_pdata
(cusnarks src/python/groth_prover.py:1690-1747):

* .json proof: {"pi_a": [x, y, "1"], "pi_b": [[x0, x1], [y0, y1], ["1", "0"]],
  "pi_c": [x, y, "1"], "protocol": "groth"} with affine coordinates as
  normal-form decimal strings (G2 coordinates as [c0, c1] of
  Fq2 = Fq[u]/(u^2 + 1)), dumped with indent 4 and sorted keys, plus a
  newline;
* .json public inputs: a list of normal-form decimal strings, indent 4, plus
  a newline;
* .bin proof: the normal-form words of pi_a (X, Y), pi_b (X.c0, X.c1, Y.c0,
  Y.c1) and pi_c, u32 little-endian (writeU256DataFile_h,
  src/cuda/file_utils.cpp:480-486);
* .bin public inputs: [count, signals...] as u32 words.

Proof points come in as this library's affine Montgomery words (16 / 32
u32, infinity (0, 1*R)).  The point at infinity, which a proof with random
r, s never contains, is written in projective form ("0", "1", "0").
"""
from __future__ import annotations

import json

import numpy as np

from . import inputs

Q = inputs.Q
_RINV_Q = pow(1 << 256, -1, Q)
_RINV_R = pow(1 << 256, -1, inputs.R)


def _words_to_int(w) -> int:
    w = np.ascontiguousarray(w, dtype=np.uint32).reshape(8)
    return sum(int(x) << (32 * i) for i, x in enumerate(w))


def _int_to_words(x: int) -> np.ndarray:
    return np.array([(x >> (32 * i)) & 0xFFFFFFFF for i in range(8)], dtype=np.uint32)


def from_mont_q(w) -> int:
    return _words_to_int(w) * _RINV_Q % Q


def to_mont_q(x: int) -> np.ndarray:
    return _int_to_words((x << 256) % Q)


def _is_inf(words: np.ndarray, nw: int) -> bool:
    x, y = words[:nw], words[nw:]
    one = to_mont_q(1)
    y_is_one = np.array_equal(y[:8], one) and not y[8:].any()
    return not x.any() and (not y.any() or y_is_one)


def g1_coords(words) -> list[str]:
    w = np.ascontiguousarray(words, dtype=np.uint32).reshape(16)
    if _is_inf(w, 8):
        return ["0", "1", "0"]
    return [str(from_mont_q(w[0:8])), str(from_mont_q(w[8:16])), "1"]


def g2_coords(words) -> list[list[str]]:
    w = np.ascontiguousarray(words, dtype=np.uint32).reshape(32)
    if _is_inf(w, 16):
        return [["0", "0"], ["1", "0"], ["0", "0"]]
    c = [str(from_mont_q(w[8 * i:8 * i + 8])) for i in range(4)]
    return [[c[0], c[1]], [c[2], c[3]], ["1", "0"]]


def proof_json(pi_a, pi_b, pi_c) -> str:
    proof = {"pi_a": g1_coords(pi_a), "pi_b": g2_coords(pi_b), "pi_c": g1_coords(pi_c), "protocol": "groth"}
    return json.dumps(proof, indent=4, sort_keys=True) + "\n"


def public_json(signals) -> str:
    """signals: (n, 8) u32 normal-form field elements (the public inputs)."""
    s = np.ascontiguousarray(signals, dtype=np.uint32).reshape(-1, 8)
    return json.dumps([str(_words_to_int(x)) for x in s], indent=4) + "\n"


def proof_bin(pi_a, pi_b, pi_c) -> np.ndarray:
    parts = [np.ascontiguousarray(p, dtype=np.uint32).reshape(-1, 8) for p in (pi_a, pi_b, pi_c)]
    return np.concatenate([_int_to_words(from_mont_q(w)) for p in parts for w in p])


def write_proof(path: str, pi_a, pi_b, pi_c) -> None:
    if path.endswith(".json"):
        with open(path, "w") as fh:
            fh.write(proof_json(pi_a, pi_b, pi_c))
    elif path.endswith(".bin"):
        proof_bin(pi_a, pi_b, pi_c).astype("<u4").tofile(path)
    else:
        raise ValueError("proof file must end in .json or .bin: %r" % path)


def write_public(path: str, signals) -> None:
    s = np.ascontiguousarray(signals, dtype=np.uint32).reshape(-1, 8)
    if path.endswith(".json"):
        with open(path, "w") as fh:
            fh.write(public_json(s))
    elif path.endswith(".bin"):
        np.concatenate([np.array([s.shape[0]], np.uint32), s.reshape(-1)]).astype("<u4").tofile(path)
    else:
        raise ValueError("public-input file must end in .json or .bin: %r" % path)


def read_proof_json(path: str):
    """(pi_a, pi_b, pi_c) as affine Montgomery words from a .json proof."""
    with open(path) as fh:
        d = json.load(fh)

    def g1(c):
        if c[2] == "0":
            return np.concatenate([np.zeros(8, np.uint32), to_mont_q(1)])
        return np.concatenate([to_mont_q(int(c[0])), to_mont_q(int(c[1]))])

    def g2(c):
        if c[2] == ["0", "0"]:
            return np.concatenate([np.zeros(16, np.uint32), to_mont_q(1), np.zeros(8, np.uint32)])
        return np.concatenate([to_mont_q(int(c[i][j])) for i in (0, 1) for j in (0, 1)])
    return g1(d["pi_a"]), g2(d["pi_b"]), g1(d["pi_c"])


def read_public_json(path: str) -> np.ndarray:
    with open(path) as fh:
        return np.stack([_int_to_words(int(x)) for x in json.load(fh)])
