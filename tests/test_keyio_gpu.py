"""Proving keys and witnesses read from the reference's file formats into HBM.

  * csk_ec_check accepts key points and both infinity encodings, and flags a
    point off the curve, a coordinate not below q and (G2) a twist point
    outside the order-r subgroup (built here from a square root in Fq2 and
    confirmed outside the subgroup by the textbook curve);
  * a key written to .bin (Montgomery, and normal form) and read back proves
    exactly what the in-memory key proves, and the proof is valid for the
    trapdoor the key came from; the witness goes through a .dat file;
  * a corrupted key point is rejected with the array and index named.
"""
import numpy as np
import pytest

from cusnarks_amd import _capi, inputs, keyio
from cusnarks_amd.device import DeviceBuffer
from cusnarks_amd.msm import fixed_base
from cusnarks_amd.prover import GrothProver, ProvingKey
from cusnarks_amd.synth import Circuit, KeyScalars, Trapdoor
from oracle import pyoracle as P

pytestmark = pytest.mark.gpu

Q = P.Q


def _check(words: np.ndarray, ec2: bool, subgroup: bool):
    ctx = _capi.default_context()
    d = DeviceBuffer.from_array(ctx, np.ascontiguousarray(words, dtype=np.uint32))
    res = (_capi.U64 * 2)()
    n = words.shape[0]
    _capi.check(_capi.lib().csk_ec_check(ctx.ptr, int(ec2), d.ptr, n, int(subgroup), res, None), "csk_ec_check")
    d.close()
    return int(res[0]), int(res[1])


def _f2_pow(a, e):
    out, base = (1, 0), a
    while e:
        if e & 1:
            out = P.f2_mul(out, base)
        base = P.f2_mul(base, base)
        e >>= 1
    return out


def _f2_sqrt(a):
    """Square root in Fq2 for q = 3 mod 4 (Adj and Rodriguez-Henriquez 2012,
    Algorithm 9); None for a non-square."""
    a1 = _f2_pow(a, (Q - 3) // 4)
    alpha = P.f2_mul(a1, P.f2_mul(a1, a))
    if P.f2_mul(_f2_pow(alpha, Q), alpha) == (Q - 1, 0):
        return None
    x0 = P.f2_mul(a1, a)
    if alpha == (Q - 1, 0):
        return P.f2_mul((0, 1), x0)
    return P.f2_mul(_f2_pow(P.f2_add((1, 0), alpha), (Q - 1) // 2), x0)


def _gen_points(scalars, g2):
    ctx = _capi.default_context()
    k = inputs.from_ints(scalars)
    d_k = DeviceBuffer.from_array(ctx, k)
    out = DeviceBuffer(ctx, len(scalars) * (128 if g2 else 64))
    fixed_base(d_k, len(scalars), out, g2=g2, ctx=ctx)
    return out.download((len(scalars), 32 if g2 else 16))


def test_point_check_g1():
    pts = _gen_points(list(range(1, 200)), False)
    assert _check(pts, False, False) == (0, 2 ** 64 - 1)
    inf = np.zeros((2, 16), np.uint32)
    inf[0, 8:] = P.to_limbs(P.to_mont(1))  # (0, 1*R) and (0, 0)
    assert _check(inf, False, False)[0] == 0
    bad = pts.copy()
    bad[17, 8] ^= 4  # y changed: off the curve
    bad[150, 0:8] = P.to_limbs(Q)  # x = q: not canonical
    assert _check(bad, False, False) == (2, 17)


def test_point_check_g2_subgroup():
    pts = _gen_points(list(range(5, 70)), True)
    assert _check(pts, True, True) == (0, 2 ** 64 - 1)
    # a twist point outside the order-r subgroup
    k = 1
    while True:
        x = (k, 1)
        y = _f2_sqrt(P.f2_add(P.f2_mul(P.f2_mul(x, x), x), P.B2))
        if y is not None:
            break
        k += 1
    assert P.f2_mul(y, y) == P.f2_add(P.f2_mul(P.f2_mul(x, x), x), P.B2)
    E2 = P.Curve(True)
    assert E2.on_curve((x, y)) and E2.mul(P.RR, (x, y)) is not None
    odd = np.array(P.g2_to_words((x, y)), np.uint32)
    mixed = pts.copy()
    mixed[40] = odd
    assert _check(mixed, True, False) == (0, 2 ** 64 - 1)  # on the twist
    assert _check(mixed, True, True) == (1, 40)
    mixed[3, 16] ^= 1  # y.c0 changed
    assert _check(mixed, True, False) == (1, 3)


def _trapdoor_key(seed):
    circ = Circuit(700, 600, 2, seed=seed, satisfied=True)
    td = Trapdoor(seed)
    return circ, td, ProvingKey(circ, KeyScalars.from_trapdoor(circ, td))


def _valid(proof, circ, td, r, s):
    aff = (P.g1_from_words(proof[0]), P.g2_from_words(proof[1]), P.g1_from_words(proof[2]))
    return P.groth16_check(aff, circ.triples("A"), circ.triples("B"), circ.triples("C"),
                           inputs.to_ints(circ.witness), circ.n_public, circ.log2m,
                           (td.tau, td.alpha, td.beta, td.gamma, td.delta), r, s)


def _to_normal(words: np.ndarray, p: int) -> np.ndarray:
    vals = [P.from_mont(v, p) for v in inputs.to_ints(words.reshape(-1, 8))]
    return inputs.from_ints(vals).reshape(words.shape)


def test_key_and_witness_files_prove_like_the_resident_key(tmp_path):
    circ, td, pk = _trapdoor_key(41)
    r, s = P.reference_rs(41)
    want = GrothProver(pk).prove(circ.witness, r, s)
    assert _valid(want, circ, td, r, s)

    words = keyio.proving_key_words(pk, circ)
    path = str(tmp_path / "pk.bin")
    words.astype("<u4").tofile(path)
    wpath = str(tmp_path / "w.dat")
    keyio.write_witness(wpath, circ.witness)
    pk2 = keyio.load_proving_key(path)
    assert (pk2.n_vars, pk2.n_public, pk2.m) == (pk.n_vars, pk.n_public, pk.m)
    got = GrothProver(pk2).prove(keyio.read_witness(wpath, pk2.n_vars), r, s)
    for g, e in zip(got, want):
        assert np.array_equal(g, e)

    # the same key in normal form (k_binformat 0): converted on the GPU
    pkf = keyio.read_pk_bin(path)
    nv, m = pkf.n_vars, pkf.domain_size
    arr = keyio.prover_arrays(pkf)
    hx = np.concatenate([np.asarray(pkf.hExps[:m + 1])])
    polsA, polsB = (keyio.csr_to_mpoly(c, nv) for c in (circ.A, circ.B))
    for pols in (polsA, polsB):
        npoly, nc, used = keyio.mpoly_counts(pols)
        vpos = []
        start = 1 + npoly
        for c in nc:
            vpos.append(np.arange(start + c, start + 9 * c))
            start += 9 * int(c)
        idx = np.concatenate(vpos)
        pols[idx] = _to_normal(pols[idx], P.RR).reshape(-1)
    norm = keyio.pk_bin_words(nv, pkf.n_public, pkf.domain_bits, _to_normal(arr["A"], Q), _to_normal(arr["B1"], Q),
                              _to_normal(arr["B2"], Q), _to_normal(arr["C"], Q), _to_normal(hx, Q), polsA, polsB,
                              binformat=keyio.FMT_EXT)
    npath = str(tmp_path / "pk_normal.bin")
    norm.astype("<u4").tofile(npath)
    got = GrothProver(keyio.load_proving_key(npath)).prove(circ.witness, r, s)
    for g, e in zip(got, want):
        assert np.array_equal(g, e)


def test_corrupted_key_point_is_rejected(tmp_path):
    circ, td, pk = _trapdoor_key(42)
    words = keyio.proving_key_words(pk, circ)
    counts = words[26:40].view("<u8")
    nv = circ.n_vars
    b2 = 40 + int(counts[0]) + int(counts[1]) + 112 + int(counts[2]) + int(counts[3])
    words[b2 + 32 * 7 + 16] ^= 2  # B2[7].y.c0
    path = str(tmp_path / "bad.bin")
    words.astype("<u4").tofile(path)
    with pytest.raises(keyio.KeyFormatError, match=r"B2: 1 of %d points .* index 7" % (nv + 2)):
        keyio.load_proving_key(path)
