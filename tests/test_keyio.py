"""Proving-key and witness files (CPU): the word layout of the reference's
README tables and pkbin_get offsets, round trips through the writers, and the
reader's rejections of malformed files.  Point values here are arbitrary words:
the curve checks run on the GPU (tests/test_keyio_gpu.py)."""
import json

import numpy as np
import pytest

from cusnarks_amd import inputs, keyio
from cusnarks_amd.poly import mpoly_to_csr
from cusnarks_amd.synth import Circuit

R, Q = inputs.R, keyio.Q


def _words(x):
    return [(x >> (32 * i)) & 0xFFFFFFFF for i in range(8)]


def _key(nc=200, nv=150, npub=2, seed=3):
    circ = Circuit(nc, nv, npub, seed=seed)
    rng = np.random.default_rng(seed)
    pts = lambda n, w: rng.integers(0, 1 << 32, size=(n, w), dtype=np.uint64).astype(np.uint32)  # noqa: E731
    nv, m = circ.n_vars, circ.m
    arrs = dict(A=pts(nv + 2, 16), B1=pts(nv + 2, 16), B2=pts(nv + 2, 32), C=pts(nv, 16), hExps=pts(m + 1, 16))
    arrs["B1"][nv + 1] = arrs["A"][nv + 1]  # both end with delta_1
    words = keyio.pk_bin_words(nv, npub, circ.log2m, arrs["A"], arrs["B1"], arrs["B2"], arrs["C"], arrs["hExps"],
                               keyio.csr_to_mpoly(circ.A, nv), keyio.csr_to_mpoly(circ.B, nv))
    return circ, arrs, words


def _triples(csr):
    rows = np.repeat(np.arange(csr.nrows), np.diff(csr.rowptr.astype(np.int64)))
    t = [(int(r), int(c), tuple(int(x) for x in v)) for r, c, v in zip(rows, csr.col, csr.val)]
    return sorted(t)


def test_header_layout_matches_the_reference_tables(tmp_path):
    circ, arrs, w = _key()
    nv, m = circ.n_vars, circ.m
    # README.md:733-752 / PKBIN_H_* offsets (src/python/constants.py:246-256)
    assert list(w[1:10]) == [1, 1, 256, 1, 2, nv, circ.n_public, circ.log2m, m]
    assert w[0] == w.size
    assert list(w[10:18]) == _words(R) and list(w[18:26]) == _words(Q)
    counts = w[26:40].view("<u8")
    pa, pb = keyio.csr_to_mpoly(circ.A, nv), keyio.csr_to_mpoly(circ.B, nv)
    # polsA / polsB padded to at least domainSize x 8 words (pkvars_to_bin ext=True)
    assert counts[0] == max(pa.size, 8 * m) and counts[1] == max(pb.size, 8 * m)
    assert list(counts[2:]) == [16 * (nv + 2), 16 * (nv + 2), 32 * (nv + 2), 16 * nv, 16 * (m + 1) + 96]
    # pkbin_get: offset_data = PKBIN_H_N_OFFSET + 2*8 + 14, EC data after both polys
    off = 10 + 16 + 14
    assert np.array_equal(w[off:off + pa.size], pa)
    ec = off + int(counts[0]) + int(counts[1])
    A = arrs["A"]
    assert np.array_equal(w[ec:ec + 16], A[nv]) and np.array_equal(w[ec + 32:ec + 48], A[nv + 1])  # alfa_1, delta_1
    assert np.array_equal(w[ec + 16:ec + 32], arrs["B1"][nv])  # beta_1
    assert np.array_equal(w[ec + 48:ec + 80], arrs["B2"][nv])  # beta_2
    assert np.array_equal(w[ec + 80:ec + 112], arrs["B2"][nv + 1])  # delta_2
    assert np.array_equal(w[ec + 112:ec + 112 + 16 * (nv + 2)].reshape(-1, 16), A)
    # hExps' extension: six points, the fourth delta_1 (pysnarks_utils.py:676)
    tail = w[-96:].reshape(6, 16)
    assert np.array_equal(tail[3], A[nv + 1]) and not tail[[0, 1, 2, 4, 5]].any()


def test_mpoly_round_trip_keeps_every_coefficient():
    circ = Circuit(300, 120, 3, seed=5)
    for csr in (circ.A, circ.B, circ.C):
        words = keyio.csr_to_mpoly(csr, circ.n_vars)
        npoly, nc, used = keyio.mpoly_counts(words)
        assert npoly == circ.n_vars and used == words.size and nc.sum() == csr.col.size
        back, wire = mpoly_to_csr(words, circ.m, 0, circ.n_vars)
        assert _triples(back) == _triples(csr)
        # per-wire rows ascending, as r1cs_to_mpoly_h writes them
        start = 1 + npoly
        for c in nc[:20]:
            rows = words[start:start + c]
            assert np.all(np.diff(rows.astype(np.int64)) >= 0)
            start += 9 * int(c)


def test_pk_round_trip(tmp_path):
    circ, arrs, w = _key()
    path = str(tmp_path / "pk.bin")
    w.astype("<u4").tofile(path)
    pkf = keyio.read_pk_bin(path)
    assert (pkf.n_vars, pkf.n_public, pkf.domain_size, pkf.binformat) == (circ.n_vars, circ.n_public, circ.m, 1)
    for k in ("A", "B1", "B2", "C"):
        assert np.array_equal(pkf.__dict__[k], arrs[k])
    # hExps' region also holds the six extension points
    assert np.array_equal(pkf.hExps[:circ.m + 1], arrs["hExps"]) and pkf.hExps.shape[0] == circ.m + 7
    pa = keyio.prover_arrays(pkf)
    m, nv = circ.m, circ.n_vars
    assert np.array_equal(pa["hExps"][:m - 1], arrs["hExps"][:m - 1])
    assert np.array_equal(pa["hExps"][m - 1], arrs["A"][nv + 1])  # delta_1 appended
    assert _triples(pa["qapA"]) == _triples(circ.A) and _triples(pa["qapB"]) == _triples(circ.B)


def test_pk_without_appended_points_gets_them_from_the_header(tmp_path):
    circ, arrs, w = _key()
    nv = circ.n_vars
    counts = w[26:40].view("<u8").copy()
    ec = 40 + int(counts[0]) + int(counts[1])
    head = w[:ec + 112].copy()
    hc = head[26:40].view("<u8")
    hc[2], hc[3], hc[4], hc[6] = 16 * nv, 16 * nv, 32 * nv, 16 * (circ.m + 1)
    short = np.concatenate([head, arrs["A"][:nv].reshape(-1), arrs["B1"][:nv].reshape(-1),
                            arrs["B2"][:nv].reshape(-1), arrs["C"].reshape(-1), arrs["hExps"].reshape(-1)])
    path = str(tmp_path / "short.bin")
    short.astype("<u4").tofile(path)
    pa = keyio.prover_arrays(keyio.read_pk_bin(path))
    for k in ("A", "B1", "B2"):
        assert np.array_equal(pa[k], arrs[k])


@pytest.mark.parametrize("what", ["ftype", "field", "truncated", "partial_point", "row", "npoly", "domain"])
def test_pk_reader_rejects_malformed_files(tmp_path, what):
    circ, arrs, w = _key()
    w = w.copy()
    if what == "ftype":
        w[1] = 3
    elif what == "field":
        w[10] ^= 1
    elif what == "truncated":
        w = w[:-1]
    elif what == "partial_point":
        w[26 + 2 * 5] = 16 * circ.n_vars + 8  # C count not whole points
    elif what == "row":
        w[40 + 1 + circ.n_vars] = circ.m  # first row of wire 0 outside the domain
    elif what == "npoly":
        w[40] = circ.n_vars - 1
    elif what == "domain":
        w[9] = circ.m + 1
    path = str(tmp_path / "bad.bin")
    w.astype("<u4").tofile(path)
    with pytest.raises(keyio.KeyFormatError):
        keyio.prover_arrays(keyio.read_pk_bin(path))


@pytest.mark.parametrize("ext", [".bin", ".dat", ".json", ".txt"])
def test_witness_round_trip(tmp_path, ext):
    w = inputs.field_elements(77, 333)
    path = str(tmp_path / ("w" + ext))
    keyio.write_witness(path, w)
    assert np.array_equal(keyio.read_witness(path, 333), w)
    assert np.array_equal(keyio.read_witness(path, 100), w[:100])
    if ext in (".bin", ".dat"):
        assert np.array_equal(keyio.read_witness(path), w)  # count from the header


def test_witness_binary_layouts(tmp_path):
    w = inputs.field_elements(78, 5)
    keyio.write_witness(str(tmp_path / "w.bin"), w)
    raw = np.fromfile(str(tmp_path / "w.bin"), dtype="<u4")
    # writeWitnessFile_h: u64 count, u32 8, u32 0, values from word 4
    assert raw[0] == 5 and raw[1] == 0 and raw[2] == 8 and raw[3] == 0
    assert np.array_equal(raw[4:].reshape(-1, 8), w)
    keyio.write_witness(str(tmp_path / "w.dat"), w)
    raw = np.fromfile(str(tmp_path / "w.dat"), dtype="<u4")
    # readWitnessFile_h(fmt=0) seeks to byte 32
    assert raw[0] == 5 and raw[1] == 8 and not raw[2:8].any()
    assert np.array_equal(raw[8:].reshape(-1, 8), w)
    # a .dat written by another tool, read from byte 32 whatever its header holds
    other = np.concatenate([np.arange(8, dtype=np.uint32) + 100, w.reshape(-1)])
    other.astype("<u4").tofile(str(tmp_path / "o.dat"))
    assert np.array_equal(keyio.read_witness(str(tmp_path / "o.dat"), 5), w)


def test_witness_reader_rejections(tmp_path):
    w = inputs.field_elements(79, 4)
    p = str(tmp_path / "w.bin")
    keyio.write_witness(p, w)
    with pytest.raises(keyio.KeyFormatError):
        keyio.read_witness(p, 5)  # more values than the file holds
    big = str(tmp_path / "big.json")
    with open(big, "w") as fh:
        json.dump([str(R)], fh)  # not below r
    with pytest.raises(keyio.KeyFormatError):
        keyio.read_witness(big)
    with pytest.raises(keyio.KeyFormatError):
        keyio.read_witness(str(tmp_path / "w.csv"))
