"""Proving-key and witness files of the reference prover, read into HBM.

Proving key, .bin (cusnarks README.md:729-786; reader pkbin_get,
src/python/pysnarks_utils.py:544-618; writer pkvars_to_bin, :619-785;
header reader readU256PKFileHeader_h, src/cuda/file_utils.cpp:447-469).
Little-endian 32-bit words:

  0 nWords (total words; "unused" by readers)   5 k_ecformat (2 = affine)
  1 ftype (1 = proving key)                     6 nVars
  2 protocol (1 = Groth16)                      7 nPublic
  3 Rbitlen (256)                               8 domainBits
  4 k_binformat (1 Montgomery, 0 normal)        9 domainSize
  10..17 field_r, 18..25 group_q (8 words each)
  26..39 seven u64 word counts: polsA, polsB, A, B1, B2, C, hExps
  then polsA, polsB (the per-wire sparse format of src/cuda/mpoly.cpp:195-200,
  zero-padded to their counts), alfa_1, beta_1, delta_1 (G1, 16 words each),
  beta_2, delta_2 (G2, 32 words each), A, B1, B2, C, hExps (affine points).

The setup writes the extended layout (pkvars_to_bin(..., ext=True),
src/python/groth_setup.py:961): A, B1 and B2 hold nVars points followed by
(alfa_1 | beta_1 | beta_2, delta_1 | delta_2), exactly the nVars + 2 points the
prover's multiexps read (src/python/groth_prover.py:623, 648, 697); C holds
nVars points, the first nPublic + 1 zero (infinity); hExps holds domainSize + 1
points, of which the prover reads the first domainSize - 1 and appends delta_1
(groth_prover.py:253, 744-750).  Keys without the appended points (nVars
points per array) are read too: the reader appends them from the header's
copies, as the prover's own layout needs.

Witness files (README.md:814-849; readWitnessFile_h / writeWitnessFile_h,
src/cuda/file_utils.cpp:550-592; dispatch by extension,
src/python/groth_prover.py:963-1005): .bin has a 16-byte header (u64 count,
u32 word size 8, u32 0), .dat a 32-byte header, both followed by 8-word
little-endian values in normal form; .json is a list of decimal strings and
.txt one decimal per line.  As in the reference, nVars values are read (from
the key), whatever the header's count says.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass

import numpy as np

from . import _capi, inputs
from ._capi import P, check, lib
from .device import DeviceBuffer
from .poly import CSR, mpoly_to_csr, upload_csr

FTYPE_PK = 1          # SNARKSFILE_T_PK (src/python/constants.py:239)
PROTOCOL_GROTH = 1    # PROTOCOL_T_GROTH (constants.py:230)
FMT_EXT = 0           # normal-form values (constants.py:222)
FMT_MONT = 1          # Montgomery values (constants.py:223)
EC_AFFINE = 2         # EC_T_AFFINE (constants.py:235)
HEADER_WORDS = 40     # 10 fields + field_r + group_q + 7 u64 counts

Q = 21888242871839275222246405745257275088696311157297823662689037894645226208583
G1W, G2W = 16, 32


class KeyFormatError(ValueError):
    """A proving-key or witness file that does not follow the reference format."""


def _int_words(x: int) -> np.ndarray:
    return np.array([(x >> (32 * i)) & 0xFFFFFFFF for i in range(8)], dtype=np.uint32)


def _words_int(w) -> int:
    return sum(int(v) << (32 * i) for i, v in enumerate(w))


@dataclass
class PkFile:
    """The arrays of a proving-key .bin file (views into the file's words)."""
    n_vars: int
    n_public: int
    domain_bits: int
    domain_size: int
    binformat: int
    polsA: np.ndarray   # u32 words, per-wire sparse format (+ zero padding)
    polsB: np.ndarray
    alfa_1: np.ndarray  # (16,)
    beta_1: np.ndarray
    delta_1: np.ndarray
    beta_2: np.ndarray  # (32,)
    delta_2: np.ndarray
    A: np.ndarray       # (nA, 16)
    B1: np.ndarray
    B2: np.ndarray      # (nB2, 32)
    C: np.ndarray
    hExps: np.ndarray


# ------------------------------------------------------------------ sparse polys
def mpoly_counts(words: np.ndarray) -> tuple[int, np.ndarray, int]:
    """(npoly, coefficient count per poly, words used) of a per-wire sparse
    block [npoly, ncoeff_0 .. ncoeff_{npoly-1}, per poly: rows, values]."""
    if words.size < 1:
        raise KeyFormatError("empty sparse polynomial block")
    npoly = int(words[0])
    if words.size < 1 + npoly:
        raise KeyFormatError("sparse polynomial block shorter than its %d counts" % npoly)
    nc = np.asarray(words[1:1 + npoly], dtype=np.int64)
    used = 1 + npoly + 9 * int(nc.sum())
    if used > words.size:
        raise KeyFormatError("sparse polynomial block needs %d words, has %d" % (used, words.size))
    return npoly, nc, used


def csr_to_mpoly(csr: CSR, n_vars: int) -> np.ndarray:
    """The per-wire sparse format of a row-major CSR over n_vars wires: for each
    wire, its rows ascending and their coefficients (r1cs_to_mpoly_h's output,
    src/cuda/mpoly.cpp:201-263)."""
    nnz = csr.col.size
    rows = np.repeat(np.arange(csr.nrows, dtype=np.int64), np.diff(csr.rowptr.astype(np.int64)))
    order = np.argsort(csr.col, kind="stable")  # rows stay ascending within a wire
    wires = csr.col[order].astype(np.int64)
    if nnz and int(wires[-1]) >= n_vars:
        raise ValueError("column %d outside %d wires" % (int(wires[-1]), n_vars))
    counts = np.bincount(wires, minlength=n_vars).astype(np.int64)
    out = np.zeros(1 + n_vars + 9 * nnz, dtype=np.uint32)
    out[0] = n_vars
    out[1:1 + n_vars] = counts
    starts = 1 + n_vars + 9 * np.concatenate([[0], np.cumsum(counts)[:-1]])
    first = np.concatenate([[0], np.cumsum(counts)[:-1]])
    within = np.arange(nnz) - np.repeat(first, counts)
    base = np.repeat(starts, counts)
    out[base + within] = rows[order]
    vpos = base + np.repeat(counts, counts) + 8 * within
    out[vpos[:, None] + np.arange(8)[None, :]] = csr.val[order]
    return out


# ------------------------------------------------------------------ proving key
def read_pk_bin(path: str) -> PkFile:
    """Parse a proving-key .bin (memory-mapped: nothing is copied until the
    arrays are uploaded).  Raises KeyFormatError on any inconsistency."""
    size = os.path.getsize(path)
    if size % 4 or size < 4 * HEADER_WORDS:
        raise KeyFormatError("%s: %d bytes is not a proving-key file" % (path, size))
    w = np.memmap(path, dtype="<u4", mode="r")
    ftype, protocol, rbits, binfmt, ecfmt, nv, npub, dbits, dsize = (int(x) for x in w[1:10])
    if ftype != FTYPE_PK or protocol != PROTOCOL_GROTH:
        raise KeyFormatError("%s: file type %d / protocol %d, expected a Groth16 proving key (1 / 1)"
                             % (path, ftype, protocol))
    if rbits != 256 or ecfmt != EC_AFFINE or binfmt not in (FMT_EXT, FMT_MONT):
        raise KeyFormatError("%s: Rbitlen %d, ecformat %d, binformat %d unsupported" % (path, rbits, ecfmt, binfmt))
    if _words_int(w[10:18]) != inputs.R or _words_int(w[18:26]) != Q:
        raise KeyFormatError("%s: field_r / group_q are not BN254's" % path)
    if dbits >= 32 or dsize != 1 << dbits or dbits < 1:
        raise KeyFormatError("%s: domainSize %d != 2^domainBits (%d)" % (path, dsize, dbits))
    if nv < npub + 2:
        raise KeyFormatError("%s: nVars %d too small for nPublic %d" % (path, nv, npub))
    counts = [int(x) for x in np.asarray(w[26:40]).view("<u8")]
    names = ("polsA", "polsB", "A", "B1", "B2", "C", "hExps")
    need = HEADER_WORDS + counts[0] + counts[1] + 112 + sum(counts[2:])
    if need > w.size:
        raise KeyFormatError("%s: header counts need %d words, file has %d" % (path, need, w.size))
    for name, cnt, pw in zip(names[2:], counts[2:], (G1W, G1W, G2W, G1W, G1W)):
        if cnt % pw:
            raise KeyFormatError("%s: %s holds %d words, not whole points" % (path, name, cnt))
    off = HEADER_WORDS
    polsA = w[off:off + counts[0]]
    off += counts[0]
    polsB = w[off:off + counts[1]]
    off += counts[1]
    alfa_1, beta_1, delta_1 = w[off:off + 16], w[off + 16:off + 32], w[off + 32:off + 48]
    beta_2, delta_2 = w[off + 48:off + 80], w[off + 80:off + 112]
    off += 112
    arrs = []
    for cnt, pw in zip(counts[2:], (G1W, G1W, G2W, G1W, G1W)):
        arrs.append(w[off:off + cnt].reshape(-1, pw))
        off += cnt
    return PkFile(nv, npub, dbits, dsize, binfmt, polsA, polsB, alfa_1, beta_1, delta_1, beta_2, delta_2, *arrs)


def _with_tail(arr: np.ndarray, n_vars: int, tail, name: str) -> np.ndarray:
    if arr.shape[0] == n_vars + 2:
        return arr
    if arr.shape[0] == n_vars:
        return np.concatenate([arr, np.stack(tail)])
    raise KeyFormatError("%s holds %d points; nVars = %d needs %d or %d" % (name, arr.shape[0], n_vars,
                                                                        n_vars, n_vars + 2))


def prover_arrays(pkf: PkFile) -> dict:
    """The arrays in the prover's device layout: A, B1, B2 (nVars + 2 points),
    C (nVars), hExps (domainSize - 1 points + delta_1) and the QAP A / B as
    row-major CSR over domainSize rows."""
    nv, m = pkf.n_vars, pkf.domain_size
    A = _with_tail(pkf.A, nv, (pkf.alfa_1, pkf.delta_1), "A")
    B1 = _with_tail(pkf.B1, nv, (pkf.beta_1, pkf.delta_1), "B1")
    B2 = _with_tail(pkf.B2, nv, (pkf.beta_2, pkf.delta_2), "B2")
    if pkf.C.shape[0] < nv:
        raise KeyFormatError("C holds %d points, nVars = %d" % (pkf.C.shape[0], nv))
    if pkf.hExps.shape[0] < m - 1:
        raise KeyFormatError("hExps holds %d points, domainSize - 1 = %d needed" % (pkf.hExps.shape[0], m - 1))
    hExps = np.concatenate([pkf.hExps[:m - 1], pkf.delta_1.reshape(1, G1W)])
    qap = []
    for name, words in (("polsA", pkf.polsA), ("polsB", pkf.polsB)):
        npoly, nc, used = mpoly_counts(np.asarray(words))
        if npoly != nv:
            raise KeyFormatError("%s has %d polynomials, nVars = %d" % (name, npoly, nv))
        try:
            csr, _ = mpoly_to_csr(np.asarray(words[:used]), m, 0, nv)
        except ValueError as e:
            raise KeyFormatError("%s: %s" % (name, e)) from None
        qap.append(csr)
    return {"A": A, "B1": B1, "B2": B2, "C": pkf.C[:nv], "hExps": hExps, "qapA": qap[0], "qapB": qap[1]}


def load_proving_key(path: str, ctx=None, check_points: bool = True, stream=None):
    """Read a proving-key .bin into a device-resident ProvingKey.

    Normal-form files (k_binformat 0) are converted to Montgomery form on the
    GPU.  With check_points every point is checked on the GPU (coordinates
    below q, on the curve, G2 points in the order-r subgroup) and a bad key
    raises KeyFormatError naming the array and index."""
    from .prover import ProvingKey
    ctx = ctx or _capi.default_context()
    pkf = read_pk_bin(path)
    arr = prover_arrays(pkf)
    L = lib()
    dev = {}
    try:
        for name in ("A", "B1", "B2", "C", "hExps"):
            a = np.ascontiguousarray(arr[name], dtype=np.uint32)
            buf = DeviceBuffer.from_array(ctx, a)
            dev[name] = buf
            if pkf.binformat == FMT_EXT:
                check(L.csk_field_op(ctx.ptr, _capi.OP_TOMONT, 0, buf.ptr, None, buf.ptr, a.size // 8, stream),
                      "csk_field_op")
            if check_points:
                ec2 = 1 if name == "B2" else 0
                res = (_capi.U64 * 2)()
                check(L.csk_ec_check(ctx.ptr, ec2, buf.ptr, a.shape[0], ec2, res, stream), "csk_ec_check")
                if res[0]:
                    raise KeyFormatError("%s: %d of %d points fail the curve%s check (first at index %d)"
                                         % (name, res[0], a.shape[0], " / subgroup" if ec2 else "", res[1]))
        qap = []
        for name in ("qapA", "qapB"):
            d = upload_csr(arr[name], ctx)
            if pkf.binformat == FMT_EXT and arr[name].val.size:
                v = d["val"]
                check(L.csk_field_op(ctx.ptr, _capi.OP_TOMONT, 1, v.ptr, None, v.ptr, arr[name].val.shape[0],
                                     stream), "csk_field_op")
            qap.append(d)
        check(L.csk_stream_sync(ctx.ptr, stream), "csk_stream_sync")
    except BaseException:
        for b in dev.values():
            b.close()
        raise
    return ProvingKey.from_device(ctx, pkf.n_vars, pkf.n_public, pkf.domain_bits, dev["A"], dev["B1"], dev["B2"],
                                  dev["C"], dev["hExps"], qap[0], qap[1])


def pk_bin_words(n_vars: int, n_public: int, domain_bits: int, A, B1, B2, C, hExps, polsA, polsB,
                 binformat: int = FMT_MONT) -> np.ndarray:
    """The .bin words of a proving key, in the extended layout the setup writes
    (pkvars_to_bin(FMT_MONT, EC_T_AFFINE, pk, ext=True)): A, B1, B2 with nVars + 2
    points (the last two are alfa_1|beta_1|beta_2 and delta_1|delta_2, which also
    fill the header slots), C with nVars points, hExps with domainSize + 1
    points followed by six zero points of which the fourth is delta_1, and
    polsA / polsB zero-padded to at least domainSize x 8 words.  Values are
    written as given (Montgomery for FMT_MONT)."""
    m = 1 << domain_bits
    A, B1, C, hExps = (np.ascontiguousarray(x, dtype=np.uint32).reshape(-1, G1W) for x in (A, B1, C, hExps))
    B2 = np.ascontiguousarray(B2, dtype=np.uint32).reshape(-1, G2W)
    for name, a in (("A", A), ("B1", B1), ("B2", B2)):
        if a.shape[0] != n_vars + 2:
            raise ValueError("%s needs nVars + 2 points" % name)
    if C.shape[0] != n_vars or hExps.shape[0] != m + 1:
        raise ValueError("C needs nVars points and hExps domainSize + 1")
    polsA = np.ascontiguousarray(polsA, dtype=np.uint32)
    polsB = np.ascontiguousarray(polsB, dtype=np.uint32)
    padA = max(polsA.size, 8 * m) - polsA.size
    padB = max(polsB.size, 8 * m) - polsB.size
    alfa_1, delta_1 = A[n_vars], A[n_vars + 1]
    if not np.array_equal(B1[n_vars + 1], delta_1):
        raise ValueError("A and B1 must both end with delta_1")
    beta_1, beta_2, delta_2 = B1[n_vars], B2[n_vars], B2[n_vars + 1]
    h_ext = np.zeros((6, G1W), dtype=np.uint32)
    h_ext[3] = delta_1
    counts = np.array([polsA.size + padA, polsB.size + padB, A.size, B1.size, B2.size, C.size,
                       hExps.size + h_ext.size], dtype="<u8")
    head = np.array([0, FTYPE_PK, PROTOCOL_GROTH, 256, binformat, EC_AFFINE, n_vars, n_public, domain_bits, m],
                    dtype=np.uint32)
    out = np.concatenate([head, _int_words(inputs.R), _int_words(Q), counts.view(np.uint32),
                          polsA, np.zeros(padA, np.uint32), polsB, np.zeros(padB, np.uint32),
                          alfa_1, beta_1, delta_1, beta_2, delta_2,
                          A.reshape(-1), B1.reshape(-1), B2.reshape(-1), C.reshape(-1), hExps.reshape(-1),
                          h_ext.reshape(-1)])
    out[0] = out.size
    return out


def write_pk_bin(path: str, *args, **kw) -> None:
    """Write pk_bin_words(*args, **kw) to path."""
    pk_bin_words(*args, **kw).astype("<u4").tofile(path)


def proving_key_words(pk, circuit) -> np.ndarray:
    """The .bin words of a device-resident ProvingKey built for `circuit`
    (downloads its points).  hExps[domainSize - 1] and hExps[domainSize],
    which the prover never reads, are written as infinity (0, 1*R)."""
    m, nv = pk.m, pk.n_vars
    A = pk.A.download((nv + 2, G1W))
    B1 = pk.B1.download((nv + 2, G1W))
    B2 = pk.B2.download((nv + 2, G2W))
    C = pk.C.download((nv, G1W))
    h = pk.hExps.download((m, G1W))[:m - 1]
    inf = np.zeros((2, G1W), dtype=np.uint32)
    inf[:, 8:16] = _int_words((1 << 256) % Q)
    return pk_bin_words(nv, pk.n_public, pk.log2m, A, B1, B2, C, np.concatenate([h, inf]),
                        csr_to_mpoly(circuit.A, nv), csr_to_mpoly(circuit.B, nv))


# ------------------------------------------------------------------ witness
def _check_witness(w: np.ndarray, path: str) -> np.ndarray:
    ints = inputs.to_ints(w)
    bad = [i for i, v in enumerate(ints) if v >= inputs.R]
    if bad:
        raise KeyFormatError("%s: witness value %d is not below r" % (path, bad[0]))
    return w


def read_witness(path: str, n_vars: int | None = None, pinned: bool = False, ctx=None) -> np.ndarray:
    """(n_vars, 8) normal-form witness from a .bin, .dat, .json or .txt file.
    Binary files give their header count when n_vars is None.  pinned: return
    it in page-locked host memory (device.pinned_empty, in ctx), from which
    the prover's upload runs at the link's full rate."""
    ext = os.path.splitext(path)[1].lower()
    if ext in (".bin", ".dat"):
        start = 16 if ext == ".bin" else 32
        size = os.path.getsize(path)
        if size < start:
            raise KeyFormatError("%s: shorter than its %d-byte header" % (path, start))
        raw = np.memmap(path, dtype="<u4", mode="r")
        if n_vars is None:
            n_vars = int(raw[0]) if ext == ".dat" else int(np.asarray(raw[0:2]).view("<u8")[0])
        if start + 32 * n_vars > size:
            raise KeyFormatError("%s: %d values need %d bytes, file has %d" % (path, n_vars, start + 32 * n_vars,
                                                                              size))
        w = np.array(raw[start // 4:start // 4 + 8 * n_vars], dtype=np.uint32).reshape(-1, 8)
    elif ext in (".json", ".txt"):
        with open(path) as fh:
            vals = json.load(fh) if ext == ".json" else [ln.strip() for ln in fh if ln.strip()]
        vals = [int(v) for v in vals]
        if n_vars is not None:
            if len(vals) < n_vars:
                raise KeyFormatError("%s: %d values, %d needed" % (path, len(vals), n_vars))
            vals = vals[:n_vars]
        if any(v < 0 for v in vals):
            raise KeyFormatError("%s: negative witness value" % path)
        if any(v >= 1 << 256 for v in vals):
            raise KeyFormatError("%s: witness value wider than 256 bits" % path)
        w = inputs.from_ints(vals)
    else:
        raise KeyFormatError("%s: unknown witness extension (.bin, .dat, .json, .txt)" % path)
    w = _check_witness(w, path)
    if pinned:
        from .device import pinned_empty
        out = pinned_empty(w.shape, np.uint32, ctx=ctx)
        out[...] = w
        return out
    return w


def write_witness(path: str, w: np.ndarray) -> None:
    """Write a normal-form witness in the format the extension names: .bin as
    writeWitnessFile_h does (u64 count, u32 8, u32 0, values), .dat with the
    32-byte header (u32 count, u32 8, zeros), .json / .txt as decimals."""
    w = np.ascontiguousarray(w, dtype=np.uint32).reshape(-1, 8)
    ext = os.path.splitext(path)[1].lower()
    if ext == ".bin":
        head = np.zeros(4, dtype=np.uint32)
        head[0:2] = np.array([w.shape[0]], dtype="<u8").view(np.uint32)
        head[2] = 8
    elif ext == ".dat":
        head = np.zeros(8, dtype=np.uint32)
        head[0], head[1] = w.shape[0], 8
    elif ext in (".json", ".txt"):
        vals = [str(v) for v in inputs.to_ints(w)]
        with open(path, "w") as fh:
            fh.write(json.dumps(vals) if ext == ".json" else "\n".join(vals) + "\n")
        return
    else:
        raise ValueError("unknown witness extension %r" % ext)
    np.concatenate([head, w.reshape(-1)]).astype("<u4").tofile(path)
