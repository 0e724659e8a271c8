"""GPU NTT / INTT / H step / QAP evaluation parity against the CPU oracle
(restated ntt_h, intt_h, ntt_interpolandmul and mpoly_eval_h), bit-exact."""
import numpy as np
import pytest

import oracle as O
from cusnarks_amd import inputs
from cusnarks_amd.poly import CSR, interpolandmul_host, mpoly_eval_host, ntt_host
from oracle import pyoracle as P

pytestmark = pytest.mark.gpu


def mont_elems(seed, n):
    return O.to_montgomeryN(inputs.field_elements(seed, n), 1)


@pytest.mark.parametrize("lg", [0, 1, 2, 5, 10, 11, 12, 13, 15])
def test_ntt_vs_oracle(lg):
    a = mont_elems(lg, 1 << lg)
    assert np.array_equal(ntt_host(a), O.ntt(a, lg))


@pytest.mark.parametrize("lg", [0, 1, 4, 10, 11, 14])
@pytest.mark.parametrize("fmt", [0, 1])
def test_intt_vs_oracle(lg, fmt):
    a = mont_elems(100 + lg, 1 << lg)
    assert np.array_equal(ntt_host(a, inverse=True, fmt=fmt), O.intt(a, lg, fmt=fmt))


@pytest.mark.parametrize("lg", [20, 21])
def test_ntt_roundtrip_and_dft_points(lg):
    # full-size properties: INTT(NTT(a)) == a, and two outputs against the DFT definition
    a = mont_elems(7, 1 << lg)
    X = ntt_host(a)
    back = ntt_host(X, inverse=True, fmt=1)
    assert np.array_equal(back, a)
    av = [P.from_mont(v, P.RR) for v in inputs.to_ints(a)]
    w = P.root_of_unity(lg)
    for k in (1, (1 << lg) - 3):
        wk = pow(w, k, P.RR)
        acc, t = 0, 1
        for x in av:
            acc = (acc + x * t) % P.RR
            t = t * wk % P.RR
        assert P.from_mont(inputs.to_ints(X[k:k + 1])[0], P.RR) == acc


@pytest.mark.parametrize("lg", [11, 12, 14])
def test_interpolandmul_vs_oracle(lg):
    m = 1 << lg
    pa, pb = mont_elems(200 + lg, m), mont_elems(300 + lg, m)
    assert np.array_equal(interpolandmul_host(pa, pb), O.interpolandmul(pa, pb, lg))


def test_interpolandmul_high_half_is_polynomial_product():
    # H coefficients are the schoolbook product of the interpolants (checked at m = 2^11)
    lg, m = 11, 1 << 11
    pa, pb = mont_elems(1, m), mont_elems(2, m)
    out = inputs.to_ints(interpolandmul_host(pa, pb))
    a = P.ntt([P.from_mont(v, P.RR) for v in inputs.to_ints(pa)], lg, inverse=True)
    b = P.ntt([P.from_mont(v, P.RR) for v in inputs.to_ints(pb)], lg, inverse=True)
    # coefficient k of A*B for a few k in the high half
    for k in (m, m + 1, 2 * m - 2):
        ck = sum(a[i] * b[k - i] for i in range(k - m + 1, m)) % P.RR
        assert out[k] == ck
    assert out[2 * m - 1] == 0


def _random_csr(seed, nrows, nwires, per_row):
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(nrows), per_row)
    cols = rng.integers(0, nwires, size=rows.size)
    vals = mont_elems(seed + 1, rows.size)
    return rows, cols, vals


@pytest.mark.parametrize("nrows,nwires", [(1, 1), (64, 50), (4096, 3000)])
def test_mpoly_eval_csr_vs_oracle(nrows, nwires):
    rows, cols, vals = _random_csr(nrows, nrows, nwires, 3)
    csr = CSR.from_triples(rows, cols, vals, nrows)
    w = inputs.field_elements(9, nwires)
    exp = O.mpoly_eval_csr(csr.rowptr, csr.col, csr.val, w, nrows)
    from cusnarks_amd.device import DeviceBuffer
    from cusnarks_amd.poly import mpoly_eval_device, upload_csr
    from cusnarks_amd import _capi
    ctx = _capi.default_context()
    dev = upload_csr(csr, ctx)
    dw = DeviceBuffer.from_array(ctx, w)
    dout = DeviceBuffer(ctx, nrows * 32)
    mpoly_eval_device(dev, dw, dout, ctx=ctx)
    assert np.array_equal(dout.download((nrows, 8)), exp)


def test_mpoly_eval_reference_format():
    # per-wire polys in the reference layout, evaluated through mpoly_eval_h's path
    nw, m = 40, 32
    rng = np.random.default_rng(5)
    polys = []
    for wire in range(nw):
        k = int(rng.integers(0, 4))
        rws = rng.choice(m, size=k, replace=False)
        polys.append((rws, mont_elems(1000 + wire, k)))
    hdr = [nw] + [len(p[0]) for p in polys]
    body = []
    for rws, vals in polys:
        body += list(rws) + list(vals.reshape(-1))
    pdata = np.array(hdr + body, dtype=np.uint32)
    w = inputs.field_elements(11, nw)
    got = mpoly_eval_host(w, pdata, m, 3, nw)
    exp = [0] * m
    wv = inputs.to_ints(w)
    for wire in range(3, nw):
        rws, vals = polys[wire]
        for r, v in zip(rws, inputs.to_ints(vals)):
            exp[r] = (exp[r] + P.from_mont(v, P.RR) * wv[wire]) % P.RR
    assert [P.from_mont(v, P.RR) for v in inputs.to_ints(got)] == exp


def test_ntt_tables_shared_across_threads():
    # two host threads on two contexts grow the shared twiddle table while the
    # other transforms (the tables are per device; outgrown ones are retired,
    # not freed, so a launch never reads a freed table)
    import threading
    from cusnarks_amd import _capi
    from cusnarks_amd.poly import ntt_host
    sizes = {0: [8, 12, 16, 19], 1: [9, 13, 17, 20]}
    data = {t: [inputs.field_elements(500 + 10 * t + i, 1 << lg) for i, lg in enumerate(ls)]
            for t, ls in sizes.items()}
    out, errs = {}, []

    def work(t):
        try:
            ctx = _capi.Context(0)
            out[t] = [ntt_host(a, ctx=ctx) for a in data[t]]
            ctx.close()
        except Exception as e:  # reported below
            errs.append(e)

    th = [threading.Thread(target=work, args=(t,)) for t in sizes]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errs, errs
    for t, ls in sizes.items():
        for a, lg, got in zip(data[t], ls, out[t]):
            assert np.array_equal(got, O.ntt(a, lg)), (t, lg)


@pytest.mark.parametrize("lg,count", [(4, 3), (11, 1), (11, 5), (12, 33), (15, 4)])
def test_interpolandmul_batch_equals_single(lg, count):
    # members batched in one launch per NTT pass (33: two groups of at most 32);
    # lg 4 takes the member-by-member path below the radix-4 passes
    from cusnarks_amd import _capi
    from cusnarks_amd.device import DeviceBuffer
    from cusnarks_amd.poly import interpolandmul_batch_device
    ctx = _capi.default_context()
    m = 1 << lg
    pa = [inputs.field_elements(7000 + 2 * t + lg, m) for t in range(count)]
    pb = [inputs.field_elements(7001 + 2 * t + lg, m) for t in range(count)]
    da = [DeviceBuffer.from_array(ctx, a) for a in pa]
    db = [DeviceBuffer.from_array(ctx, b) for b in pb]
    do = [DeviceBuffer(ctx, 2 * m * 32) for _ in range(count)]
    interpolandmul_batch_device(da, db, lg, do, ctx=ctx)
    for t in range(count):
        got = do[t].download((2 * m, 8))
        assert np.array_equal(got, interpolandmul_host(pa[t], pb[t])), t
    if lg <= 12:
        assert np.array_equal(do[0].download((2 * m, 8)), O.interpolandmul(pa[0], pb[0], lg))
