"""Host-side tests that need no GPU: input generator, the C ABI library's
exports, reference-format parsing and the bench's single-GPU contract."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from cusnarks_amd import inputs
from tests.conftest import ROOT


def test_splitmix64_reference_values():
    # SplitMix64 (Steele, Lea, Flood 2014) from seed 0: published first outputs
    v = inputs.splitmix64(0, 0, 3)
    assert [int(x) for x in v] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_field_elements_deterministic_and_in_range():
    a = inputs.field_elements(5, 1000)
    b = inputs.field_elements(5, 1000)
    assert np.array_equal(a, b)
    vals = inputs.to_ints(a)
    assert all(0 <= v < inputs.R for v in vals)
    assert all(v < (1 << 254) for v in vals)
    # prefix property: the first k draws do not depend on how many are asked for
    assert np.array_equal(inputs.field_elements(5, 10), a[:10])
    nz = inputs.to_ints(inputs.field_elements(6, 500, inputs.Q, nonzero=True))
    assert all(0 < v < inputs.Q for v in nz)


def _declared_symbols():
    with open(os.path.join(ROOT, "include", "csk.h")) as fh:
        src = fh.read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*(csk_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    so = os.path.join(ROOT, "cusnarks_amd", "libcsk.so")
    if not os.path.exists(so):
        pytest.skip("libcsk.so not built")
    import ctypes
    lib = ctypes.CDLL(so)  # loads without a GPU; no compute is called
    missing = [s for s in _declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert len(_declared_symbols()) >= 20


def test_mpoly_format_parse_matches_direct_reading():
    from cusnarks_amd.poly import mpoly_to_csr
    rng = np.random.default_rng(1)
    nw, m = 25, 16
    polys = []
    for wire in range(nw):
        k = int(rng.integers(0, 5))
        polys.append((rng.choice(m, size=k, replace=False),
                      rng.integers(0, 2**32, size=(k, 8), dtype=np.uint64).astype(np.uint32)))
    pdata = [nw] + [len(p[0]) for p in polys]
    for rws, vals in polys:
        pdata += list(rws) + list(vals.reshape(-1))
    pdata = np.array(pdata, dtype=np.uint32)
    csr, _ = mpoly_to_csr(pdata, m, 2, nw - 1)
    trip = sorted((int(r), int(wire), tuple(v)) for wire in range(2, nw - 1)
                  for r, v in zip(polys[wire][0], polys[wire][1]))
    got = []
    for r in range(m):
        for e in range(int(csr.rowptr[r]), int(csr.rowptr[r + 1])):
            got.append((r, int(csr.col[e]), tuple(csr.val[e])))
    assert sorted(got) == trip


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_product_has_no_oracle_imports():
    # the product package never imports, links or reads the oracle
    pkg = os.path.join(ROOT, "cusnarks_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in txt and "from oracle" not in txt, f
                assert "liboracle" not in txt, f


def _header_params():
    """{name: [param C types]} for every function declared in include/csk.h."""
    with open(os.path.join(ROOT, "include", "csk.h")) as fh:
        src = re.sub(r"/\*.*?\*/", "", fh.read(), flags=re.S)
    out = {}
    for m in re.finditer(r"(?:int|void|const char \*)\s*(csk_\w+)\s*\(([^)]*)\)\s*;", src):
        params = [p.strip() for p in m.group(2).split(",") if p.strip() and p.strip() != "void"]
        out[m.group(1)] = params
    return out


def test_ctypes_signatures_match_header():
    # every ctypes binding passes exactly the header's parameters, with pointers
    # as pointers and integers of the declared width (a pointer bound as u32
    # would be truncated)
    import ctypes
    from cusnarks_amd import _capi
    decl = _header_params()
    sigs = dict(_capi._SIGS)
    sigs.update(_capi._OPTIONAL)
    for name, (argtypes, _) in sigs.items():
        assert name in decl, name
        params = decl[name]
        assert len(params) == len(argtypes), (name, params, argtypes)
        for c_param, ct in zip(params, argtypes):
            is_ptr = "*" in c_param
            ct_ptr = ct is ctypes.c_void_p or ct is ctypes.c_char_p or hasattr(ct, "_type_") and \
                issubclass(ct, ctypes._Pointer)
            assert is_ptr == bool(ct_ptr), (name, c_param, ct)
            if not is_ptr:
                ctype = c_param.rsplit(" ", 1)[0].strip()
                width = {"uint32_t": 4, "int": 4, "uint64_t": 8}[ctype]
                assert ctypes.sizeof(ct) == width, (name, c_param, ct)
    assert set(decl) == set(sigs), set(decl) ^ set(sigs)


def test_log_uniform_sizes():
    from cusnarks_amd import inputs
    s = inputs.log_uniform_sizes(7, 64, 10, 22)
    assert len(s) == 64 and min(s) >= 1 << 10 and max(s) <= 1 << 22
    assert s == inputs.log_uniform_sizes(7, 64, 10, 22)
    # the uniforms are draw >> 11 scaled by 2^-53 (first draw of seed 7)
    z = int(inputs.splitmix64(7, 0, 1)[0]) >> 11
    assert s[0] == int(round(2.0 ** (10 + 12 * (z / 2.0 ** 53))))
    # log2 sizes spread over the whole range
    import math
    lg = sorted(math.log2(x) for x in s)
    assert lg[0] < 11 and lg[-1] > 21


def test_tuning_table_from_json():
    from cusnarks_amd import tune
    doc = {"classes": {"12": {"window_bits": 11, "chunk": 16}, "20": {"window_bits": 16, "chunk": 0}}}
    tab = tune.table_from_json(doc)
    assert len(tab) == 64
    assert (tab[12].window_bits, tab[12].chunk, tab[12].reduce_seg) == (11, 16, 0)
    assert tab[20].window_bits == 16 and tab[0].window_bits == 0
    # a persisted table, if present, names only sane settings
    import json
    for g2 in (False, True):
        path = tune.tuning_path(g2)
        if os.path.exists(path):
            with open(path) as fh:
                d = json.load(fh)
            for lg, e in d["classes"].items():
                assert 2 <= e["window_bits"] <= 22 and 0 <= int(lg) < 64
                assert e["chunk"] == 0 or 1 <= e["chunk"] <= 1 << 20


def test_default_chunk_mirrors_library():
    from cusnarks_amd import tune
    assert tune.default_chunk(1 << 24, 16) == 256
    assert tune.default_chunk(1 << 10, 10) == 8
    assert tune.default_chunk(1 << 20, 16) == 64


def test_field_elements_prefix_across_generation_chunks():
    # the stream is consumed in passes of 2^20 candidates; a longer draw
    # starts with the shorter one whatever the pass boundaries
    a = inputs.field_elements(9, (1 << 20) + 12345)
    b = inputs.field_elements(9, 700000)
    assert np.array_equal(a[:700000], b)
    assert len({bytes(r) for r in a[-1000:]}) == 1000


def test_product_fails_loudly_without_a_gpu():
    # no HIP device here: the C ABI reports an error code and message, the
    # Python layer raises -- nothing computes on the CPU instead
    import ctypes
    from cusnarks_amd import _capi
    if not os.path.exists(_capi.LIB_PATH):
        pytest.skip("libcsk.so not built")
    L = _capi.lib()
    n = ctypes.c_int(-1)
    rc = L.csk_device_count(ctypes.byref(n))
    if rc == 0 and n.value > 0:
        pytest.skip("a GPU is visible")
    ptr = _capi.P()
    assert L.csk_ctx_create(0, ctypes.byref(ptr)) in (_capi.CSK_ERR_NODEV, _capi.CSK_ERR_HIP)
    assert L.csk_last_error()
    with pytest.raises(_capi.CskError):
        _capi.Context(0)
    from cusnarks_amd.msm import msm_host
    with pytest.raises(_capi.CskError):
        msm_host(np.zeros((1, 8), np.uint32), np.zeros((1, 16), np.uint32))
    from cusnarks_amd import pycusnarks
    with pytest.raises(_capi.CskError):
        pycusnarks.montmultN_h(np.ones((2, 8), np.uint32), np.ones((2, 8), np.uint32), 1)


def test_null_arguments_are_errors_not_crashes():
    # argument checks that run before any device work
    import ctypes
    from cusnarks_amd import _capi
    if not os.path.exists(_capi.LIB_PATH):
        pytest.skip("libcsk.so not built")
    L = _capi.lib()
    assert L.csk_msm_batch_create(None, 0, 1, None, None, 0, None) == _capi.CSK_ERR_ARG
    assert L.csk_msm_set_tuning(None, 0, None, 0) == _capi.CSK_ERR_ARG
    assert L.csk_ctx_create(0, None) == _capi.CSK_ERR_ARG
    assert L.csk_event_query(None, None) == _capi.CSK_ERR_ARG
