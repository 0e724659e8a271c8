"""Regenerate the golden fixtures under tests/golden/ from the reference tree.

Run only in a container where the reference sources are mounted (default
/root/reference, or pass a path).  It copies DATA only:

  * the binary known-answer files of the reference's C unit tests
    (test/c/aux_data/*.tbin: 1024-point G1/G2 scalar-mul samples, per-point
    products and the multi-exponentiation results), byte for byte;
  * the numeric known-answer tables held in test/c/test_utils_host.cpp
    (Fr Montgomery products A_test * B_test = C_test, lines 94-163; the
    128th root of unity p_root128 and its powers roots_128_test, 90-297),
    written as JSON lists of u32 words;
  * the sample proof and public-input files written by the reference prover
    (src/python/proof.json, public.json: the write_proof / write_pdata output
    format, groth_prover.py:1690-1747), byte for byte, as
    sample_proof.json / sample_public.json.

Nothing of the reference is imported or executed.
"""
from __future__ import annotations

import json
import os
import re
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

TBIN = [
    "test_scmul_samples.tbin",
    "test_scmul_soln.tbin",
    "test_scmul_reduction_soln.tbin",
    "test2_scmul_samples.tbin",
    "test2_scmul_soln.tbin",
    "test2_scmul_reduction_soln.tbin",
]
TABLES = ["A_test", "B_test", "C_test", "p_root128", "roots_128_test"]


def parse_tables(src: str) -> dict[str, list[int]]:
    out = {}
    for name in TABLES:
        m = re.search(r"static uint32_t %s\[\]\s*=\s*\{(.*?)\};" % name, src, re.S)
        if not m:
            raise SystemExit("table %s not found" % name)
        body = re.sub(r"//[^\n]*", "", m.group(1))
        out[name] = [int(tok, 0) for tok in re.findall(r"0x[0-9a-fA-F]+|\d+", body)]
    return out


def main(ref: str = "/root/reference") -> None:
    aux = os.path.join(ref, "test", "c", "aux_data")
    for f in TBIN:
        shutil.copyfile(os.path.join(aux, f), os.path.join(HERE, f))
    with open(os.path.join(ref, "test", "c", "test_utils_host.cpp")) as fh:
        tables = parse_tables(fh.read())
    with open(os.path.join(HERE, "kat_tables.json"), "w") as fh:
        json.dump(tables, fh)
    for src, dst in (("proof.json", "sample_proof.json"), ("public.json", "sample_public.json")):
        shutil.copyfile(os.path.join(ref, "src", "python", src), os.path.join(HERE, dst))
    print("wrote", len(TBIN), "tbin files,", len(tables), "tables and 2 sample output files")


if __name__ == "__main__":
    main(*sys.argv[1:])
