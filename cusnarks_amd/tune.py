"""Autotuner for the multiexp launch settings per size class.

For each size class lg (sizes with floor(log2 n) == lg) it times the Pippenger
pipeline at n = 2^(lg + 1/2) over a sweep of window widths c, then of the
accumulation chunk length, and keeps the fastest.  The table is persisted as
JSON under cusnarks_amd/tuning/ and installed into a context with
csk_msm_set_tuning; single calls and batch plans then use it whenever the
caller gives no explicit setting.  The reference fixes its launch geometry per
multiexp batch instead (get_shfl_blockD, cusnarks
src/python/cuda_wrapper.py:1570-1587).

    python -m cusnarks_amd.tune --curve g1 --lg 10:25 --out cusnarks_amd/tuning/msm_g1.json
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import statistics
import time

from . import _capi, inputs
from ._capi import MsmCfg, P, check, lib
from .device import DeviceBuffer

TUNING_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning")
G1W, G2W = 16, 32


def tuning_path(g2: bool) -> str:
    return os.path.join(TUNING_DIR, "msm_g2.json" if g2 else "msm_g1.json")


def table_from_json(doc: dict) -> list[MsmCfg]:
    """64-entry csk_msm_cfg table (index = lg) from a tuning document."""
    tab = [MsmCfg(0, 0, 0, 0) for _ in range(64)]
    for lg, e in doc.get("classes", {}).items():
        lg = int(lg)
        if 0 <= lg < 64:
            tab[lg] = MsmCfg(int(e.get("window_bits", 0)), int(e.get("chunk", 0)), int(e.get("reduce_seg", 0)), 0)
    return tab


def install(ctx, g2: bool, doc: dict) -> None:
    tab = (MsmCfg * 64)(*table_from_json(doc))
    check(lib().csk_msm_set_tuning(ctx.ptr, int(g2), ctypes.cast(tab, P), 64), "csk_msm_set_tuning")


def apply(ctx) -> list[str]:
    """Install the persisted tables (if any) into ctx; returns the files used."""
    used = []
    for g2 in (False, True):
        path = tuning_path(g2)
        if os.path.exists(path):
            with open(path) as fh:
                install(ctx, g2, json.load(fh))
            used.append(path)
    return used


def default_chunk(n: int, c: int) -> int:
    """The library's built-in chunk length for one multiexp of n points."""
    return max(8, min(256, (n * (254 // c + 1)) >> 18))


def _time_once(ctx, g2, ds, dp, n, out, cfg) -> float:
    fn = lib().csk_msm_g2 if g2 else lib().csk_msm_g1
    check(fn(ctx.ptr, ds.ptr, dp.ptr, n, out.ptr, ctypes.byref(cfg), None), "csk_msm")
    ms = (ctypes.c_float * 4)()
    check(lib().csk_last_stage_times(ctx.ptr, ms, 4), "csk_last_stage_times")
    return float(ms[3])


def _measure(ctx, g2, ds, dp, n, out, cfg, reps) -> float:
    _time_once(ctx, g2, ds, dp, n, out, cfg)  # warm (workspace growth, caches)
    return statistics.median(_time_once(ctx, g2, ds, dp, n, out, cfg) for _ in range(reps))


def autotune(g2: bool, lgs, ctx=None, reps: int = 3, log=print) -> dict:
    """Sweep c, then the chunk length, for every class in lgs; returns the tuning document."""
    from .msm import fixed_base
    ctx = ctx or _capi.Context(0)
    check(lib().csk_ctx_set_timing(ctx.ptr, 1), "csk_ctx_set_timing")
    lgs = sorted(int(x) for x in lgs)
    nmax = int(round(2 ** (lgs[-1] + 0.5)))
    t0 = time.time()
    k = DeviceBuffer.from_array(ctx, inputs.field_elements(9001, nmax, nonzero=True))
    pts = DeviceBuffer(ctx, nmax * (G2W if g2 else G1W) * 4)
    fixed_base(k, nmax, pts, g2=g2, ctx=ctx)
    check(lib().csk_stream_sync(ctx.ptr, None), "sync")
    k.close()
    sc = DeviceBuffer.from_array(ctx, inputs.field_elements(9002, nmax))
    out = DeviceBuffer(ctx, 4 * (G2W if g2 else G1W))
    log("tune %s: %d points ready in %.1f s" % ("g2" if g2 else "g1", nmax, time.time() - t0))
    cmax = 18 if g2 else 20
    classes = {}
    for lg in lgs:
        n = int(round(2 ** (lg + 0.5)))
        sweep = {}
        for c in range(max(4, lg - 10), min(lg + 1, cmax) + 1):
            sweep[c] = _measure(ctx, g2, sc, pts, n, out, MsmCfg(c, 0, 0, 0), reps)
        c = min(sweep, key=sweep.get)
        l0 = default_chunk(n, c)
        chunks = {}
        for L in sorted({max(4, l0 // 2), l0, min(1024, 2 * l0)}):
            chunks[L] = _measure(ctx, g2, sc, pts, n, out, MsmCfg(c, L, 0, 0), reps)
        L = min(chunks, key=chunks.get)
        classes[str(lg)] = {"window_bits": c, "chunk": L if L != l0 else 0, "reduce_seg": 0,
                            "n": n, "ms": round(chunks[L], 4),
                            "sweep_c_ms": {str(a): round(b, 4) for a, b in sweep.items()},
                            "sweep_chunk_ms": {str(a): round(b, 4) for a, b in chunks.items()}}
        log("  lg %2d n %9d: c=%d chunk=%d %.3f ms   (c sweep %s)" % (
            lg, n, c, L, chunks[L], " ".join("%d:%.2f" % kv for kv in sweep.items())))
    for b in (pts, sc, out):
        b.close()
    check(lib().csk_ctx_set_timing(ctx.ptr, 0), "csk_ctx_set_timing")
    return {"curve": "g2" if g2 else "g1", "device": "gfx950", "reps": reps,
            "method": "median device time of the full pipeline (HIP events); c swept, then chunk",
            "classes": classes}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--curve", choices=("g1", "g2"), default="g1")
    ap.add_argument("--lg", default="10:25", help="classes lo:hi (hi exclusive)")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    lo, hi = (int(x) for x in a.lg.split(":"))
    g2 = a.curve == "g2"
    doc = autotune(g2, range(lo, hi), reps=a.reps, log=lambda s: print(s, flush=True))
    path = a.out or tuning_path(g2)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as fh:
        json.dump(doc, fh, indent=1)
    print("wrote", path, flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
