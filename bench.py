#!/usr/bin/env python3
"""G1 multi-exponentiation throughput on one MI355X (BN254, N = 2^24).

One step = one full Pippenger multiexp (digits, sort, bucket accumulation,
reductions, final affine point) over 2^24 device-resident affine Montgomery
points k_i*G (k_i uniform in [1, r), seed 0) with scalars uniform in [0, r)
(seed 1).  Prints one JSON line.

  python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--lg 24] [--cpu-seconds 10]

The roofline object describes the dominant kernel (bucket accumulation):
achieved = algorithmic MAC32 per launch (entries x 11 Fp products x 136
MAC32, the mixed-addition accounting of the multiexp cost model) divided by
its mean device time, read from HIP events recorded on its launch stream.
The peak is the measured v_mad_u64_u32 rate of this GPU (integer VALU; the
work is exact modular arithmetic, no MFMA).  cpu_baseline times the CPU
oracle (a C restatement of the reference's ec_jacreduce_server_h) on a
bounded sample of the same workload on this host.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "G1 MSM points/s at 2^24; proofs/s + p99 latency, served vs plain"
ISSUE_PEAK = 256 * 4 * 16 * 2.4e9  # lane-ops/s at the 2.4 GHz issue rate


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--lg", type=int, default=24)
    ap.add_argument("--window", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-sample-lg", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested); run with --gpus 1" % args.gpus, file=sys.stderr)
        return 2

    import numpy as np
    from cusnarks_amd import _capi, inputs
    from cusnarks_amd.device import DeviceBuffer
    from cusnarks_amd.msm import fixed_base, msm_device, valu_probe

    ctx = _capi.default_context(0)
    L = _capi.lib()

    def sync():
        _capi.check(L.csk_stream_sync(ctx.ptr, None), "sync")

    n = 1 << args.lg
    k = inputs.field_elements(0, n, inputs.R, nonzero=True)
    sc = inputs.field_elements(1, n)
    d_k = DeviceBuffer.from_array(ctx, k)
    d_s = DeviceBuffer.from_array(ctx, sc)
    d_p = DeviceBuffer(ctx, n * 64)
    d_out = DeviceBuffer(ctx, 64)
    fixed_base(d_k, n, d_p, ctx=ctx)
    sync()
    d_k.close()
    del k

    mad_rate, _ = valu_probe(0, ctx=ctx)
    mul_rate, _ = valu_probe(1, ctx=ctx)

    cfg = {"window_bits": args.window} if args.window else {}
    for _ in range(args.warmup):
        msm_device(d_s, d_p, n, d_out, ctx=ctx, **cfg)
    sync()

    L.csk_ctx_set_timing(ctx.ptr, 1)
    stage = np.zeros(4)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        msm_device(d_s, d_p, n, d_out, ctx=ctx, **cfg)
        ms = (ctypes.c_float * 4)()
        _capi.check(L.csk_last_stage_times(ctx.ptr, ms, 4), "stage times")
        stage += np.array(ms[:4])
    sync()
    elapsed = time.perf_counter() - t0
    L.csk_ctx_set_timing(ctx.ptr, 0)
    stage /= args.steps
    plan = (ctypes.c_uint32 * 8)()
    L.csk_last_msm_plan(ctx.ptr, plan, 8)
    c, W = plan[0], plan[1]
    ms_step = elapsed / args.steps * 1e3
    value = n / (elapsed / args.steps)

    entries = n * W  # nonzero signed digits (zero digits occur with probability ~2^-c)
    acc_mac = entries * 11 * 136
    achieved = acc_mac / (stage[1] * 1e-3)
    peak = min(mad_rate, ISSUE_PEAK)
    roofline = {
        "bound": "valu",
        "achieved": achieved / 1e12,
        "peak": peak / 1e12,
        "unit": "TOP/s",
        "frac": achieved / peak,
        "traffic": None,
        "kernel": "k_accumulate<Fq>",
        "kernel_ms": float(stage[1]),
        "algorithmic_ops_per_launch": acc_mac,
        "op": "MAC32 (32x32+64 multiply-add); 1 Fq product = 136, mixed add = 11 products",
        "peak_measured_mad64": mad_rate / 1e12,
        "peak_issue": ISSUE_PEAK / 1e12,
        "mulmont_per_s": mul_rate,
    }

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ctx, d_s, d_p, args.cpu_sample_lg, args.cpu_seconds)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "BN254 G1 multiexp, N=2^%d affine Montgomery points k_i*G (k_i~U[1,r) seed 0), "
                        "scalars ~U[0,r) seed 1, Pippenger c=%d (%d windows)" % (args.lg, c, W),
            "n": n,
            "window_bits": c,
            "windows": W,
            "chunk": plan[3],
            "stages_ms": {"digits_sort": float(stage[0]), "accumulate": float(stage[1]),
                          "fixup_reduce_final": float(stage[2]), "total": float(stage[3])},
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def cpu_baseline(ctx, d_s, d_p, lg, seconds):
    """Time the CPU oracle on the first 2^lg (scalar, point) pairs of the
    workload; also checks the GPU multiexp on that sample bit for bit."""
    import numpy as np
    import oracle
    from cusnarks_amd.msm import msm_host

    m = 1 << lg
    sc = d_s.download((m, 8))
    pts = d_p.download((m, 16))
    gpu = msm_host(sc, pts, ctx=ctx)
    reps, t = 0, 0.0
    out, threads = None, 0
    while reps < 2 or t < seconds:
        t0 = time.perf_counter()
        out, threads = oracle.ec_jacreduce(sc, pts, order=6)
        t += time.perf_counter() - t0
        reps += 1
    return {
        "value": m * reps / t,
        "unit": "points/s",
        "cores": threads,
        "kind": "port",
        "sample": "first 2^%d points/scalars of the workload, %d reps; CPU oracle = C restatement of "
                  "ec_jacreduce_server_h (Straus order 6, threads=min(nproc,32)); nproc=%d; cpu=%s"
                  % (lg, reps, os.cpu_count() or 0, _cpu_model()),
        "sample_parity": bool(np.array_equal(out, gpu)),
    }


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    sys.exit(main())
