#!/usr/bin/env python3
"""G1 multi-exponentiation throughput on one MI355X (BN254, N = 2^24).

One step = one full Pippenger multiexp (digits, sort, bucket accumulation,
reductions, final affine point) over 2^24 device-resident affine Montgomery
points k_i*G (k_i uniform in [1, r), seed 0) with scalars uniform in [0, r)
(seed 1).  Prints one JSON line.

  python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--lg 24] [--cpu-seconds 10]

The roofline object describes the dominant kernel (bucket accumulation,
k_accumulate29, which computes on 29-bit limbs): achieved = algorithmic
MAC32 per launch (entries x 11 Fp products x 136 MAC32, the mixed-addition
accounting of the multiexp cost model on 32-bit limbs) divided by its mean
device time, read from HIP events recorded on its launch stream.
The peak is the measured v_mad_u64_u32 rate of this GPU (integer VALU; the
work is exact modular arithmetic, no MFMA).  cpu_baseline times the CPU
oracle (a C restatement of the reference's ec_jacreduce_server_h) on a
bounded sample of the same workload on this host; cpu_baseline.more adds the
2^20 multiexp, the H step at m = 2^21 (the threaded 4-step restatement of
ntt_interpolandmul_server_h) and one 2^20-constraint Groth16 proof (the
oracle's routines composed as the reference's CPU prover), each beside the
GPU's time on the same inputs and checked equal.

single_calls: the plain single-call kernels (Fr NTT + INTT at 2^20, G1 multiexp
at 2^20, G2 multiexp at 2^18), each checked (round trip, closed form).

prove: Groth16 proofs of a 2^20-constraint, 2^20-wire circuit (m = 2^21):
the plain prover, the overlapped prover, and both again with the key's
fixed-base tables built (--no-prove skips it; --prove-only LG for others).

served: the proof server against the plain per-request prover over circuits
of 2^14 .. 2^24 constraints (uniform mix: capacity, then p50/p99 at half the
plain capacity; mix weighted by 1/size: capacity), 60 requests per run
(--serve: 100 requests and three loads; --no-serve skips it).
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
TRAFFIC_FILE = os.path.join("profiles", "msm_g1_2e24_traffic.json")
VALU_FILE = os.path.join("profiles", "msm_g1_2e24_valu.json")


def measured_traffic(n, c, name=TRAFFIC_FILE):
    """A committed rocprofv3 counter summary of the bucket accumulation (HBM
    bytes, or VALU counters), if it was measured on this workload (same n and c)."""
    path = os.path.join(ROOT, name)
    try:
        with open(path) as fh:
            d = json.load(fh)
    except (OSError, ValueError):
        return None
    return d if d.get("n") == n and d.get("window_bits") == c else None


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
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-prove", action="store_true",
                    help="skip the 2^20-constraint Groth16 prove leg (plain, overlapped, fixed-base tables)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--no-single", action="store_true",
                    help="skip the single-call legs (NTT + INTT 2^20, G1 2^20, G2 2^18)")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the proof-serving comparison (Poisson load over several circuits)")
    ap.add_argument("--serve", action="store_true",
                    help="full serving comparison: three offered loads and --serve-requests per run "
                         "(the default run serves 60 requests per run at one load)")
    ap.add_argument("--serve-lgs", default="14,16,18,20,22,24")
    ap.add_argument("--serve-requests", type=int, default=100)
    ap.add_argument("--serve-budget-gb", type=float, default=96.0)
    ap.add_argument("--serve-only", action="store_true")
    ap.add_argument("--g2", action="store_true", help="--table-only: G2 points")
    ap.add_argument("--table-widths", default="0,15,16,17,20",
                    help="--table-only: window widths to build tables of (0 = the library's default)")
    ap.add_argument("--table-only", type=int, default=0, metavar="LG",
                    help="time the fixed-base table multiexp against the plain one at 2^LG G1 points")
    ap.add_argument("--h-only", type=int, default=0, metavar="LG",
                    help="time the H step (csk_interpolandmul) at m = 2^LG and print one JSON line")
    ap.add_argument("--prove-only", type=int, default=0, metavar="LG",
                    help="time full Groth16 proofs of a synthetic circuit with 2^LG constraints and wires")
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
    if args.table_only:
        widths = tuple(int(x) for x in args.table_widths.split(","))
        print(json.dumps(table_bench(ctx, args.table_only, max(args.steps, 1), widths=widths, g2=args.g2)),
              flush=True)
        return 0
    if args.h_only:
        print(json.dumps(h_bench(ctx, args.h_only, max(args.steps, 1))), flush=True)
        return 0
    if args.prove_only:
        print(json.dumps(prove_bench(args.prove_only, max(args.steps, 1))), flush=True)
        return 0
    if args.serve_only:
        print(json.dumps(serve_bench([int(x) for x in args.serve_lgs.split(",")], args.serve_requests,
                                     args.serve_budget_gb)), flush=True)
        return 0

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
    copy_rate, _ = valu_probe(3, ctx=ctx)  # 1 GiB device copy: HBM reference + counter calibration

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
        "kernel": "k_accumulate29<Fq>",
        "kernel_ms": float(stage[1]),
        "algorithmic_ops_per_launch": acc_mac,
        "op": "MAC32 (32x32+64 multiply-add); 1 Fq product = 136, mixed add = 11 products "
              "(cost-model units; the kernel itself runs on 29-bit limbs)",
        "peak_measured_mad64": mad_rate / 1e12,
        "peak_issue": ISSUE_PEAK / 1e12,
        "mulmont_per_s": mul_rate,
        "hbm_copy_GBps": copy_rate / 1e9,
        "traffic_note": "HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel, "
                        "scaled by the counters' reading for the 1 GiB copy kernel (see profiles/)",
    }
    tr = measured_traffic(n, c)
    if tr is not None:
        roofline["traffic"] = tr["traffic_bytes"]
        roofline["traffic_source"] = TRAFFIC_FILE
    roofline["traffic_algorithmic_bytes"] = entries * 68
    va = measured_traffic(n, c, VALU_FILE)
    if va is not None:
        roofline["valu_busy_pct"] = va["valu_busy_pct"]
        roofline["valu_utilization_pct"] = va["valu_utilization_pct"]
        roofline["valu_instr_per_entry"] = va["valu_instr_per_entry_lane"]
        roofline["valu_source"] = VALU_FILE

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ctx, d_s, d_p, args.cpu_sample_lg, args.cpu_seconds)
        cpu["more"] = [cpu_msm_2e20(ctx, d_s, d_p), cpu_h_step(ctx, 21)]
    for b in (d_s, d_p):
        b.close()

    batch = None
    if not args.no_batch:
        batch = batch_bench(ctx, args.batch_steps)
    single = None if args.no_single else single_calls(ctx)
    prove = None
    if not args.no_prove:
        prove = prove_bench(20, 3, cpu=not args.no_cpu)
        if cpu is not None:
            cpu["more"].append(prove["cpu"])
    served = None
    if args.serve:
        served = serve_bench([int(x) for x in args.serve_lgs.split(",")], args.serve_requests,
                             args.serve_budget_gb)
    elif not args.no_serve:
        served = serve_bench([int(x) for x in args.serve_lgs.split(",")], 60, args.serve_budget_gb,
                             loads=(0.5,))

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
            "stages_ms": {"digits_sort_to29": float(stage[0]), "accumulate": float(stage[1]),
                          "from29_fixup_reduce_final": float(stage[2]), "total": float(stage[3])},
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batch,
        "single_calls": single,
        "prove": prove,
        "served": served,
    }
    print(json.dumps(line), flush=True)
    return 0


def batch_bench(ctx, steps, count=64, seed=7):
    """64 independent G1 multiexps, sizes round(2^u) with u ~ U[10, 22] (seed 7);
    member j: points k_i*G (k_i ~ U[1,r) seed 7000+2j), scalars ~ U[0,r) (seed
    7001+2j).  Times one batched run (graph replay) against the same 64
    multiexps as back-to-back single calls on one stream; results must agree."""
    import numpy as np
    from cusnarks_amd import _capi, inputs
    from cusnarks_amd.batch import MsmBatch
    from cusnarks_amd.device import DeviceBuffer
    from cusnarks_amd.msm import fixed_base, msm_device

    L = _capi.lib()

    def sync():
        _capi.check(L.csk_stream_sync(ctx.ptr, None), "sync")

    sizes = inputs.log_uniform_sizes(seed, count, 10, 22)
    pts, scl = [], []
    for j, n in enumerate(sizes):
        k = DeviceBuffer.from_array(ctx, inputs.field_elements(1000 * seed + 2 * j, n, nonzero=True))
        p = DeviceBuffer(ctx, n * 64)
        fixed_base(k, n, p, ctx=ctx)
        sync()
        k.close()
        pts.append(p)
        scl.append(DeviceBuffer.from_array(ctx, inputs.field_elements(1000 * seed + 2 * j + 1, n)))
    seq_out = DeviceBuffer(ctx, count * 64)
    bat_out = DeviceBuffer(ctx, count * 64)

    def sequential():
        for j, n in enumerate(sizes):
            msm_device(scl[j], pts[j], n, seq_out.ptr + 64 * j, ctx=ctx)

    batch = MsmBatch(sizes, ctx=ctx)
    outs = [bat_out.ptr + 64 * j for j in range(count)]

    def batched():
        batch.run(scl, pts, outs, use_graph=True)

    res = {}
    for name, fn in (("sequential", sequential), ("batched", batched)):
        fn()  # warm: workspace growth / graph capture
        sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        sync()
        res[name] = (time.perf_counter() - t0) / steps * 1e3
    same = bool(np.array_equal(seq_out.download((count, 16)), bat_out.download((count, 16))))
    total = sum(sizes)
    info = batch.info()
    batch.close()
    for b in pts + scl + [seq_out, bat_out]:
        b.close()
    return {
        "workload": "64 G1 multiexps, n=round(2^u), u~U[10,22] (seed 7), one batch (graph replay) "
                    "vs 64 back-to-back single calls",
        "points": total,
        "sizes_lg_min_max": [round(float(np.log2(min(sizes))), 2), round(float(np.log2(max(sizes))), 2)],
        "sequential_ms": res["sequential"],
        "batched_ms": res["batched"],
        "speedup": res["sequential"] / res["batched"],
        "batched_points_per_s": total / (res["batched"] * 1e-3),
        "sequential_points_per_s": total / (res["sequential"] * 1e-3),
        "results_equal": same,
        "groups": [{"c": g[0], "members": g[1], "chunk": g[2], "entries": g[3]} for g in info["groups"]],
        "device_bytes": info["bytes"],
        "steps": steps,
    }


def table_bench(ctx, lg, steps, widths=(0, 15, 16, 17, 20), g2=False):
    """G1 (or G2) multiexp over 2^lg points k_i*G (seed 0) with scalars
    ~U[0,r) (seed 1), plain (csk_msm_g1/g2) against fixed-base tables of
    several window widths (table build timed apart); every result compared
    with the plain one.  Mean device time over `steps` calls, HIP events on
    the null stream."""
    import numpy as np
    from cusnarks_amd import _capi, inputs
    from cusnarks_amd.device import DeviceBuffer, Event
    from cusnarks_amd.msm import MsmTable, fixed_base, msm_device

    n, nw = 1 << lg, 32 if g2 else 16
    dk = DeviceBuffer.from_array(ctx, inputs.field_elements(0, n, inputs.R, nonzero=True))
    dp = DeviceBuffer(ctx, n * nw * 4)
    fixed_base(dk, n, dp, g2=g2, ctx=ctx)
    _capi.check(_capi.lib().csk_stream_sync(ctx.ptr, None), "sync")
    dk.close()
    ds = DeviceBuffer.from_array(ctx, inputs.field_elements(1, n))
    ref, out = DeviceBuffer(ctx, nw * 4), DeviceBuffer(ctx, nw * 4)

    def timed(fn):
        fn()
        e0, e1 = Event(ctx), Event(ctx)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        e1.sync()
        return e0.elapsed_ms(e1) / steps

    plain_ms = timed(lambda: msm_device(ds, dp, n, ref, g2=g2, ctx=ctx))
    want = ref.download((nw,))
    rows = []
    for c in widths:
        t0 = time.perf_counter()
        t = MsmTable(dp, n, g2=g2, window_bits=c, ctx=ctx)
        build_s = time.perf_counter() - t0
        ms = timed(lambda: t.run(ds, out))
        info = t.info()
        rows.append({"window_bits": info["window_bits"], "windows": info["windows"], "ms": ms,
                     "points_per_s": n / (ms * 1e-3), "build_s": build_s, "table_gb": info["table_bytes"] / 1e9,
                     "equal": bool(np.array_equal(out.download((nw,)), want))})
        t.close()
    for b in (dp, ds, ref, out):
        b.close()
    return {"workload": "%s multiexp 2^%d: plain vs fixed-base tables" % ("G2" if g2 else "G1", lg),
            "plain_ms": plain_ms,
            "plain_points_per_s": n / (plain_ms * 1e-3), "tables": rows}


def single_calls(ctx, steps=10):
    """The plain single-call kernels of the BN128 single-call workload, each on
    device-resident inputs, mean device time over `steps` calls (HIP events on
    the null stream) and checked: an Fr NTT then INTT of 2^20 coefficients
    (uniform mod r, seed 41; the round trip must return the input), a G1
    multiexp of 2^20 points and a G2 multiexp of 2^18 points (points k_i*G with
    k_i ~U[1,r) and scalars ~U[0,r), seeds 42/43 and 44/45; each result must
    equal (sum k_i s_i mod r) * G, computed from the same seeds on the host)."""
    import numpy as np
    from cusnarks_amd import _capi, inputs
    from cusnarks_amd.device import DeviceBuffer, Event
    from cusnarks_amd.msm import fixed_base, msm_device
    from cusnarks_amd.poly import ntt_device

    def timed(fn):
        fn()
        e0, e1 = Event(ctx), Event(ctx)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        e1.sync()
        return e0.elapsed_ms(e1) / steps

    out = {}
    n = 1 << 20
    a = inputs.field_elements(41, n)
    da, dt, db = DeviceBuffer.from_array(ctx, a), DeviceBuffer(ctx, n * 32), DeviceBuffer(ctx, n * 32)

    def rt():
        ntt_device(da, dt, 20, ctx=ctx)
        ntt_device(dt, db, 20, inverse=True, fmt=1, ctx=ctx)
    ms = timed(rt)
    out["ntt_intt_2e20"] = {"ms": ms, "roundtrip_equal": bool(np.array_equal(db.download((n, 8)), a))}
    for b in (da, dt, db):
        b.close()
    for key, lg, g2, seeds in (("msm_g1_2e20", 20, False, (42, 43)), ("msm_g2_2e18", 18, True, (44, 45))):
        n = 1 << lg
        nw = 32 if g2 else 16
        k = inputs.field_elements(seeds[0], n, inputs.R, nonzero=True)
        sc = inputs.field_elements(seeds[1], n)
        dk, dp = DeviceBuffer.from_array(ctx, k), DeviceBuffer(ctx, n * nw * 4)
        fixed_base(dk, n, dp, g2=g2, ctx=ctx)
        ds, do = DeviceBuffer.from_array(ctx, sc), DeviceBuffer(ctx, nw * 4)
        ms = timed(lambda: msm_device(ds, dp, n, do, g2=g2, ctx=ctx))
        got = do.download((nw,))
        dot = sum(x * y for x, y in zip(inputs.to_ints(k), inputs.to_ints(sc))) % inputs.R
        one = DeviceBuffer.from_array(ctx, inputs.from_ints([dot]))
        ref = DeviceBuffer(ctx, nw * 4)
        fixed_base(one, 1, ref, g2=g2, ctx=ctx)
        out[key] = {"ms": ms, "points_per_s": n / (ms * 1e-3),
                    "closed_form_equal": bool(np.array_equal(got, ref.download((nw,))))}
        for b in (dk, dp, ds, do, one, ref):
            b.close()
    out["workload"] = ("plain single calls: Fr NTT + INTT 2^20, G1 multiexp 2^20, G2 multiexp 2^18 "
                       "(device-resident inputs, mean of %d calls)" % steps)
    return out


def h_bench(ctx, log2m, steps, seed=31):
    """The H step alone: polH from m = 2^log2m evaluations each of A and B
    (2 INTT_m + 2 NTT_m + pointwise + INTT_2m), device-resident inputs,
    mean device time over `steps` calls (HIP events on the launch stream),
    with the Fr products it needs at the radix-2 count as its algorithmic work."""
    from cusnarks_amd import inputs
    from cusnarks_amd.device import DeviceBuffer, Event
    from cusnarks_amd.msm import valu_probe
    from cusnarks_amd.poly import interpolandmul_device

    m = 1 << log2m
    dA = DeviceBuffer.from_array(ctx, inputs.field_elements(seed, m))
    dB = DeviceBuffer.from_array(ctx, inputs.field_elements(seed + 1, m))
    dO = DeviceBuffer(ctx, 2 * m * 32)
    interpolandmul_device(dA, dB, log2m, dO, ctx=ctx)
    e0, e1 = Event(ctx), Event(ctx)
    e0.record()
    for _ in range(steps):
        interpolandmul_device(dA, dB, log2m, dO, ctx=ctx)
    e1.record()
    e1.sync()
    ms = e0.elapsed_ms(e1) / steps
    # butterflies: 4 transforms of m (m/2 log2 m each) + one of 2m; finalize scaling/coset/pointwise
    prods = 4 * (m // 2) * log2m + m * (log2m + 1) + 4 * m + m + m + 2 * m
    mul_rate, _ = valu_probe(1, ctx=ctx)
    for b in (dA, dB, dO):
        b.close()
    return {"workload": "H step, m = 2^%d (2 INTT_m + 2 NTT_m + pointwise + INTT_2m), inputs ~U[0,r) seeds %d/%d"
                        % (log2m, seed, seed + 1),
            "ms": ms, "steps": steps, "fr_products": prods, "fr_products_per_s": prods / (ms * 1e-3),
            "mulmont_per_s_probe": mul_rate, "frac_of_probe": prods / (ms * 1e-3) / mul_rate}


def prove_bench(lg, steps, seed=0, cpu=False):
    """Full Groth16 proofs (configs as the serving circuits: 2^lg constraints
    and wires, 3 nonzeros per row, nPublic 1, seeded keys and witness): one
    warm proof, then `steps` timed proofs, the last with per-stage times
    (host wall time between stream synchronisations).  With cpu, the same
    proof is also computed once by the CPU oracle (cpu_prove) and compared."""
    import numpy as np
    from cusnarks_amd.prover import GrothProver, ProvingKey
    from cusnarks_amd.synth import Circuit, KeyScalars
    t0 = time.perf_counter()
    circ = Circuit(1 << lg, 1 << lg, 1, seed=seed)
    pk = ProvingKey(circ, KeyScalars(circ, seed=seed))
    setup = time.perf_counter() - t0
    pr = GrothProver(pk)
    pr.prove(circ.witness, 1234567, 7654321)
    times = []
    for i in range(steps):
        t = time.perf_counter()
        pr.prove(circ.witness, 1234567 + i, 7654321, timed=(i == steps - 1))
        times.append((time.perf_counter() - t) * 1e3)
    from cusnarks_amd.prover import OverlappedProver
    ov = OverlappedProver(pk)
    same = all(np.array_equal(a, b) for a, b in zip(ov.prove(circ.witness, 5, 7), pr.prove(circ.witness, 5, 7)))
    ov_times = []
    for i in range(steps):
        t = time.perf_counter()
        ov.prove(circ.witness, 1234567 + i, 7654321)
        ov_times.append((time.perf_counter() - t) * 1e3)
    plain_stages = {k: round(v * 1e3, 3) for k, v in pr.times.items()}
    # the same provers with the five multiexps run from fixed-base tables
    want = pr.prove(circ.witness, 5, 7)
    t0 = time.perf_counter()
    tbytes = pk.build_tables()
    build_s = time.perf_counter() - t0
    tp, tov = GrothProver(pk), OverlappedProver(pk)
    t_same = all(np.array_equal(a, b) for p_ in (tp, tov) for a, b in zip(p_.prove(circ.witness, 5, 7), want))
    t_times, t_ov_times = [], []
    for i in range(steps):
        t = time.perf_counter()
        tp.prove(circ.witness, 1234567 + i, 7654321, timed=(i == steps - 1))
        t_times.append((time.perf_counter() - t) * 1e3)
    for i in range(steps):
        t = time.perf_counter()
        tov.prove(circ.witness, 1234567 + i, 7654321)
        t_ov_times.append((time.perf_counter() - t) * 1e3)
    tables = {"build_s": build_s, "table_gb": tbytes / 1e9,
              "window_bits": {k: t.info()["window_bits"] for k, t in pk.tables.items()},
              "proof_ms": t_times, "proofs_per_s": 1e3 / min(t_times),
              "overlapped_proof_ms": t_ov_times, "overlapped_proofs_per_s": 1e3 / min(t_ov_times),
              "equals_plain": bool(t_same), "stages_ms": {k: round(v * 1e3, 3) for k, v in tp.times.items()}}
    # the witness in page-locked host memory (as keyio.read_witness(..., pinned=True) returns it)
    from cusnarks_amd.device import pinned_empty
    wp = pinned_empty(circ.witness.shape, np.uint32, ctx=pk.ctx)
    wp[...] = circ.witness
    p_same = all(np.array_equal(a, b) for p_ in (tp, tov) for a, b in zip(p_.prove(wp, 5, 7), want))
    p_times, p_ov_times = [], []
    for i in range(steps):
        t = time.perf_counter()
        tp.prove(wp, 1234567 + i, 7654321, timed=(i == steps - 1))
        p_times.append((time.perf_counter() - t) * 1e3)
    for i in range(steps):
        t = time.perf_counter()
        tov.prove(wp, 1234567 + i, 7654321)
        p_ov_times.append((time.perf_counter() - t) * 1e3)
    tables["pinned_witness"] = {"proof_ms": p_times, "overlapped_proof_ms": p_ov_times,
                                "overlapped_proofs_per_s": 1e3 / min(p_ov_times), "equals_plain": bool(p_same),
                                "upload_ms": round(tp.times["upload"] * 1e3, 3)}
    del wp
    pk.close_tables()
    out = {"workload": "Groth16 prove, synthetic R1CS 2^%d constraints x 2^%d wires (m = 2^%d)" % (lg, lg, circ.log2m),
           "proof_ms": times, "proofs_per_s": 1e3 / min(times),
           "overlapped_proof_ms": ov_times, "overlapped_proofs_per_s": 1e3 / min(ov_times),
           "overlapped_equals_plain": bool(same),
           "stages_ms": plain_stages, "setup_s": setup, "tables": tables}
    if cpu:
        out["cpu"] = cpu_prove(lg, circ, pk, want, 5, 7, gpu_ms=min(times), gpu_best_ms=min(t_ov_times))
    return out


def cpu_prove(lg, circ, pk, want, r, s, gpu_ms, gpu_best_ms):
    """One Groth16 proof on the CPU from the oracle's restatements of the
    reference's CPU path (the prover's composition, groth_prover.py:575-790):
    QAP evaluation (mpoly_eval), the threaded H step
    (ntt_interpolandmul_server_h), the five multiexps and the final
    4-term combination with the Straus multiexp server (ec_jacreduce_server_h),
    on the key points the GPU prover used; compared with the GPU proof `want`
    for the same r, s.  Key download not timed."""
    import numpy as np
    import oracle
    from cusnarks_amd import inputs

    nv, m, npub = circ.n_vars, circ.m, circ.n_public
    A, B1 = pk.A.download((nv + 2, 16)), pk.B1.download((nv + 2, 16))
    B2, C = pk.B2.download((nv + 2, 32)), pk.C.download((nv, 16))
    hE = pk.hExps.download((m, 16))
    w = circ.witness
    one = inputs.from_ints([1])
    th = _cpu_threads()
    st = {}
    t0 = time.perf_counter()
    t = t0
    pA = oracle.mpoly_eval_csr(circ.A.rowptr, circ.A.col, circ.A.val, w, m)
    pB = oracle.mpoly_eval_csr(circ.B.rowptr, circ.B.col, circ.B.val, w, m)
    st["qap"], t = (time.perf_counter() - t) * 1e3, time.perf_counter()
    polH = oracle.interpolandmul(pA, pB, circ.log2m, threads=th)
    st["h"], t = (time.perf_counter() - t) * 1e3, time.perf_counter()
    rs_ = inputs.from_ints([r, s, (-r * s) % inputs.R])
    pa, _ = oracle.ec_jacreduce(np.concatenate([w, one, rs_[0:1]]), A, max_threads=th)
    pb1, _ = oracle.ec_jacreduce(np.concatenate([w, one, rs_[1:2]]), B1, max_threads=th)
    pcp, _ = oracle.ec_jacreduce(w[npub + 1:], C[npub + 1:], max_threads=th)
    pc2, _ = oracle.ec_jacreduce(np.concatenate([polH[m:2 * m - 1], rs_[2:3]]), hE, max_threads=th)
    st["msm_g1"], t = (time.perf_counter() - t) * 1e3, time.perf_counter()
    pb, _ = oracle.ec_jacreduce(np.concatenate([w, one, rs_[1:2]]), B2, ec2=True, max_threads=th)
    st["msm_g2"], t = (time.perf_counter() - t) * 1e3, time.perf_counter()
    pc, _ = oracle.ec_jacreduce(np.concatenate([one, one, rs_[1:2], rs_[0:1]]), np.stack([pc2, pcp, pa, pb1]))
    cpu_ms = (time.perf_counter() - t0) * 1e3
    same = all(np.array_equal(a, b) for a, b in zip((pa, pb, pc), want))
    return {"workload": "Groth16 prove, 2^%d constraints x 2^%d wires (m = 2^%d)" % (lg, lg, circ.log2m),
            "value": 1e3 / cpu_ms, "unit": "proofs/s", "cores": th, "kind": "port", "cpu_ms": cpu_ms,
            "cpu_stages_ms": {k: round(v, 1) for k, v in st.items()}, "gpu_ms": gpu_ms,
            "gpu_best_ms": gpu_best_ms, "gpu_over_cpu": cpu_ms / gpu_ms, "sample_parity": bool(same),
            "sample": "one proof: C restatements of mpoly_eval, ntt_interpolandmul_server_h and "
                      "ec_jacreduce_server_h (Straus order 6) composed as the reference's CPU prover; "
                      "gpu_ms = plain GrothProver, gpu_best_ms = OverlappedProver with fixed-base tables"}


def serve_bench(lgs, n_req, budget_gb, loads=(0.5, 0.8, 1.1), seed=11):
    """Groth16 proof requests over synthetic circuits of 2^lg constraints and
    2^lg wires (3 nonzeros per row, nPublic = 1); witnesses and keys seeded
    per circuit.  Uniform mix: the capacity (a burst of n_req requests at
    t=0) of the continuous-batching server and of two plain per-request
    baselines (keys resident in HBM; keys in host memory, copied per
    multiexp call as in the reference's flow), then all three under Poisson
    arrivals at 0.5x, 0.8x and 1.1x the HBM-resident plain capacity.  A second
    mix draws circuits with probability proportional to 1/size (every circuit
    gets the same share of the work) and measures capacity.  Every served
    proof is compared with the plain proof of the same request."""
    import dataclasses
    import numpy as np
    from cusnarks_amd.prover import ProvingKey
    from cusnarks_amd.serve import PlainServer, ProofServer, Request, poisson_requests, summarize
    from cusnarks_amd.synth import Circuit, KeyScalars

    def log(msg):
        print("serve: " + msg, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    keys, wit = {}, {}
    for lg in lgs:
        name = "2^%d" % lg
        circ = Circuit(1 << lg, 1 << lg, 1, seed=100 + lg)
        keys[name] = ProvingKey(circ, KeyScalars(circ, seed=100 + lg))
        wit[name] = circ.witness
        del circ
        log("key %s ready (%.1f s)" % (name, time.perf_counter() - t0))
    setup_s = time.perf_counter() - t0
    servers = {"served": ProofServer(keys, budget_bytes=int(budget_gb * (1 << 30)), max_inflight=64),
               "plain": PlainServer(keys), "plain_host_keys": PlainServer(keys, host_keys=True)}

    def fresh(reqs):
        return [dataclasses.replace(q, t_admit=None, t_done=None, proof=None, error=None, slot=None, ev_h=None)
                for q in reqs]

    same = [True]

    def run_all(reqs, names=("served", "plain", "plain_host_keys")):
        out, got = {}, {}
        for nm in names:
            rr = fresh(reqs)
            out[nm] = summarize(servers[nm].run(rr))
            got[nm] = rr
            log("%s: %s" % (nm, out[nm]))
        ref = got[names[1]]
        for nm in names:
            same[0] &= all(x.proof is not None and all(np.array_equal(u, v) for u, v in zip(x.proof, y.proof))
                           for x, y in zip(got[nm], ref))
        return out

    warm = [Request(i, name, wit[name], 3 + i, 5 + i) for i, name in enumerate(sorted(keys))]
    run_all(warm)
    log("warm")
    cap = run_all(poisson_requests(0, n_req, wit, seed=seed))
    base = cap["plain"]["proofs_per_s"]
    runs = []
    for i, f in enumerate(loads):
        # the same request sequence at every load; only the arrival times scale
        res = run_all(poisson_requests(f * base, n_req, wit, seed=seed + 1))
        runs.append({"offered_x_plain_capacity": f, "offered_proofs_per_s": f * base, **res})
    weights = {"2^%d" % lg: 2.0 ** -lg for lg in lgs}
    cap_w = run_all(poisson_requests(0, n_req, wit, seed=seed + 9, weights=weights))

    def ratios(c):
        return {"served_over_plain": c["served"]["proofs_per_s"] / c["plain"]["proofs_per_s"],
                "served_over_plain_host_keys": c["served"]["proofs_per_s"] / c["plain_host_keys"]["proofs_per_s"]}
    return {
        "workload": "Groth16 requests over circuits %s (2^lg constraints = wires, 3 nnz/row, nPublic 1), "
                    "Poisson arrivals" % ",".join("2^%d" % x for x in lgs),
        "requests_per_run": n_req,
        "hbm_budget_gb": budget_gb,
        "uniform_mix": {"capacity": {**cap, **ratios(cap)}, "loads": runs},
        "equal_work_mix": {"weights": "p(circuit) ~ 1/size", "capacity": {**cap_w, **ratios(cap_w)}},
        "proofs_equal_served_vs_plain": bool(same[0]),
        "server": servers["served"].stats(),
        "setup_s": setup_s,
    }


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


def cpu_msm_2e20(ctx, d_s, d_p, lg=20):
    """The CPU oracle's multiexp on the first 2^20 pairs of the workload (one
    run, ~5 s on 16 threads), beside the GPU multiexp of the same pairs."""
    import numpy as np
    import oracle
    from cusnarks_amd import _capi
    from cusnarks_amd.device import DeviceBuffer
    from cusnarks_amd.msm import msm_device

    m = 1 << lg
    sc = d_s.download((m, 8))
    pts = d_p.download((m, 16))
    out = DeviceBuffer(ctx, 64)
    msm_device(d_s, d_p, m, out, ctx=ctx)
    _capi.check(_capi.lib().csk_stream_sync(ctx.ptr, None), "sync")
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        msm_device(d_s, d_p, m, out, ctx=ctx)
    _capi.check(_capi.lib().csk_stream_sync(ctx.ptr, None), "sync")
    gpu_ms = (time.perf_counter() - t0) / reps * 1e3
    t0 = time.perf_counter()
    ref, threads = oracle.ec_jacreduce(sc, pts, order=6, max_threads=_cpu_threads())
    cpu_s = time.perf_counter() - t0
    same = bool(np.array_equal(ref, out.download((16,))))
    out.close()
    return {"workload": "G1 multiexp, first 2^%d pairs of the workload" % lg, "value": m / cpu_s,
            "unit": "points/s", "cores": threads, "kind": "port", "cpu_ms": cpu_s * 1e3, "gpu_ms": gpu_ms,
            "gpu_over_cpu": cpu_s * 1e3 / gpu_ms, "sample_parity": same,
            "sample": "one run of the C restatement of ec_jacreduce_server_h (Straus order 6)"}


def cpu_h_step(ctx, log2m, seed=31):
    """The H step (polH = A(x) B(x) from m evaluations each, m = 2^log2m) on
    the CPU, computed as the reference's threaded ntt_interpolandmul_server_h
    (4-step transforms, threads = min(nproc, 32) capped at 16), beside the
    GPU's csk_interpolandmul on the same inputs; results compared."""
    import numpy as np
    import oracle
    from cusnarks_amd import _capi, inputs
    from cusnarks_amd.device import DeviceBuffer, Event
    from cusnarks_amd.poly import interpolandmul_device

    m = 1 << log2m
    pA = inputs.field_elements(seed, m)
    pB = inputs.field_elements(seed + 1, m)
    dA, dB = DeviceBuffer.from_array(ctx, pA), DeviceBuffer.from_array(ctx, pB)
    dO = DeviceBuffer(ctx, 2 * m * 32)
    interpolandmul_device(dA, dB, log2m, dO, ctx=ctx)
    e0, e1 = Event(ctx), Event(ctx)
    reps = 10
    e0.record()
    for _ in range(reps):
        interpolandmul_device(dA, dB, log2m, dO, ctx=ctx)
    e1.record()
    e1.sync()
    gpu_ms = e0.elapsed_ms(e1) / reps
    got = dO.download((2 * m, 8))
    for b in (dA, dB, dO):
        b.close()
    threads = _cpu_threads()
    rts = oracle.roots(log2m + 1)
    t0 = time.perf_counter()
    ref = oracle.interpolandmul(pA, pB, log2m, threads=threads, rts=rts)
    cpu_s = time.perf_counter() - t0
    return {"workload": "H step (2 INTT_m + 2 NTT_m + pointwise + INTT_2m), m = 2^%d" % log2m,
            "value": 1.0 / cpu_s, "unit": "H steps/s", "cores": threads, "kind": "port",
            "cpu_ms": cpu_s * 1e3, "gpu_ms": gpu_ms, "gpu_over_cpu": cpu_s * 1e3 / gpu_ms,
            "sample_parity": bool(np.array_equal(ref, got)),
            "sample": "one run of the C restatement of ntt_interpolandmul_server_h (4-step transforms "
                      "split over the threads)"}


def _cpu_threads() -> int:
    return max(1, min(os.cpu_count() or 1, 32, 16))


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
