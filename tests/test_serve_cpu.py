"""Host logic of the proof server: admission by HBM budget, load generation
and latency summaries (no GPU)."""
import numpy as np

from cusnarks_amd.serve import Admission, Request, poisson_requests, summarize


def test_admission_budget_and_cap():
    a = Admission(budget_bytes=4000, max_inflight=3, shares=(0.25, 0.75))
    assert a.budget == [1000, 3000]
    assert a.never_fits(0, 600, 401) and not a.never_fits(0, 600, 400)
    assert not a.never_fits(1, 600, 401)
    assert a.fits(0, 300, 200)
    a.take(0, 300, 200)            # slot 300 allocated, lane-0 need 200
    assert a.projected(0) == 500
    assert a.fits(0, 0, 500)       # reuses an idle slot: 300 + 700
    assert not a.fits(0, 1, 500)
    a.take(0, 0, 500)
    assert not a.fits(0, 0, 1)     # lane 0 full: 300 + 700 + 1 > 1000
    assert a.fits(1, 1000, 2000)   # the other lane has its own share
    a.release(0, 500)
    a.cohort(0, 650)               # lane 0 workspace grew to 650 (high-water mark)
    assert a.projected(0) == 300 + 650
    assert a.projected(0, 0, 450) == 300 + 650        # within the high-water mark
    assert a.fits(0, 0, 500) and not a.fits(0, 0, 501)  # 300 + 200 + 500 = 1000
    a.take(0, 0, 10)
    assert a.inflight[0] == 2
    a.take(0, 0, 10)
    assert not a.fits(0, 0, 0)     # cap on proofs in flight
    a.free_slot(0, 300)
    assert a.slot_bytes[0] == 0


def test_admission_idle_lane_is_never_stalled_by_its_high_water_mark():
    a = Admission(budget_bytes=4000, max_inflight=3, shares=(0.25, 0.75))
    a.cohort(0, 5000)              # a cohort's workspace grew past the lane's share
    assert a.fits(0, 300, 200)     # nothing in flight: admitted anyway
    a.take(0, 300, 200)
    assert not a.fits(0, 0, 100)   # one in flight: the high-water mark counts again
    a.release(0, 200)
    assert a.fits(0, 0, 100)


def test_poisson_requests_rate_and_determinism():
    wit = {"a": np.zeros((4, 8), np.uint32), "b": np.ones((4, 8), np.uint32)}
    r1 = poisson_requests(50.0, 4000, wit, seed=2)
    r2 = poisson_requests(50.0, 4000, wit, seed=2)
    assert [(q.t_arrival, q.circuit, q.r, q.s) for q in r1] == [(q.t_arrival, q.circuit, q.r, q.s) for q in r2]
    gaps = np.diff([0.0] + [q.t_arrival for q in r1])
    assert abs(gaps.mean() - 1 / 50.0) < 0.002          # mean inter-arrival 1/rate
    assert abs(np.std(gaps) / gaps.mean() - 1) < 0.1     # exponential: cv 1
    frac = sum(q.circuit == "a" for q in r1) / len(r1)
    assert 0.45 < frac < 0.55
    burst = poisson_requests(0, 5, wit)
    assert all(q.t_arrival == 0 for q in burst)


def test_summarize_percentiles():
    reqs = [Request(i, "a", None, 1, 1, t_arrival=0.0) for i in range(100)]
    for i, q in enumerate(reqs):
        q.t_done = (i + 1) * 1e-3
    st = summarize(reqs)
    assert st["proofs"] == 100
    assert abs(st["p50_ms"] - 50.5) < 1e-6 and abs(st["p99_ms"] - 99.01) < 1e-6
    assert abs(st["proofs_per_s"] - 1000.0) < 1e-6
