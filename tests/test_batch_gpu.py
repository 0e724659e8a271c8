"""Batched multi-exponentiation: every member's result is the same alone
(single call) and in the mix, and matches the oracle, bit for bit."""
import numpy as np
import pytest

import oracle as O
from cusnarks_amd import _capi, inputs
from cusnarks_amd.batch import MsmBatch
from cusnarks_amd.device import DeviceBuffer
from cusnarks_amd.msm import fixed_base, msm_device
from tests.gpu_util import g1_fixture

pytestmark = pytest.mark.gpu


class Member:
    """Device inputs of one multiexp: points k_i G (seeded), scalars (seeded)."""

    def __init__(self, ctx, seed, n, g2=False):
        self.n, self.g2 = n, g2
        self.w = 32 if g2 else 16
        m = max(n, 1)
        k = DeviceBuffer.from_array(ctx, inputs.field_elements(seed, m, nonzero=True))
        self.pts = DeviceBuffer(ctx, m * self.w * 4)
        fixed_base(k, m, self.pts, g2=g2, ctx=ctx)
        self.scl_host = inputs.field_elements(seed + 1, m)
        self.scl = DeviceBuffer.from_array(ctx, self.scl_host)
        self.out = DeviceBuffer(ctx, self.w * 4)

    def alone(self, ctx):
        o = DeviceBuffer(ctx, self.w * 4)
        msm_device(self.scl, self.pts, self.n, o, g2=self.g2, ctx=ctx)
        return o.download((self.w,))

    def oracle(self):
        pts = self.pts.download((max(self.n, 1), self.w))[: self.n]
        exp, _ = O.ec_jacreduce(self.scl_host[: self.n], pts, ec2=self.g2)
        return exp


def _run(batch, members, use_graph=True):
    batch.run([m.scl for m in members], [m.pts for m in members], [m.out for m in members],
              use_graph=use_graph)
    return [m.out.download((m.w,)) for m in members]


SIZES = [0, 1, 3, 1000, 4096, 4097, 20000, 70000, 5, 1 << 17, 2, 300]


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_g1_alone_vs_mix(use_graph):
    ctx = _capi.default_context()
    members = [Member(ctx, 1000 + 2 * i, n) for i, n in enumerate(SIZES)]
    batch = MsmBatch(SIZES, ctx=ctx)
    assert len(batch.info()["groups"]) >= 2  # several window widths in the mix
    got = _run(batch, members, use_graph)
    for m, g in zip(members, got):
        assert np.array_equal(g, m.alone(ctx)), m.n
    for m, g in zip(members, got):
        if m.n <= 20000:
            assert np.array_equal(g, m.oracle()), m.n


def test_batch_graph_replay_and_new_buffers():
    ctx = _capi.default_context()
    sizes = [700, 9000, 3, 33000]
    a = [Member(ctx, 50 + 2 * i, n) for i, n in enumerate(sizes)]
    b = [Member(ctx, 90 + 2 * i, n) for i, n in enumerate(sizes)]
    batch = MsmBatch(sizes, ctx=ctx)
    first = _run(batch, a)
    again = _run(batch, a)  # graph replay
    other = _run(batch, b)  # same graph, new descriptor pointers
    back = _run(batch, a)
    for m, x, y, z in zip(a, first, again, back):
        assert np.array_equal(x, y) and np.array_equal(x, z)
        assert np.array_equal(x, m.alone(ctx))
    for m, y in zip(b, other):
        assert np.array_equal(y, m.alone(ctx))
    # eager run after graph runs
    assert all(np.array_equal(x, y) for x, y in zip(_run(batch, b, use_graph=False), other))


def test_batch_forced_windows_many_groups():
    ctx = _capi.default_context()
    # two 20-bit members: the first one's top window (2^14 reachable buckets
    # of 2^19) is followed by the second one's keys
    sizes = [257, 3000, 1, 12000, 64, 5000, 900]
    widths = [2, 16, 7, 11, 20, 7, 20]
    members = [Member(ctx, 300 + 2 * i, n) for i, n in enumerate(sizes)]
    batch = MsmBatch(sizes, window_bits=widths, ctx=ctx)
    assert sorted(g[0] for g in batch.info()["groups"]) == [2, 7, 11, 16, 20]
    for m, g in zip(members, _run(batch, members)):
        assert np.array_equal(g, m.oracle()), m.n


def test_batch_width_cap_for_crowded_class():
    # nine members in one size class [2^14, 2^15), whose single calls take
    # wider windows: the planner caps theirs at 11 bits (per-member bucket
    # reduction grows as 2^(c-1)); results stay those of the single calls
    # and the oracle
    ctx = _capi.default_context()
    sizes = [16384 + 1500 * i for i in range(9)]
    members = [Member(ctx, 1300 + 2 * i, n) for i, n in enumerate(sizes)]
    def widths(ns):
        return {g[0] for g in MsmBatch(ns, ctx=ctx).info()["groups"]}

    single = [widths([n]).pop() for n in sizes]  # a lone member: the single-call width
    assert widths(sizes[:7]) == set(single[:7])  # seven in the class: no cap
    batch = MsmBatch(sizes, ctx=ctx)
    groups = batch.info()["groups"]
    assert sum(g[1] for g in groups) == len(sizes)
    assert {g[0] for g in groups} == {min(c, 11) for c in single}, (groups, single)
    got = _run(batch, members)
    for m, g in zip(members, got):
        assert np.array_equal(g, m.alone(ctx)), m.n
        assert np.array_equal(g, m.oracle()), m.n


def test_batch_same_inputs_repeated():
    # aliased members (the same buffers several times) give identical results
    ctx = _capi.default_context()
    m = Member(ctx, 77, 5000)
    outs = [DeviceBuffer(ctx, 64) for _ in range(5)]
    batch = MsmBatch([5000] * 5, ctx=ctx)
    batch.run([m.scl] * 5, [m.pts] * 5, outs)
    exp = m.oracle()
    for o in outs:
        assert np.array_equal(o.download((16,)), exp)


def test_batch_fixture_split():
    # the reference fixture cut into members: each piece against the oracle,
    # covering infinity (index 1023) and scalars >= r inside a batch
    ctx = _capi.default_context()
    scl, pts, _ = g1_fixture()
    cuts = [0, 100, 101, 600, 1024]
    sizes = [b - a for a, b in zip(cuts, cuts[1:])]
    ds = [DeviceBuffer.from_array(ctx, scl[a:b]) for a, b in zip(cuts, cuts[1:])]
    dp = [DeviceBuffer.from_array(ctx, pts[a:b]) for a, b in zip(cuts, cuts[1:])]
    outs = [DeviceBuffer(ctx, 64) for _ in sizes]
    MsmBatch(sizes, ctx=ctx).run(ds, dp, outs)
    for (a, b), o in zip(zip(cuts, cuts[1:]), outs):
        exp, _ = O.ec_jacreduce(scl[a:b], pts[a:b])
        assert np.array_equal(o.download((16,)), exp)


def test_batch_g2_alone_vs_mix():
    ctx = _capi.default_context()
    sizes = [1, 7, 300, 0, 2000, 4500]
    members = [Member(ctx, 500 + 2 * i, n, g2=True) for i, n in enumerate(sizes)]
    batch = MsmBatch(sizes, g2=True, ctx=ctx)
    for m, g in zip(members, _run(batch, members)):
        assert np.array_equal(g, m.alone(ctx)), m.n
        if m.n <= 2000:
            assert np.array_equal(g, m.oracle()), m.n


def test_tuned_single_calls_match_oracle():
    # the persisted tuning table (if any) changes launch settings only
    from cusnarks_amd import tune
    ctx = _capi.Context(0)
    used = tune.apply(ctx)
    m = Member(ctx, 901, 40000)
    assert np.array_equal(m.alone(ctx), m.oracle()), used


def test_shared_workspace_batches_and_regrowth():
    # batches on one context's shared workspace, run alternately; a larger
    # batch grows (moves) the workspace, after which the smaller batch's
    # captured graph must be rebuilt against the new addresses
    from cusnarks_amd.device import Event, Stream
    ctx = _capi.Context(0)
    small = [Member(ctx, 700 + 2 * i, n) for i, n in enumerate([100, 2000, 7])]
    big = [Member(ctx, 800 + 2 * i, n) for i, n in enumerate([60000, 3, 150000])]
    bs = MsmBatch([m.n for m in small], ctx=ctx, shared_workspace=True)
    exp_small = [m.alone(ctx) for m in small]
    exp_big = [m.alone(ctx) for m in big]
    st = Stream(ctx)
    ev = Event(ctx)
    for _ in range(2):
        bs.run([m.scl for m in small], [m.pts for m in small], [m.out for m in small], stream=st.ptr)
    ev.record(st)
    ev.sync()
    assert ev.done()
    assert all(np.array_equal(m.out.download((16,)), e) for m, e in zip(small, exp_small))
    bb = MsmBatch([m.n for m in big], ctx=ctx, shared_workspace=True)  # grows the workspace
    assert bb.workspace_bytes([m.n for m in big], ctx=ctx) > MsmBatch.workspace_bytes(
        [m.n for m in small], ctx=ctx)
    bb.run([m.scl for m in big], [m.pts for m in big], [m.out for m in big], stream=st.ptr)
    st.sync()
    assert all(np.array_equal(m.out.download((16,)), e) for m, e in zip(big, exp_big))
    for m in small:
        m.out.upload(np.zeros(16, np.uint32))
    bs.run([m.scl for m in small], [m.pts for m in small], [m.out for m in small], stream=st.ptr)
    st.sync()
    assert all(np.array_equal(m.out.download((16,)), e) for m, e in zip(small, exp_small))
