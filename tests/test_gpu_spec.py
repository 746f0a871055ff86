"""Shape-specialised kernels: the same queries through the specialised path,
the generic engine and the oracle must agree bit for bit."""
import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import DeviceColumn, make_agg_query, _capi as A
from turbogp_amd import workloads as W

pytestmark = pytest.mark.gpu
EQ, NE, LT, LE, GT, GE = range(6)


@pytest.fixture(scope="module")
def fact(ctx):
    n = 200_003
    host = W.host_fact_columns(n, B.synth_uniform)
    dev = {k: DeviceColumn.from_numpy(ctx, v) for k, v in host.items()}
    yield host, dev
    for c in dev.values():
        c.free()


def kernel_forms(ctx, q):
    """The query through every kernel form: the default choice, the
    replicated-histogram specialised kernel (static interleave, then one-slice
    chunks from a counter), the per-slice-fold specialised kernel and the
    generic engine."""
    out = []
    try:
        ctx.tune(A.TUNE_SPECIALIZE, 1)
        out.append(ctx.group_agg(q))
        ctx.tune(A.TUNE_SPEC_KERNEL, 2)
        out.append(ctx.group_agg(q))
        ctx.tune(A.TUNE_CHUNK, 1)
        out.append(ctx.group_agg(q))
        ctx.tune(A.TUNE_CHUNK, 0)
        ctx.tune(A.TUNE_SPEC_KERNEL, 1)
        out.append(ctx.group_agg(q))
        ctx.tune(A.TUNE_SPECIALIZE, 0)
        out.append(ctx.group_agg(q))
    finally:
        ctx.tune(A.TUNE_SPECIALIZE, 1)
        ctx.tune(A.TUNE_SPEC_KERNEL, 0)
        ctx.tune(A.TUNE_CHUNK, 0)
    return out


def both_paths(ctx, host, dev, spec, b=0, e=None):
    names, preds, parts, aggs = spec
    q = make_agg_query([dev[c] for c in names], preds, parts, aggs, b, e)
    forms = kernel_forms(ctx, q)
    cnt, sums, oor = B.group_agg(B.make_query([host[c] for c in names], preds, parts, aggs, b, e))
    for got in forms:
        assert [int(x) for x in got.count] == [int(x) for x in cnt]
        assert got.sums() == sums
        assert got.n_out_of_range == oor


@pytest.mark.parametrize("group", ["ab", "a", "b", "month", ""])
@pytest.mark.parametrize("b,e", [(0, None), (1, 200_003), (255, 257), (300, 300), (12_345, 199_999)])
def test_q1_family(ctx, fact, group, b, e):
    host, dev = fact
    both_paths(ctx, host, dev, W.q1_spec(date_cutoff=1800, group=group), b, e)


@pytest.mark.parametrize("op", range(6))
@pytest.mark.parametrize("signed", [True, False])
@pytest.mark.parametrize("lit", [-(2**63), -5, 0, 1, 1200, 2525, 2526, 2**31, 2**63 - 1])
def test_date_predicate_intervals(ctx, fact, op, signed, lit):
    host, dev = fact
    names, _, parts, aggs = W.q1_spec(group="ab")
    both_paths(ctx, host, dev, (names, [(0, op, lit, signed)], parts, aggs), 7, 150_001)


def test_q6_and_variants(ctx, fact):
    host, dev = fact
    both_paths(ctx, host, dev, W.q6_spec())
    names, preds, parts, aggs = W.q6_spec()
    preds2 = preds + [(1, NE, 6, True)]  # adds a single-point exclusion
    both_paths(ctx, host, dev, (names, preds2, parts, aggs))
    preds3 = [(0, GE, 900, False), (0, LT, 100, False)]  # empty intersection
    both_paths(ctx, host, dev, (names, preds3, parts, aggs))


def test_big_values_take_the_exact_slow_path(ctx):
    rng = np.random.default_rng(3)
    n = 70_001
    host = W.host_fact_columns(n, B.synth_uniform)
    host["price"] = rng.integers(2**60, 2**62, n).astype(np.int64)   # |value| >= 2^55
    host["qty"] = rng.integers(-2**62, 2**62, n).astype(np.int64)
    dev = {k: DeviceColumn.from_numpy(ctx, v) for k, v in host.items()}
    try:
        both_paths(ctx, host, dev, W.q1_spec(group="ab"))
        both_paths(ctx, host, dev, W.q1_spec(group="month"))
    finally:
        for c in dev.values():
            c.free()


def test_short_columns_tail_slices(ctx):
    for n in (1, 63, 255, 256, 257, 1000):
        host = W.host_fact_columns(n, B.synth_uniform)
        dev = {k: DeviceColumn.from_numpy(ctx, v) for k, v in host.items()}
        try:
            both_paths(ctx, host, dev, W.q1_spec(group="ab"))
            both_paths(ctx, host, dev, W.q6_spec())
        finally:
            for c in dev.values():
                c.free()


def test_wave_kernel_flushes_every_2_20_rows(ctx):
    """4 waves over 5.3M rows: in 4096-slice chunks a wave takes a second chunk
    and flushes its LDS partials in between; statically interleaved, every wave
    adds 1.3M rows and flushes once on the way."""
    n = 5_300_000
    table = W.make_fact_table(ctx, n)
    try:
        for spec in (W.q1_spec(date_cutoff=2000, group=g) for g in ("ab", "b", "month")):
            q = W.build_query(table, spec, 3, n - 5)
            want = ctx.group_agg(q)  # automatic settings
            try:
                ctx.tune(A.TUNE_SPEC_KERNEL, 2)
                ctx.tune(A.TUNE_GRID_MAX, 1)
                static = ctx.group_agg(q)
                ctx.tune(A.TUNE_CHUNK, 4096)
                got = ctx.group_agg(q)
                ctx.tune(A.TUNE_SPECIALIZE, 0)
                generic = ctx.group_agg(q)
            finally:
                ctx.tune(A.TUNE_GRID_MAX, 0)
                ctx.tune(A.TUNE_CHUNK, 0)
                ctx.tune(A.TUNE_SPECIALIZE, 1)
                ctx.tune(A.TUNE_SPEC_KERNEL, 0)
            for r in (static, got, generic):
                assert np.array_equal(r.count, want.count)
                assert np.array_equal(r.sums_words, want.sums_words)
                assert r.n_out_of_range == want.n_out_of_range
    finally:
        for c in table.values():
            c.free()
