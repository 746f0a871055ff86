"""GPU parity of the radix-partitioned GROUP BY (strategy 5: rows sorted into
partitions of consecutive group ids, each aggregated in an LDS table) against the
oracle: exact int128 sums with carries and signs, MIN/MAX in both orders,
COUNT(col), out-of-range keys, selection bitmaps, ragged row ranges, partition
boundaries and hundreds of partitions."""
import ctypes as C

import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import DeviceBitmap, DeviceColumn, make_agg_query, _capi as A

pytestmark = pytest.mark.gpu
EQ, NE, LT, LE, GT, GE = range(6)


def explain(q):
    st, sh = C.c_int32(), C.c_int32()
    A.check(A.lib().tgp_group_agg_explain(C.byref(q), C.byref(st), C.byref(sh)), "explain")
    return st.value


def run_both(ctx, host_cols, preds, parts, aggs, b=0, e=None, sel_words=None, mode=2,
             want_strategy=None):
    """mode: TGP_TUNE_AGG_PART for the device run (2 forces the partitioned form)."""
    cols = [DeviceColumn.from_numpy(ctx, c) for c in host_cols]
    dsel = None
    ctx.tune(A.TUNE_AGG_PART, mode)
    try:
        if sel_words is not None:
            dsel = DeviceBitmap(ctx, (e - b))
            dsel.buf.upload(sel_words)
        q = make_agg_query(cols, preds, parts, aggs, b, e, dsel)
        if want_strategy is not None:
            assert explain(q) == want_strategy
        got = ctx.group_agg(q)
    finally:
        ctx.tune(A.TUNE_AGG_PART, 0)
        for c in cols:
            c.free()
        if dsel:
            dsel.free()
    oq = B.make_query(host_cols, preds, parts, aggs, b, e, sel_words)
    cnt, sums, oor = B.group_agg(oq)
    assert got.n_out_of_range == oor
    assert np.array_equal(np.asarray(got.count, np.uint64), np.asarray(cnt, np.uint64))
    assert got.sums() == sums
    return got


def columns(rng, n, kmax):
    return [rng.integers(-2**62, 2**62, n).astype(np.int64),    # wide: wrapping products
            rng.integers(0, 7, n).astype(np.uint8),
            rng.integers(-1000, 1000, n).astype(np.int16),
            rng.integers(-3, kmax, n).astype(np.int32),         # group key, a few below 0
            rng.integers(-5, 5, n).astype(np.int8),
            rng.integers(0, 2**64, n, dtype=np.uint64)]


SUMS = [("sum", 1, [(0, 0, 1)]), ("sum", 1, [(0, 3, -1), (2, 1, 1)]),
        ("sum", -3, [(2, 5, 7), (4, 0, 1), (3, 1, -1)]), ("count_nonzero", 4),
        ("sum", 1, [(0, 0, 1), (0, 0, 1)]), ("sum", 9, [])]
MINMAX = [("min", 2), ("max", 0), ("min_u64", 5), ("max_u64", 4), ("sum", 1, [(2, 0, 1)])]


@pytest.mark.parametrize("ngroups,aggs", [(5000, SUMS), (5000, MINMAX), (40_000, SUMS[:1]),
                                          (300_000, SUMS[:2]), (70_001, []),
                                          (2_000_000, SUMS[:1])])
def test_partitioned_vs_oracle(ctx, ngroups, aggs):
    rng = np.random.default_rng(ngroups + len(aggs))
    n = 400_000
    cols = columns(rng, n, ngroups + 3)   # keys ngroups .. ngroups + 2 are out of range
    preds = [(2, NE, 17, True)]
    parts = [(3, 1, 0, ngroups)]
    run_both(ctx, cols, preds, parts, aggs, want_strategy=5)  # >= 2^18 rows: the default
    run_both(ctx, cols, preds, parts, aggs, 12_345, n - 6_789)
    # an input selection, and a range that is not a multiple of 64 at either end
    sel = rng.integers(0, 2**64, (n - 70 + 63) // 64, dtype=np.uint64)
    run_both(ctx, cols, [(3, LT, ngroups // 2, True)], parts, aggs, 33, n - 37, sel)


def test_partition_boundaries_and_composite_keys(ctx):
    # 2 aggregates (5 table words): 1024 groups per partition; keys on both sides of
    # partition boundaries
    n = 300_000
    rng = np.random.default_rng(5)
    edges = np.array([0, 1023, 1024, 1025, 2047, 2048, 4095, 4096, 6143, 6144, 6999], np.int32)
    key = edges[rng.integers(0, len(edges), n)]
    val = rng.integers(-2**63, 2**63, n, dtype=np.int64)
    run_both(ctx, [key, val], [], [(0, 1, 0, 7000)],
             [("sum", 1, [(1, 0, 1)]), ("max", 1)])
    # a composite key (a x b, one part with a divisor) and a base offset
    a = rng.integers(100, 400, n).astype(np.int64)
    bcol = rng.integers(0, 90_000, n).astype(np.int32)
    run_both(ctx, [a, bcol, val], [(0, GE, 120, True)],
             [(0, 1, 100, 300), (1, 30, 0, 3000)], [("sum", 1, [(2, 0, 1)])])


def test_nothing_selected_and_tiny_ranges(ctx):
    rng = np.random.default_rng(9)
    n = 50_000
    cols = columns(rng, n, 9000)
    parts = [(3, 1, 0, 9000)]
    got = run_both(ctx, cols, [(3, LT, -100, True)], parts, SUMS[:2])
    assert not np.asarray(got.count).any()
    for b, e in ((0, 0), (5, 6), (64, 129), (n - 1, n)):
        run_both(ctx, cols, [], parts, MINMAX, b, e)


def test_int128_carries_in_partitions(ctx):
    n = 1_000_000
    v = np.full(n, 2**62 + 12345, np.int64)
    v[::3] = -(2**62) - 7
    key = (np.arange(n) % 20_000).astype(np.int32)
    key[::7] = 19_999                        # one heavy group
    run_both(ctx, [v, key], [], [(1, 1, 0, 20_000)], [("sum", 1, [(0, 0, 1)]), ("min", 0)],
             mode=0, want_strategy=5)


def test_strategy_choice(ctx):
    rng = np.random.default_rng(1)
    n = 1 << 18
    cols = [DeviceColumn.from_numpy(ctx, c) for c in columns(rng, n, 5000)[:4]]
    try:
        q = make_agg_query(cols, [], [(3, 1, 0, 5000)], [("sum", 1, [(0, 0, 1)])])
        assert explain(q) == 5
        q = make_agg_query(cols, [], [(3, 1, 0, 5000)], [("sum", 1, [(0, 0, 1)])], 0, n - 1)
        assert explain(q) == 3                # below 2^18 rows: the global table
        q = make_agg_query(cols, [], [(3, 1, 0, 500)], [("sum", 1, [(0, 0, 1)])])
        assert explain(q) == 2                # fits an LDS table
        # more partitions than the scatter holds: the global table
        q = make_agg_query(cols, [], [(3, 1, 0, 1 << 21), (1, 1, 0, 7)],
                           [("sum", 1, [(0, 0, 1)])] * 6)
        assert explain(q) == 3
    finally:
        for c in cols:
            c.free()
