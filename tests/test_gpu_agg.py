"""GPU parity of the fused filter + GROUP BY + exact aggregation against the
oracle (bit-exact int128 sums and counts, compared per group id)."""
import json
import os

import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import DeviceBitmap, DeviceColumn, make_agg_query, _capi as A
from turbogp_amd import workloads as W

pytestmark = pytest.mark.gpu
EQ, NE, LT, LE, GT, GE = range(6)


def run_both(ctx, host_cols, preds, parts, aggs, b=0, e=None, sel_words=None):
    cols = [DeviceColumn.from_numpy(ctx, c) for c in host_cols]
    dsel = None
    try:
        if sel_words is not None:
            dsel = DeviceBitmap(ctx, (e - b))
            dsel.buf.upload(sel_words)
        q = make_agg_query(cols, preds, parts, aggs, b, e, dsel)
        got = ctx.group_agg(q)
    finally:
        for c in cols:
            c.free()
        if dsel:
            dsel.free()
    oq = B.make_query(host_cols, preds, parts, aggs, b, e, sel_words)
    cnt, sums, oor = B.group_agg(oq)
    assert got.n_out_of_range == oor
    assert [int(x) for x in got.count] == [int(x) for x in cnt]
    assert got.sums() == sums
    return got


def test_q1_golden_1m(ctx):
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fact_1m.json")))
    n = fx["rows"]
    table = W.make_fact_table(ctx, n)   # generated on the device
    try:
        for k, c in table.items():      # device generator == oracle generator
            cid, dt, lo, hi = W.FACT_SCHEMA[k]
            assert np.array_equal(c.to_numpy(), B.synth_uniform(dt, W.SEED, cid, lo, hi, 0, n)), k
        res = ctx.group_agg(W.build_query(table, W.q1_spec()))
        for gid, g in enumerate(fx["q1"]["groups"]):
            assert int(res.count[gid]) == g["count"]
            assert res.sums()[gid] == [int(g[k]) for k in ("sum_qty", "sum_price",
                                                           "sum_disc_price", "sum_charge")]
        r6 = ctx.group_agg(W.build_query(table, W.q6_spec()))
        assert int(r6.count[0]) == fx["q6"]["count"]
        assert r6.sums()[0][0] == int(fx["q6"]["sum_price_disc"])
    finally:
        for c in table.values():
            c.free()


@pytest.mark.parametrize("group", ["ab", "a", "b", "month", ""])
@pytest.mark.parametrize("b,e", [(0, 300_000), (12_345, 250_001), (1, 2)])
def test_q1_shapes_vs_oracle(ctx, group, b, e):
    cols = W.host_fact_columns(300_000, B.synth_uniform)
    names, preds, parts, aggs = W.q1_spec(date_cutoff=2000, group=group)
    run_both(ctx, [cols[c] for c in names], preds, parts, aggs, b, e)


def test_q6_vs_oracle(ctx):
    cols = W.host_fact_columns(500_000, B.synth_uniform)
    names, preds, parts, aggs = W.q6_spec()
    run_both(ctx, [cols[c] for c in names], preds, parts, aggs)


@pytest.mark.parametrize("seed", range(6))
def test_random_queries_vs_oracle(ctx, seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 200_000))
    cols = [rng.integers(-2**62, 2**62, n).astype(np.int64),   # wide, wraps in products
            rng.integers(0, 7, n).astype(np.uint8),
            rng.integers(-1000, 1000, n).astype(np.int16),
            rng.integers(0, 5000, n).astype(np.int32),
            rng.integers(-5, 5, n).astype(np.int8)]
    ncard = [1, 2, 7, 40, 300, 5000][seed]
    preds = [(2, int(rng.integers(0, 6)), int(rng.integers(-500, 500)), True),
             (3, LT, int(rng.integers(100, 5000)), False)]
    parts = [(1, 1, 0, 7)] if ncard == 7 else [(3, max(1, 5000 // ncard), 0, ncard)]
    if ncard == 1:
        parts = []
    aggs = [("sum", 1, [(0, 0, 1)]), ("sum", 1, [(0, 3, -1), (2, 1, 1)]),
            ("sum", -3, [(2, 5, 7), (4, 0, 1), (3, 1, -1)]), ("count_nonzero", 4),
            ("sum", 1, [(0, 0, 1), (0, 0, 1)]), ("sum", 9, [])]
    b = int(rng.integers(0, n))
    e = int(rng.integers(b, n + 1))
    run_both(ctx, cols, preds, parts, aggs, b, e)


def test_group_out_of_range_and_selection_bitmap(ctx):
    rng = np.random.default_rng(7)
    n = 150_000
    key = rng.integers(-3, 12, n).astype(np.int64)
    val = rng.integers(-10**12, 10**12, n).astype(np.int64)
    sel = rng.integers(0, 2**63, (n + 63) // 64, dtype=np.int64).astype(np.uint64)
    run_both(ctx, [key, val], [(1, NE, 0, True)], [(0, 1, 0, 10)],
             [("sum", 1, [(1, 0, 1)])], 0, n, sel)


def test_group_by_kat(ctx):  # tests/joins_groupby_orderby.rs:82-106
    grp = np.array([1, 1, 1, 2, 2], np.int64)
    a = np.array([10, 20, 30, 40, 50], np.int64)
    got = run_both(ctx, [grp, a], [], [(0, 1, 1, 2)], [("sum", 1, [(1, 0, 1)])])
    assert got.rows() == [(0, 3, [60]), (1, 2, [90])]


def test_arithmetic_kats(ctx):  # tests/arithmetic_aggregates.rs:14-73
    price = np.array([100, 200, 300], np.int64)
    disc = np.array([10, 20, 30], np.int64)
    got = run_both(ctx, [price, disc], [], [], [
        ("sum", 1, [(0, 0, 2)]), ("sum", 1, [(0, 0, 1)]), ("sum", 1, [(1, 0, 1)]),
        ("sum", 1, [(0, 0, 1), (1, 0, 1)])])
    s = got.sums()[0]
    assert s[0] == 1200 and s[1] + s[2] == 660 and s[1] - s[2] == 540 and s[3] == 14000


def test_int128_carries_exact(ctx):
    # 3M rows of ~2^62: sums need more than 64 bits; negatives exercise the sign words
    n = 3_000_000
    v = np.full(n, 2**62 + 12345, np.int64)
    v[::3] = -(2**62) - 7
    grp = (np.arange(n) % 5).astype(np.uint8)
    run_both(ctx, [v, grp], [], [(1, 1, 0, 5)], [("sum", 1, [(0, 0, 1)])])
    run_both(ctx, [v, grp], [], [(1, 1, 0, 5)] * 0, [("sum", 1, [(0, 0, 1)])])
    run_both(ctx, [v, (np.arange(n) % 1000).astype(np.int32)], [], [(1, 1, 0, 1000)],
             [("sum", 1, [(0, 0, 1)])])


def test_determinism_repeated_runs(ctx):
    cols = W.host_fact_columns(200_000, B.synth_uniform)
    names, preds, parts, aggs = W.q1_spec()
    dcols = [DeviceColumn.from_numpy(ctx, cols[c]) for c in names]
    q = make_agg_query(dcols, preds, parts, aggs)
    r = [ctx.group_agg(q) for _ in range(3)]
    assert all(np.array_equal(x.count, r[0].count) and np.array_equal(x.sums_words,
                                                                       r[0].sums_words)
               for x in r)
    for c in dcols:
        c.free()


@pytest.mark.parametrize("ncard", [1, 5, 7, 300, 5000])
def test_min_max_vs_oracle(ctx, ncard):
    """MIN/MAX in signed and raw-u64 order, over columns of every width and a
    wrapping expression, through each accumulation strategy (registers for
    <= 8 groups, LDS table, global table)."""
    rng = np.random.default_rng(ncard)
    n = 180_000
    cols = [rng.integers(-2**62, 2**62, n).astype(np.int64),
            rng.integers(0, 7, n).astype(np.uint8),
            rng.integers(-1000, 1000, n).astype(np.int16),
            rng.integers(0, 5000, n).astype(np.int32),
            rng.integers(-5, 5, n).astype(np.int8),
            rng.integers(0, 2**64, n, dtype=np.uint64)]
    parts = [] if ncard == 1 else [(1, 1, 0, 7)] if ncard == 7 else \
        [(3, max(1, 5000 // ncard), 0, ncard)]
    preds = [(3, LT, 4_000, False)]
    small = [("min", 2), ("max", 4), ("min_u64", 4), ("max_u64", 0)]
    run_both(ctx, cols, preds, parts, small, 3, n - 5)
    wide = small + [("max", 5), ("min", 1, [(0, 3, -1), (2, 1, 1)])]
    run_both(ctx, cols, preds, parts, wide, 0, n)
    run_both(ctx, cols, preds, parts, [("sum", 1, [(2, 0, 1)]), ("min_u64", 5)], 17, 100_017)
    # nothing selected: every group is empty, values read 0
    got = run_both(ctx, cols, [(3, LT, 0, False)], parts, small, 0, n)
    assert not got.count.any() and not got.sums_words.any()


def test_min_max_kats(ctx):  # tests/interpreter_fallback.rs min_max_together
    region = np.array([1, 1, 2, 2, 3], np.int64)
    amount = np.array([100, 200, 150, 300, 250], np.int64)
    got = run_both(ctx, [region, amount], [], [], [("min", 1), ("max", 1)])
    assert got.sums()[0] == [100, 300]
    got = run_both(ctx, [region, amount], [], [(0, 1, 1, 3)], [("min", 1), ("max", 1)])
    assert got.rows() == [(0, 2, [100, 200]), (1, 2, [150, 300]), (2, 1, [250, 250])]
    # raw u64 order (accum_grouped.rs MinCol/MaxCol): -1 is the largest cell
    v = np.array([-1, 5, 3], np.int64)
    got = run_both(ctx, [v], [], [], [("min", 0), ("max", 0), ("min_u64", 0), ("max_u64", 0)])
    assert got.sums()[0] == [-1, 5, 3, 2**64 - 1]
