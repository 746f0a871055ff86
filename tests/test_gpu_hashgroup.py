"""GPU parity of GROUP BY over arbitrary key tuples (tgp_hash_group_agg_sync) against
the oracle's tuple grouping (orc_group_by_keys): sparse int64 keys over +-2^62,
two-column keys in the packed and in the hashed form (with the hash tags cut so that
distinct tuples collide and only the full-key check separates them), NULL keys, the
INT64_MIN key, hot keys through the LDS cache, MIN/MAX, the capacity protocol, and
1.5M groups over 60M rows."""
import ctypes as C

import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import DeviceBitmap, DeviceBuffer, DeviceColumn, make_agg_query, _capi as A
from turbogp_amd import workloads as W

pytestmark = pytest.mark.gpu
EQ, NE, LT, LE, GT, GE = range(6)
SIGNED = {A.I8, A.I16, A.I32, A.I64}


@pytest.fixture(params=[1, 2], ids=["one_pass", "partitioned"])
def part_mode(request, ctx):
    """The one-pass form (every row into the global table) or the partitioned form
    (rows by key hash, per-partition LDS tables merged into it), forced."""
    ctx.tune(A.TUNE_AGG_PART, request.param)
    yield request.param
    ctx.tune(A.TUNE_AGG_PART, 0)


def _words(flags):
    n = len(flags)
    w = np.zeros(max(1, (n + 63) // 64), np.uint64)
    idx = np.flatnonzero(flags)
    np.bitwise_or.at(w, idx >> 6, np.left_shift(np.uint64(1), (idx & 63).astype(np.uint64)))
    return w


def run_both(ctx, host_cols, preds, keys, aggs, nulls=None, b=0, e=None, sel_words=None,
             capacity=None):
    n = len(host_cols[0])
    e = n if e is None else e
    cols = [DeviceColumn.from_numpy(ctx, c) for c in host_cols]
    owned = []
    try:
        dsel = None
        if sel_words is not None:
            dsel = DeviceBitmap(ctx, e - b)
            dsel.buf.upload(sel_words)
            owned.append(dsel)
        dn = []
        for f in (nulls or [None] * len(keys)):
            if f is None:
                dn.append(None)
                continue
            bm = DeviceBitmap(ctx, e - b)
            bm.buf.upload(_words(f[b:e]))
            owned.append(bm)
            dn.append(bm)
        q = make_agg_query(cols, preds, (), aggs, b, e, dsel)
        got = ctx.hash_group_agg(q, keys, dn, capacity=capacity)
    finally:
        for c in cols + owned:
            c.free()
    signed = [host_cols[k].dtype.kind == "i" for k in keys]
    oq = B.make_query(host_cols, preds, (), aggs, b, e, sel_words)
    onulls = [None if f is None else _words(f[b:e]) for f in (nulls or [None] * len(keys))]
    want = B.group_by_keys(oq, keys, onulls, signed)
    have = got.as_dict(signed)
    assert len(have) == len(got), "a key tuple came out twice"
    assert have == want
    return have


AGGS = [("sum", 1, [(2, 0, 1)]), ("sum", 1, [(2, 0, 1), (3, 100, -1)]), ("count_nonzero", 3),
        ("min", 2), ("max_u64", 2)]


def test_sparse_int64_keys(ctx, part_mode):
    rng = np.random.default_rng(11)
    n = 1_000_000
    pool = rng.integers(-(2 ** 62), 2 ** 62, 5000, dtype=np.int64)
    k = rng.choice(pool, n)
    v = rng.integers(-(2 ** 45), 2 ** 45, n).astype(np.int64)
    w = rng.integers(0, 10, n).astype(np.int64)
    got = run_both(ctx, [k, np.zeros(n, np.uint8), v, w], [(3, LT, 8, True)], [0], AGGS)
    assert 4000 < len(got) <= 5000


def test_two_column_keys_packed_and_nulls(ctx, part_mode):
    rng = np.random.default_rng(12)
    n = 300_000
    a = rng.integers(0, 3, n).astype(np.uint8)
    bk = rng.integers(-50, 50, n).astype(np.int32)
    v = rng.integers(0, 10 ** 9, n).astype(np.int64)
    w = rng.integers(0, 10, n).astype(np.int64)
    got = run_both(ctx, [a, bk, v, w], [], [0, 1], AGGS,
                   nulls=[rng.random(n) < 0.1, rng.random(n) < 0.05])
    assert any(None in t for t in got)


@pytest.mark.parametrize("tag_bits", [0, 1, 3, 32])
def test_two_wide_columns_hashed_form(ctx, part_mode, tag_bits):
    rng = np.random.default_rng(13)
    n = 200_000
    a = rng.integers(-(2 ** 62), 2 ** 62, 700, dtype=np.int64)[rng.integers(0, 700, n)]
    bk = rng.integers(0, 2 ** 63, 900, dtype=np.int64).astype(np.uint64)[rng.integers(0, 900, n)]
    v = rng.integers(-(2 ** 30), 2 ** 30, n).astype(np.int64)
    w = rng.integers(0, 10, n).astype(np.int64)
    ctx.tune(A.TUNE_GROUP_TAG_BITS, tag_bits)
    try:
        got = run_both(ctx, [a, bk, v, w], [(3, NE, 4, True)], [0, 1], AGGS,
                       nulls=[rng.random(n) < 0.02, None])
    finally:
        ctx.tune(A.TUNE_GROUP_TAG_BITS, 0)
    assert len(got) > 10_000


def test_three_and_four_keys_with_tags_cut(ctx, part_mode):
    rng = np.random.default_rng(14)
    n = 100_000
    cols = [rng.integers(0, 4, n).astype(np.int16), rng.integers(0, 5, n).astype(np.uint16),
            rng.integers(-2, 2, n).astype(np.int8), rng.integers(0, 3, n).astype(np.uint32),
            rng.integers(-1000, 1000, n).astype(np.int64)]
    aggs = [("sum", 1, [(4, 0, 1)]), ("min", 4)]
    for keys in ([0, 1, 2], [0, 1, 2, 3]):
        for bits in (0, 2):
            ctx.tune(A.TUNE_GROUP_TAG_BITS, bits)
            try:
                run_both(ctx, cols, [], keys, aggs,
                         nulls=[None, rng.random(n) < 0.1, None, rng.random(n) < 0.3][:len(keys)])
            finally:
                ctx.tune(A.TUNE_GROUP_TAG_BITS, 0)


def test_special_keys_int64_min_and_null(ctx, part_mode):
    n = 50_000
    rng = np.random.default_rng(15)
    k = rng.choice(np.array([np.iinfo(np.int64).min, -1, 0, 1, np.iinfo(np.int64).max],
                            np.int64), n)
    v = rng.integers(-100, 100, n).astype(np.int64)
    got = run_both(ctx, [k, np.zeros(n, np.uint8), v, v], [], [0], AGGS[:1],
                   nulls=[rng.random(n) < 0.25])
    assert (np.iinfo(np.int64).min,) in got and (None,) in got


def test_hot_keys_lds_cache_and_big_values(ctx, part_mode):
    # six sparse keys over 4M rows: the LDS cache takes nearly every row; values
    # near 2^62 make the int128 carries matter
    rng = np.random.default_rng(16)
    n = 4_000_000
    k = rng.choice(np.array([-7 * 2 ** 58, 3, 2 ** 40, 2 ** 61, -1, 99], np.int64), n)
    v = rng.integers(2 ** 61, 2 ** 62, n).astype(np.int64)
    v[::3] *= -1
    w = rng.integers(0, 10, n).astype(np.int64)
    got = run_both(ctx, [k, np.zeros(n, np.uint8), v, w], [], [0], AGGS)
    assert len(got) == 6


def test_selection_bitmap_row_range_and_empty(ctx, part_mode):
    rng = np.random.default_rng(17)
    n = 100_003
    k = rng.integers(0, 2 ** 40, 2000, dtype=np.int64)[rng.integers(0, 2000, n)]
    v = rng.integers(0, 1000, n).astype(np.int64)
    b, e = 77, n - 5
    sel = rng.random(e - b) < 0.5
    run_both(ctx, [k, np.zeros(n, np.uint8), v, v], [(2, GE, 100, True)], [0], AGGS[:1],
             nulls=[rng.random(n) < 0.1], b=b, e=e, sel_words=_words(sel))
    assert run_both(ctx, [k, np.zeros(n, np.uint8), v, v], [(2, LT, 0, True)], [0],
                    AGGS[:1]) == {}
    assert run_both(ctx, [k, np.zeros(n, np.uint8), v, v], [], [0], AGGS[:1], b=10, e=10) == {}


def test_capacity_protocol(ctx, part_mode):
    rng = np.random.default_rng(18)
    n = 200_000
    k = rng.integers(-(2 ** 62), 2 ** 62, 30_000, dtype=np.int64)[rng.integers(0, 30_000, n)]
    v = rng.integers(0, 1000, n).astype(np.int64)
    truth = len(np.unique(k))
    cols = [DeviceColumn.from_numpy(ctx, c) for c in (k, v)]
    bufs = []
    try:
        q = make_agg_query(cols, (), (), [("sum", 1, [(1, 0, 1)])])
        hq = A.HashGroupQuery()
        hq.agg, hq.nkeys = q, 1
        hq.key_col[0] = 0
        for cap in (0, 100, truth - 1):
            bufs = [DeviceBuffer(ctx, max(8, cap * 16)) for _ in range(4)]
            out = A.HashGroupResult(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, cap)
            ng = C.c_uint64()
            rc = A.lib().tgp_hash_group_agg_sync(ctx.handle, C.byref(hq), C.byref(out),
                                                 C.byref(ng), None)
            assert rc == A.ERR_CAPACITY and cap < ng.value <= truth, (cap, rc, ng.value)
            for bf in bufs:
                bf.free()
            bufs = []
        # the wrapper grows its guess until the groups fit
        got = ctx.hash_group_agg(q, [0], capacity=1)
        assert len(got) == truth and int(got.result.count.sum()) == n
    finally:
        for c in cols + bufs:
            c.free()


def test_partitioned_form_lds_overflow_and_hot_keys(ctx):
    """Keys whose hash puts them all into one partition, more than its LDS table holds,
    so most go to the global table directly; plus one key holding half the rows."""
    rng = np.random.default_rng(12)
    cand = rng.integers(-2**62, 2**62, 400_000, dtype=np.int64)
    # keys whose top 3 hash bits are 0: with 5000 groups (16 partitions by the top 4
    # bits) they crowd 2 partitions, 2500 keys each against 1024 LDS slots
    keys = cand[(W.mix64_np(cand.view(np.uint64)) >> np.uint64(61)) == 0][:5000]
    assert len(keys) == 5000
    n = 400_000
    k = keys[rng.integers(0, len(keys), n)]
    k[::2] = keys[7]                                   # a hot key
    v = rng.integers(-2**40, 2**40, n).astype(np.int64)
    ctx.tune(A.TUNE_AGG_PART, 2)
    try:
        run_both(ctx, [k, v], [], [0], [("sum", 1, [(1, 0, 1)]), ("max", 1), ("min_u64", 1)],
                 capacity=len(keys))
    finally:
        ctx.tune(A.TUNE_AGG_PART, 0)


def test_1_5m_groups_over_60m_rows(ctx):
    # the 1.5M-value probe key of the join workload, grouped through the hash table
    n = W.JOIN_PROBE_ROWS
    gkey = W.device_uniform(ctx, n, A.I32, 20, 0, W.JOIN_GKEY_CARD - 1)
    val = W.device_uniform(ctx, n, A.I64, 21, 90_000, 10_500_000)
    try:
        q = make_agg_query([gkey, val], [(1, LT, 5_000_000, True)], (),
                           [("sum", 1, [(1, 0, 1)])])
        got = ctx.hash_group_agg(q, [0], capacity=W.JOIN_GKEY_CARD)
        assert len(got) == W.JOIN_GKEY_CARD
        hk = B.synth_uniform(A.I32, W.SEED, 20, 0, W.JOIN_GKEY_CARD - 1, 0, n)
        hv = B.synth_uniform(A.I64, W.SEED, 21, 90_000, 10_500_000, 0, n)
        oq = B.make_query([hk, hv], [(1, LT, 5_000_000, True)], (), [("sum", 1, [(1, 0, 1)])])
        assert got.as_dict() == B.group_by_keys(oq, [0])
    finally:
        gkey.free()
        val.free()
