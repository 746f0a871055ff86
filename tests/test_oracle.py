"""CPU checks of the oracle: the reference's own known-answer tests, closed forms,
brute force on small inputs and the committed golden fixtures."""
import json
import os

import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import workloads as W

EQ, NE, LT, LE, GT, GE = range(6)
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "fact_1m.json")


def bits_list(bits, n):
    return [bool((bits[i >> 3] >> (i & 7)) & 1) for i in range(n)]


def u64(vals):
    return np.array([v & (2**64 - 1) for v in vals], dtype=np.uint64)


# ---- exec/bitmap.rs tests (bitmap.rs:1126-1470) ------------------------------------------
def test_filter_eq_u64_kat():  # bitmap.rs test_filter_eq_u64
    bm = B.filter_cells(u64([1, 5, 5, 3, 5, 7, 8, 5, 9, 10]), EQ, 0, 5)
    assert bits_list(bm, 10) == [False, True, True, False, True, False, False, True, False, False]


@pytest.mark.parametrize("op,exp", [
    (LT, [1, 1, 1, 1, 0, 0, 0, 0, 0, 0]), (GE, [0, 0, 0, 0, 1, 1, 1, 1, 1, 1]),
    (LE, [1, 1, 1, 1, 1, 0, 0, 0, 0, 0]), (GT, [0, 0, 0, 0, 0, 1, 1, 1, 1, 1])])
def test_filter_ordered_u64_kat(op, exp):  # bitmap.rs test_filter_{lt,ge,le,gt}_u64
    bm = B.filter_cells(u64(range(10, 101, 10)), op, 0, 50)
    assert bits_list(bm, 10) == [bool(x) for x in exp]


def test_filter_ne_u64_kat():
    bm = B.filter_cells(u64([1, 5, 5, 3, 5, 7, 8, 5, 9, 10]), NE, 0, 5)
    assert bits_list(bm, 10) == [True, False, False, True, False, True, True, False, True, True]


def test_filter_signed_kats():  # bitmap.rs test_filter_lt_i64_signed / ge
    col = u64([-5, -3, -1, 0, 1, 3, 5])
    assert bits_list(B.filter_cells(col, LT, 1, 0), 7) == [True] * 3 + [False] * 4
    assert bits_list(B.filter_cells(col, GE, 1, 0), 7) == [False] * 3 + [True] * 4
    # the unsigned compare of the same cells differs (negatives are huge as u64)
    assert bits_list(B.filter_cells(col, LT, 0, 0), 7) == [False] * 7


def test_large_filter_counts():  # bitmap.rs test_large_filter_u64 / test_large_and_into_bool
    col = np.arange(1_000_000, dtype=np.uint64) % 100
    eq = B.filter_cells(col, EQ, 0, 50)
    assert B.count_range(eq, len(col), 0, len(col)) == 10_000
    ge = B.filter_cells(col, GE, 0, 10)
    le = B.filter_cells(col, LE, 0, 50)
    both = ge & le
    assert B.count_range(both, len(col), 0, len(col)) == 410_000


def test_set_bits_kats():  # bitmap.rs test_iter_set_bits_*
    n = 100
    bits = np.zeros(13, np.uint8)
    for i in (5, 31, 63, 99):
        bits[i >> 3] |= 1 << (i & 7)
    assert list(B.set_bits(bits, n)) == [5, 31, 63, 99]
    bits = np.zeros(4, np.uint8)
    bits[3] = 1
    assert list(B.set_bits(bits, 32)) == [24]
    assert list(B.set_bits(np.zeros(7, np.uint8), 50)) == []
    ones = np.full(3, 0xFF, np.uint8)
    ones[2] >>= 4  # all_ones(20)
    assert list(B.set_bits(ones, 20)) == list(range(20))


def test_count_ones_range_kat():  # bitmap.rs test_count_ones_range
    bits = np.zeros(13, np.uint8)
    for i in range(10, 50):
        bits[i >> 3] |= 1 << (i & 7)
    for (a, b, e) in [(0, 100, 40), (10, 50, 40), (0, 10, 0), (50, 100, 0), (20, 30, 10),
                      (15, 45, 30)]:
        assert B.count_range(bits, 100, a, b) == e


def test_not_clears_tail():  # bitmap.rs test_not
    bits = np.zeros(2, np.uint8)
    for i in (0, 3, 7):
        bits[i >> 3] |= 1 << (i & 7)
    nb = B.bitmap_not(bits, 10)
    assert bits_list(nb, 10) == [not x for x in bits_list(bits, 10)]
    assert B.count_range(nb, 10, 0, 10) == 7
    assert nb[1] == 0b11  # bits 10..15 cleared


# ---- kernel/scan.rs scalar tests (scan.rs:641-884) ------------------------------------------
def test_scan_kats():
    assert B.scan_multi(u64([1, 2, 3, 2, 4, 2, 5]), [(2, 0)])[0] == 3
    assert B.scan_multi(np.arange(21, dtype=np.uint64), [(5, 0), (2, 1), (10, 2)])[0] == 1
    assert B.scan_multi(np.full(100, 5, np.uint64), [(5, 0), (0, 1), (100, 2)])[0] == 100
    assert B.scan_multi(np.arange(1000, dtype=np.uint64), [(1000, 1), (1000, 2), (5, 0)])[0] == 0
    assert B.scan_multi(np.zeros(0, np.uint64), [(5, 0), (2, 1), (10, 2)])[0] == 0
    assert B.scan_multi(np.arange(100, dtype=np.uint64), [(7, 0)])[0] == 1
    assert B.scan_multi(np.arange(100, dtype=np.uint64), [(50, 1), (60, 2)])[0] == 9
    assert B.scan_multi(np.arange(100, dtype=np.uint64), [(10, 1), (20, 2)])[0] == 9
    # ScanRange [100, 200] over 0..999 = 101 (scan.rs avx512_scan_range_matches_scalar)
    assert B.scan_multi(np.arange(1000, dtype=np.uint64), [(99, 1), (201, 2)])[0] == 101


# ---- aggregation: reference tests + brute force --------------------------------------------
def test_arithmetic_aggregate_kats():  # tests/arithmetic_aggregates.rs:14-73
    price = np.array([100, 200, 300], np.int64)
    disc = np.array([10, 20, 30], np.int64)
    cases = [
        ([("sum", 1, [(0, 0, 2)])], 1200),            # SUM(price * 2)
        ([("sum", 1, [(0, 0, 1)]), ("sum", 1, [(1, 0, 1)])], 660),   # SUM(price + discount)
        ([("sum", 1, [(0, 0, 1)]), ("sum", -1, [(1, 0, 1)])], 540),  # SUM(price - discount)
        ([("sum", 2, [(0, 0, 1)]), ("sum", -2, [(1, 0, 1)])], 1080),  # SUM((price-discount)*2)
        ([("sum", 1, [(0, 0, 1), (1, 0, 1)])], 14000),  # SUM(price * discount)
        ([("sum", 1, [(0, 0, 1)])], 600),             # SUM(price)
    ]
    for aggs, want in cases:
        q = B.make_query([price, disc], aggs=aggs)
        cnt, sums, _ = B.group_agg(q)
        assert int(cnt[0]) == 3
        assert sum(sums[0]) == want


def test_group_by_kat():  # tests/joins_groupby_orderby.rs:82-106 (grp 1 -> 60,3; grp 2 -> 90,2)
    grp = np.array([1, 1, 1, 2, 2], np.int64)
    a = np.array([10, 20, 30, 40, 50], np.int64)
    q = B.make_query([grp, a], parts=[(0, 1, 1, 2)], aggs=[("sum", 1, [(1, 0, 1)])])
    cnt, sums, _ = B.group_agg(q)
    assert list(cnt) == [3, 2] and sums == [[60], [90]]


def test_min_max_kats():
    # tests/interpreter_fallback.rs min_max_together: min(amount) = 100, max = 300
    region = np.array([1, 1, 2, 2, 3], np.int64)
    amount = np.array([100, 200, 150, 300, 250], np.int64)
    cnt, sums, _ = B.group_agg(B.make_query([region, amount], aggs=[("min", 1), ("max", 1)]))
    assert int(cnt[0]) == 5 and sums[0] == [100, 300]
    # tests/clickbench_executor.rs min_max_multi_aggregate: ids 1..5
    ids = np.arange(1, 6, dtype=np.int64)
    _, sums, _ = B.group_agg(B.make_query([ids], aggs=[("min", 0), ("max", 0)]))
    assert sums[0] == [1, 5]
    # the same in raw u64 order (accum_grouped.rs:269-280) differs only for negatives
    v = np.array([-1, 5, 3], np.int64)
    _, sums, _ = B.group_agg(B.make_query([v], aggs=[("min", 0), ("max", 0), ("min_u64", 0),
                                                     ("max_u64", 0)]))
    assert sums[0] == [-1, 5, 3, 2**64 - 1]
    _, f64, _ = B.group_agg(B.make_query([v], aggs=[("min", 0), ("max_u64", 0)]), mode=1)
    assert f64[0] == [-1.0, float(2**64 - 1)]


@pytest.mark.parametrize("seed", range(3))
def test_min_max_vs_numpy(seed):
    """Per-group MIN/MAX against numpy's reductions, signed and unsigned, over a
    plain column and a wrapping product expression."""
    rng = np.random.default_rng(seed)
    n = 20_000
    key = rng.integers(0, 9, n).astype(np.int64)
    a = rng.integers(-2**62, 2**62, n).astype(np.int64)
    b = rng.integers(-7, 7, n).astype(np.int8)
    aggs = [("min", 1), ("max", 1), ("min_u64", 1), ("max_u64", 1),
            ("min", 3, [(1, 0, 1), (2, 5, -1)]), ("max_u64", 3, [(1, 0, 1), (2, 5, -1)])]
    cnt, sums, _ = B.group_agg(B.make_query([key, a, b], parts=[(0, 1, 0, 9)], aggs=aggs))
    with np.errstate(over="ignore"):
        expr = (np.int64(3) * a * (np.int64(5) - b.astype(np.int64))).astype(np.int64)
    for g in range(9):
        m = key == g
        assert int(cnt[g]) == int(m.sum())
        au, eu = a[m].view(np.uint64), expr[m].view(np.uint64)
        assert sums[g] == [int(a[m].min()), int(a[m].max()), int(au.min()), int(au.max()),
                           int(expr[m].min()), int(eu.max())]


def _result(cnt, sums):
    from turbogp_amd.device import GroupAggResult
    w = np.array([[[v % 2**64, (v >> 64) % 2**64] for v in g] for g in sums], np.uint64)
    return GroupAggResult(np.asarray(cnt, np.uint64), w, 0)


def test_avg_forms():
    from fractions import Fraction
    # tests/interpreter_fallback.rs multi_aggregate_sum_and_avg: avg(amount) = 200
    amount = np.array([100, 200, 150, 300, 250], np.int64)
    cnt, sums, _ = B.group_agg(B.make_query([amount], aggs=[("sum", 1, [(0, 0, 1)])]))
    r = _result(cnt, sums)
    assert r.avg(0) == [200] and r.avg(0, "f64") == [200.0] and r.avg(0, "u64") == [200]
    # tests/joins_groupby_orderby.rs:71 data: avg(b) GROUP BY grp -> 150, 300
    grp = np.array([1, 1, 2], np.int64)
    b = np.array([100, 200, 300], np.int64)
    cnt, sums, _ = B.group_agg(B.make_query([grp, b], parts=[(0, 1, 1, 2)],
                                            aggs=[("sum", 1, [(1, 0, 1)])]))
    assert _result(cnt, sums).avg(0, "f64") == [150.0, 300.0]
    # exact vs the two reference forms on a negative, non-integral mean
    v = np.array([-3, 1, 1], np.int64)
    cnt, sums, _ = B.group_agg(B.make_query([v], aggs=[("sum", 1, [(0, 0, 1)])]))
    r = _result(cnt, sums)
    assert r.avg(0) == [Fraction(-1, 3)] and r.avg(0, "f64") == [-1 / 3]
    assert r.avg(0, "u64") == [(2**64 - 1) // 3]  # wrapping u64 sum, unsigned division
    # an empty group has no average
    cnt, sums, _ = B.group_agg(B.make_query([grp, b], parts=[(0, 1, 1, 3)],
                                            aggs=[("sum", 1, [(1, 0, 1)])]))
    assert _result(cnt, sums).avg(0) == [150, 300, None]


def brute_group_agg(cols, preds, parts, aggs, b, e, sel=None):
    """Pure-Python brute force for small inputs."""
    ng = int(np.prod([p[3] for p in parts])) if parts else 1
    cnt = [0] * ng
    sums = [[0] * len(aggs) for _ in range(ng)]
    oor = 0
    M = 2**64
    for r in range(b, e):
        ok = True
        for (ci, op, lit, sg) in preds:
            v = int(cols[ci][r])
            x, l = (v, lit) if sg else (v % M, lit % M)
            ok &= [x == l, x != l, x < l, x <= l, x > l, x >= l][op]
        if sel is not None:
            i = r - b
            ok &= bool((int(sel[i >> 6]) >> (i & 63)) & 1)
        if not ok:
            continue
        gid, stride, bad = 0, 1, False
        for (ci, div, base, card) in parts:
            v = int(cols[ci][r])
            if div > 1:
                v = int(v / div)
            p = v - base
            bad |= not (0 <= p < card)
            gid += p * stride
            stride *= card
        if bad:
            oor += 1
            continue
        cnt[gid] += 1
        for k, ag in enumerate(aggs):
            if ag[0] == "count_nonzero":
                sums[gid][k] += int(int(cols[ag[1]][r]) != 0)
                continue
            _, coef, fs = ag
            v = coef % M
            for i, (ci, fa, fb) in enumerate(fs):
                f = (fa + fb * int(cols[ci][r])) % M
                v = f if (i == 0 and coef == 1) else (v * f) % M
            sums[gid][k] += v - M if v >= M // 2 else v
    return cnt, sums, oor


@pytest.mark.parametrize("seed", range(4))
def test_oracle_vs_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 3000))
    cols = [rng.integers(-50, 50, n).astype(np.int64), rng.integers(0, 4, n).astype(np.uint8),
            rng.integers(-2**40, 2**40, n).astype(np.int64), rng.integers(0, 100, n).astype(np.int32)]
    preds = [(0, int(rng.integers(0, 6)), int(rng.integers(-30, 30)), True),
             (3, LT, int(rng.integers(10, 100)), False)]
    parts = [(1, 1, 0, 3), (3, 25, 0, 4)]
    aggs = [("sum", 1, [(2, 0, 1), (0, 7, -1)]), ("sum", 3, [(2, 5, 3), (0, 1, 1), (3, -2, 1)]),
            ("count_nonzero", 0)]
    b = int(rng.integers(0, n))
    e = int(rng.integers(b, n + 1))
    q = B.make_query(cols, preds, parts, aggs, b, e)
    cnt, sums, oor = B.group_agg(q)
    bc, bs, boor = brute_group_agg(cols, preds, parts, aggs, b, e)
    assert list(map(int, cnt)) == bc and sums == bs and oor == boor


def test_oracle_wraps_like_reference_binop():
    # per-row products wrap in 64 bits (expr.rs binop wrapping_mul) before the exact sum
    a = np.array([2**62, 2**62, 3], np.int64)
    q = B.make_query([a], aggs=[("sum", 1, [(0, 0, 1), (0, 0, 1)])])
    _, sums, _ = B.group_agg(q)
    assert sums[0][0] == 9  # 2^124 wraps to 0 twice, + 9


def test_golden_fixture_matches_oracle():
    fx = json.load(open(GOLDEN))
    n = fx["rows"]
    cols = W.host_fact_columns(n, B.synth_uniform)
    import hashlib
    for k, h in fx["column_sha256"].items():
        assert hashlib.sha256(cols[k].tobytes()).hexdigest() == h, k
    names, preds, parts, aggs = W.q1_spec()
    q = B.make_query([cols[c] for c in names], preds, parts, aggs)
    cnt, sums, _ = B.group_agg(q)
    _, f64, _ = B.group_agg(q, mode=1)
    for gid, g in enumerate(fx["q1"]["groups"]):
        assert int(cnt[gid]) == g["count"]
        assert sums[gid] == [int(g[k]) for k in ("sum_qty", "sum_price", "sum_disc_price",
                                                 "sum_charge")]
        # On this 1M-row table every per-group sum is < 2^53, so the reference's
        # f64 sums equal the exact sums (the survey's C1 parity claim).
        assert [float(x) for x in sums[gid]] == f64[gid] == g["ref_f64"]
    names6, preds6, parts6, aggs6 = W.q6_spec()
    c6, s6, _ = B.group_agg(B.make_query([cols[c] for c in names6], preds6, parts6, aggs6))
    assert int(c6[0]) == fx["q6"]["count"] and s6[0][0] == int(fx["q6"]["sum_price_disc"])


def test_generator_closed_form():
    # value = lo + floor(mix64(key ^ i) * range / 2^64), SplitMix64 finaliser
    def mix(z):
        M = 2**64 - 1
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)
    out = B.synth_uniform(6, 12345, 3, -7, 1000, 10, 20)
    key = 12345 ^ (3 << 40)
    assert [int(x) for x in out] == [-7 + ((mix(key ^ i) * 1008) >> 64) for i in range(10, 20)]
    assert B.lib().orc_mix64(0) == 0 and B.lib().orc_mix64(1) == mix(1)


def test_avx512_filter_equals_scalar():
    """The AVX-512 4x8 filter path (bitmap.rs:642-694) gives the scalar path's
    bitmap for every op x signedness, at lengths that end in each tail form."""
    lib = B.lib()
    if lib.orc_simd_level() < 512:
        pytest.skip("host without AVX-512F/BW: only the scalar path runs")
    rng = np.random.default_rng(5)
    edge = [0, 1, -1, 2 ** 63 - 1, -(2 ** 63), 12345]
    try:
        for n in (0, 1, 7, 8, 31, 32, 33, 100, 1031):
            cells = rng.choice(np.array(edge + list(rng.integers(-5, 5, 10)), np.int64), n)
            cells = cells.view(np.uint64)
            for op in range(6):
                for sg in (0, 1):
                    for lit in edge:
                        lib.orc_set_simd(512)
                        a = B.filter_cells(cells, op, sg, lit)
                        lib.orc_set_simd(0)
                        b = B.filter_cells(cells, op, sg, lit)
                        assert np.array_equal(a, b), (n, op, sg, lit)
    finally:
        lib.orc_set_simd(512)
