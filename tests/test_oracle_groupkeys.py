"""The oracle's GROUP BY over key tuples (orc_group_by_keys) pinned on the CPU:
the reference's group-by known answers, brute force over Python dicts (NULL keys,
predicates, negative values, MIN/MAX), equality with the dense-id oracle where the
keys are dense, and tuples that the reference's combined hash would merge."""
import numpy as np

from oracle import binding as B

KEY_MULT = 0x517CC1B727220A95  # accum_grouped.rs:217, dispatch.rs:1381


def test_group_by_sum_count_kat():
    # tests/joins_groupby_orderby.rs:82-106: grp 1 -> sum 60, count 3; grp 2 -> sum 90, count 2
    grp = np.array([1, 1, 1, 2, 2], np.int64)
    a = np.array([10, 20, 30, 40, 50], np.int64)
    q = B.make_query([grp, a], aggs=[("sum", 1, [(1, 0, 1)])])
    assert B.group_by_keys(q, [0]) == {(1,): (3, [60]), (2,): (2, [90])}


def test_emits_every_group_kat():
    # tests/joins_groupby_orderby.rs:64-78: (1, 10, 100), (1, 20, 200), (2, 30, 300) -> 2 groups
    grp = np.array([1, 1, 2], np.int64)
    a = np.array([10, 20, 30], np.int64)
    b = np.array([100, 200, 300], np.int64)
    q = B.make_query([grp, a, b], aggs=[("sum", 1, [(1, 0, 1)]), ("sum", 1, [(2, 0, 1)])])
    got = B.group_by_keys(q, [0])
    assert got == {(1,): (2, [30, 300]), (2,): (1, [30, 300])}
    # avg(b) per group: 150 and 300
    assert [s[1] // c for c, s in got.values()] == [150, 300]


def _brute(cols, preds, keys, nulls, aggs, n):
    """Python dict grouping by the key tuple; wrapping int64 row values."""
    def wrap(v):
        v &= (1 << 64) - 1
        return v - (1 << 64) if v >> 63 else v

    ops = [lambda x, l: x == l, lambda x, l: x != l, lambda x, l: x < l,
           lambda x, l: x <= l, lambda x, l: x > l, lambda x, l: x >= l]
    out = {}
    for i in range(n):
        if not all(ops[op](int(cols[c][i]), lit) for c, op, lit, _ in preds):
            continue
        t = tuple(None if (nulls[k] is not None and nulls[k][i]) else int(cols[c][i])
                  for k, c in enumerate(keys))
        cnt, vals = out.get(t, (0, [[] for _ in aggs]))
        for j, ag in enumerate(aggs):
            if ag[0] == "sum":
                v = ag[1]
                for c, fa, fb in ag[2]:
                    v = wrap(v * (fa + fb * int(cols[c][i])))
                vals[j].append(v)
            else:
                vals[j].append(int(cols[ag[1]][i]))
        out[t] = (cnt + 1, vals)
    res = {}
    for t, (cnt, vals) in out.items():
        r = []
        for j, ag in enumerate(aggs):
            r.append(sum(vals[j]) if ag[0] == "sum" else
                     (min(vals[j]) if ag[0] == "min" else max(vals[j])))
        res[t] = (cnt, r)
    return res


def _words(flags):
    n = len(flags)
    w = np.zeros((n + 63) // 64, np.uint64)
    for i in np.flatnonzero(flags):
        w[i >> 6] |= np.uint64(1) << np.uint64(i & 63)
    return w


def test_brute_force_with_nulls_and_predicates():
    rng = np.random.default_rng(7)
    for trial in range(12):
        n = int(rng.integers(1, 600))
        k0 = rng.choice(np.array([-(2 ** 62), -5, 0, 7, 2 ** 61, 2 ** 62 + 3], np.int64), n)
        k1 = rng.integers(-3, 3, n).astype(np.int32)
        v = rng.integers(-(2 ** 40), 2 ** 40, n).astype(np.int64)
        w = rng.integers(0, 100, n).astype(np.int64)
        cols = [k0, k1, v, w]
        nk = 1 + trial % 2
        nulls = [rng.random(n) < 0.2 if trial % 3 else None, rng.random(n) < 0.3]
        preds = [(3, 4, 20, True)] if trial % 4 else []
        aggs = [("sum", 1, [(2, 0, 1)]), ("sum", 3, [(2, 0, 1), (3, 100, -1)]), ("min", 2),
                ("max", 2)]
        q = B.make_query(cols, preds, aggs=aggs)
        keys = [0, 1][:nk]
        kn = [None if f is None else _words(f) for f in nulls[:nk]]
        got = B.group_by_keys(q, keys, kn)
        want = _brute(cols, preds, keys, nulls[:nk], aggs, n)
        assert got == want, trial


def test_equals_dense_oracle_on_dense_keys():
    rng = np.random.default_rng(3)
    n = 20000
    a = rng.integers(0, 3, n).astype(np.uint8)
    b = rng.integers(10, 14, n).astype(np.int32)
    v = rng.integers(0, 10 ** 7, n).astype(np.int64)
    aggs = [("sum", 1, [(2, 0, 1)]), ("sum", 1, [(2, 0, 1), (2, 0, 1)])]
    qd = B.make_query([a, b, v], [(2, 2, 9 * 10 ** 6, True)], [(0, 1, 0, 3), (1, 1, 10, 4)], aggs)
    cnt, sums, oor = B.group_agg(qd)
    assert oor == 0
    dense = {(g % 3, 10 + g // 3): (int(cnt[g]), sums[g]) for g in range(12) if cnt[g]}
    qk = B.make_query([a, b, v], [(2, 2, 9 * 10 ** 6, True)], aggs=aggs)
    assert B.group_by_keys(qk, [0, 1]) == dense


def test_tuples_with_one_combined_hash_stay_apart():
    # h = h * KEY_MULT + cell (wrapping): (0, KEY_MULT) and (1, 0) give the same h,
    # which the reference's accumulator would merge into one group
    def combined(t):
        h = 0
        for c in t:
            h = (h * KEY_MULT + c) % (1 << 64)
        return h

    t1, t2 = (0, KEY_MULT), (1, 0)
    assert combined(t1) == combined(t2)
    k0 = np.array([0, 1, 0, 1, 1], np.uint64)
    k1 = np.array([KEY_MULT, 0, KEY_MULT, 0, 0], np.uint64)
    v = np.array([1, 10, 100, 1000, 10000], np.int64)
    q = B.make_query([k0, k1, v], aggs=[("sum", 1, [(2, 0, 1)])])
    got = B.group_by_keys(q, [0, 1], signed=[False, False])
    assert got == {t1: (2, [101]), t2: (3, [11010])}


def test_int128_sums_past_64_bits():
    n = 4000
    k = np.zeros(n, np.int64)
    v = np.full(n, 2 ** 62 + 12345, np.int64)
    v[::2] = -(2 ** 62)
    q = B.make_query([k, v], aggs=[("sum", 1, [(1, 0, 1)]), ("sum", 3, [(1, 0, 1)])])
    s = 2000 * (2 ** 62 + 12345) - 2000 * 2 ** 62
    assert B.group_by_keys(q, [0]) == {(0,): (n, [s, 3 * s])}
    v[:] = 2 ** 62
    q = B.make_query([k, v], aggs=[("sum", 1, [(1, 0, 1)])])
    assert B.group_by_keys(q, [0]) == {(0,): (n, [n * 2 ** 62])}  # 2^74: past 64 bits
