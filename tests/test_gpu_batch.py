"""Batched executor: every query's result from the persistent batch launch
(graph and plain), from one launch per query, from a single tgp_group_agg call
and from the oracle must agree bit for bit."""
import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import DeviceColumn, make_agg_query, _capi as A
from turbogp_amd import workloads as W
from turbogp_amd.batch import Batch

pytestmark = pytest.mark.gpu
N = 3_000_017


@pytest.fixture(scope="module")
def table(ctx):
    host = W.host_fact_columns(N, B.synth_uniform)
    dev = {k: DeviceColumn.from_numpy(ctx, v) for k, v in host.items()}
    yield host, dev
    for c in dev.values():
        c.free()


def params(n):
    p = W.batch_query_params(n, N, lmin=1, lmax=N)
    # edges: empty, single rows, whole table, the last rows, an unaligned short range
    p += [(5, 5, 100, "ab"), (0, 1, 2525, "a"), (N - 1, N, 2525, "b"), (0, N, 1200, "month"),
          (N - 300, N, 2000, "ab"), (77, 333, 1500, "month")]
    return p


def same(r1, r2):
    return np.array_equal(r1.count, r2.count) and np.array_equal(r1.sums_words, r2.sums_words) \
        and r1.n_out_of_range == r2.n_out_of_range


def test_batch_matches_single_queries_and_oracle(ctx, table):
    host, dev = table
    p = params(300)
    qs = W.batch_queries(dev, p)
    # one query with a shape that has no batch kernel: a generic item of the launch
    names6, preds6, parts6, aggs6 = W.q6_spec()
    qs.append(make_agg_query([dev[c] for c in names6], preds6 + [(2, A.NE, 7, True)], parts6,
                             aggs6 + [("count_nonzero", 1)], 10, 2_000_000))
    # MIN/MAX aggregates: generic items too
    names1, preds1, parts1, aggs1 = W.q1_spec()
    qs.append(make_agg_query([dev[c] for c in names1], preds1, parts1,
                             [("min", 0), ("max_u64", 1), ("max", 1, [(1, 0, 1), (2, 100, -1)])],
                             999, 2_500_000))
    b = Batch(ctx, len(qs))
    try:
        b.set_queries(qs)
        info = b.info()
        assert info["batched"] == len(qs) and info["fallback"] == 0
        b.run(A.BATCH_PERSISTENT)
        persistent = b.results()
        b.run(A.BATCH_PERSISTENT)          # graph replay
        replay = b.results()
        single = [ctx.group_agg(q) for q in qs[:40]]   # eager work between replays
        b.run(A.BATCH_PERSISTENT)
        replay2 = b.results()
        b.run(A.BATCH_PER_QUERY)
        per_query = b.results()
        b.set_tuning(0, use_graph=False)
        b.run(A.BATCH_PERSISTENT)
        plain = b.results()
        b.set_tuning(3, use_graph=True, chunk_slices=5, order=1, cached_loads=True)
        b.run(A.BATCH_PERSISTENT)
        pipe1 = b.results()
        b.set_tuning(0, use_graph=True, chunk_slices=1, order=1, cached_loads=False)
        b.run(A.BATCH_PERSISTENT)
        table1 = b.results()
        b.set_tuning(2, use_graph=True, chunk_slices=3, order=2, cached_loads=True,
                     items_per_grab=5)
        b.run(A.BATCH_PERSISTENT)
        table2 = b.results()
        b.set_tuning(2, use_graph=True, chunk_slices=1, order=3, cached_loads=True)
        b.run(A.BATCH_PERSISTENT)
        shared1 = b.results()
        b.set_tuning(0, use_graph=True, chunk_slices=37, order=3, cached_loads=False)
        b.run(A.BATCH_PERSISTENT)
        b.run(A.BATCH_PERSISTENT)     # replay: the shared scan's counters start at zero
        shared2 = b.results()
    finally:
        b.close()
    for i in range(len(qs)):
        assert same(persistent[i], per_query[i]), i
        assert same(persistent[i], replay[i]) and same(persistent[i], replay2[i]), i
        assert same(persistent[i], plain[i]) and same(persistent[i], pipe1[i]), i
        assert same(persistent[i], table1[i]) and same(persistent[i], table2[i]), i
        assert same(persistent[i], shared1[i]) and same(persistent[i], shared2[i]), i
    for i in range(40):
        assert same(persistent[i], single[i]), i
    # oracle on a subset (all edge queries included)
    for i in list(range(0, 300, 23)) + list(range(300, len(p))):
        bb, e, c, g = p[i]
        names, preds, parts, aggs = W.q1_spec(date_cutoff=c, group=g)
        cnt, sums, oor = B.group_agg(B.make_query([host[x] for x in names], preds, parts, aggs,
                                                  bb, e))
        assert [int(x) for x in persistent[i].count] == [int(x) for x in cnt], i
        assert persistent[i].sums() == sums, i
        assert persistent[i].n_out_of_range == oor


def test_batch_flush_every_2_20_rows_with_a_tiny_grid(ctx, table):
    host, dev = table
    p = [(0, N, 2400, "ab"), (1, N - 1, 1000, "month"), (100, 2_500_000, 2525, "a")]
    qs = W.batch_queries(dev, p)
    b = Batch(ctx, len(qs))
    try:
        b.set_queries(qs)
        ctx.tune(A.TUNE_GRID_MAX, 2)   # 8 waves: each wave adds millions of rows
        got = []
        for order, chunk in ((0, 65536), (1, 7000), (2, 4500), (3, 65536), (3, 4096)):
            b.set_tuning(0, use_graph=False, chunk_slices=chunk, order=order)
            b.run(A.BATCH_PERSISTENT)
            got.append(b.results())
    finally:
        ctx.tune(A.TUNE_GRID_MAX, 0)
        b.close()
    for i, q in enumerate(qs):
        want = ctx.group_agg(q)
        assert all(same(g[i], want) for g in got), i


def test_batch_autotune_keeps_results(ctx, table):
    _, dev = table
    p = params(100)
    qs = W.batch_queries(dev, p)
    b = Batch(ctx, len(qs))
    try:
        b.set_queries(qs)
        b.run()
        before = b.results()
        best, ms = b.autotune()
        assert 0 <= best["wg_per_cu"] <= 4 and best["order"] in (0, 1, 2, 3) and ms > 0
        assert best["chunk_slices"] in (4, 8, 16, 32, 64, 128, 256) and b.tuning() == best
        b.run()
        after = b.results()
        tot, per = b.run_timed(A.BATCH_PERSISTENT)
        assert tot > 0 and np.all(per == np.float32(tot))
        tot2, per2 = b.run_timed(A.BATCH_PER_QUERY)
        assert np.all(np.diff(per2) >= 0) and abs(per2[-1] - tot2) < 1e-3 * max(1, tot2)
    finally:
        b.close()
    assert all(same(x, y) for x, y in zip(before, after))


def test_shared_scan_with_non_members_and_big_values(ctx, table):
    """Order 3 runs the family (Q1-layout queries over one set of columns with one set
    of aggregate arguments) in the shared scan and every other batched query in the
    table-ordered kernel of the same launch: queries over a second copy of the table,
    a month key with another divisor, and other aggregate coefficients are not
    members.  A copy of the table with prices near 2^50 sends its slices' sums down
    the exact int128 path.  All equal single-query launches."""
    host, dev = table
    n2 = 400_000
    big = {k: v[:n2].copy() for k, v in host.items()}
    big["price"] = big["price"] + (1 << 50)
    dev2 = {k: DeviceColumn.from_numpy(ctx, v) for k, v in big.items()}
    try:
        qs = W.batch_queries(dev, params(120))
        first_big = len(qs)
        qs += W.batch_queries(dev2, [(0, n2, 2000, "ab"), (1000, 300_000, 900, "month"),
                                     (5, 6, 2525, "a")])
        names, preds, parts, aggs = W.q1_spec(date_cutoff=1800, group="month")
        parts7 = [(parts[0][0], 31, 0, 82)]           # dt / 31: another divisor
        qs.append(make_agg_query([dev[c] for c in names], preds, parts7, aggs, 50, 2_000_000))
        aggs3 = list(aggs)
        aggs3[3] = ("sum", 1, [(2, 0, 1), (3, 99, -1), (4, 100, 1)])  # 99 - disc
        qs.append(make_agg_query([dev[c] for c in names], preds, parts, aggs3, 0, 1_000_000))
        b = Batch(ctx, len(qs))
        try:
            b.set_queries(qs)
            assert b.info()["fallback"] == 0
            got = []
            for chunk in (2, 32, 4096):
                b.set_tuning(0, use_graph=True, chunk_slices=chunk, order=3, cached_loads=True)
                b.run(A.BATCH_PERSISTENT)
                got.append(b.results())
        finally:
            b.close()
        for i, q in enumerate(qs):
            want = ctx.group_agg(q)
            assert all(same(g[i], want) for g in got), i
        # the oracle on the big-valued queries
        for j, (bb, e, c, g) in enumerate([(0, n2, 2000, "ab"), (1000, 300_000, 900, "month"),
                                           (5, 6, 2525, "a")]):
            nm, pr, pa, ag = W.q1_spec(date_cutoff=c, group=g)
            cnt, sums, oor = B.group_agg(B.make_query([big[x] for x in nm], pr, pa, ag, bb, e))
            r = got[1][first_big + j]
            assert [int(x) for x in r.count] == [int(x) for x in cnt] and r.sums() == sums
    finally:
        for c in dev2.values():
            c.free()


def test_shared_scan_signature_groups(ctx, table):
    """Queries with equal date bounds and key parts (one signature) share partial
    aggregates of the chunks they cover whole, each taking the running-sum
    difference over its own chunks and scanning its edge chunks row by row.  Many
    queries per signature over the whole table, random ranges, chunk-aligned and
    one-row ranges, the table's last rows, empty ranges, out-of-range keys and a
    table copy with prices near 2^50 (int128 partials): every result equals its
    single-query launch, for several chunk sizes and graph replays."""
    host, dev = table
    n2 = 400_000
    big = {k: v[:n2].copy() for k, v in host.items()}
    big["price"] = big["price"] + (1 << 50)
    dev2 = {k: DeviceColumn.from_numpy(ctx, v) for k, v in big.items()}
    rng = np.random.default_rng(31)

    def ranges(n, k):
        out = [(0, n), (0, n), (0, n), (n - 300, n), (5, 5), (1000, 1001), (0, 256 * 64),
               (256 * 32, 256 * 96), (256 * 7 + 3, n - 1)]
        for _ in range(k):
            L = int(np.exp(rng.uniform(0, np.log(n))))
            bb = int(rng.integers(0, n - L + 1))
            out.append((bb, bb + L))
        return out

    try:
        qs = []
        for c, g in ((2000, "ab"), (1500, "month"), (2525, "a")):
            qs += W.batch_queries(dev, [(bb, e, c, g) for bb, e in ranges(N, 25)])
        # out-of-range keys: flaga has three values, the key part takes two
        names, preds, parts, aggs = W.q1_spec(date_cutoff=2200, group="ab")
        narrow = [parts[0][:3] + (2,), parts[1]]
        for bb, e in ranges(N, 8):
            qs.append(make_agg_query([dev[x] for x in names], preds, narrow, aggs, bb, e))
        qs += W.batch_queries(dev2, [(bb, e, 2000, "ab") for bb, e in ranges(n2, 8)])
        b = Batch(ctx, len(qs))
        try:
            b.set_queries(qs)
            assert b.info()["fallback"] == 0
            got = []
            for chunk in (1, 3, 32, 64, 4096):
                b.set_tuning(0, use_graph=True, chunk_slices=chunk, order=3, cached_loads=True)
                b.run(A.BATCH_PERSISTENT)
                got.append(b.results())
                b.run(A.BATCH_PERSISTENT)  # graph replay: the partials start at zero again
                got.append(b.results())
        finally:
            b.close()
        for i, q in enumerate(qs):
            want = ctx.group_agg(q)
            assert all(same(g[i], want) for g in got), i
        assert any(r.n_out_of_range for r in got[0])
    finally:
        for c in dev2.values():
            c.free()


def test_shared_scan_first_value_packing(ctx, table):
    """The shared scan adds the row count and small first values (qty in [0, 2^16))
    with one atomic; slices whose first value is wider (at least 2^16, negative, or
    at least 2^43, the int128 path) add it on its own.  A table copy with each kind
    of qty block, in ranges that cross the blocks, equals single-query launches and
    the oracle."""
    host, dev = table
    n2 = 400_000
    wide = {k: v[:n2].copy() for k, v in host.items()}
    q = wide["qty"]
    q[:100_000] += 70_000
    q[100_000:200_000] *= -1
    q[300_000:] += 1 << 44
    q[250_000:250_300] += 1 << 16          # one slice-sized run just past the packed range
    dev2 = {k: DeviceColumn.from_numpy(ctx, v) for k, v in wide.items()}
    try:
        spec = [(0, n2, 2000, "ab"), (50_000, 150_000, 2525, "a"), (190_000, 310_000, 1500, "b"),
                (240_000, 260_000, 2400, "month"), (0, n2, 2525, ""), (99_000, 101_000, 2525, "ab")]
        qs = W.batch_queries(dev2, spec) + W.batch_queries(dev2, spec)  # twice: a signature group
        b = Batch(ctx, len(qs))
        try:
            b.set_queries(qs)
            got = []
            for chunk in (1, 16, 4096):
                b.set_tuning(0, use_graph=True, chunk_slices=chunk, order=3, cached_loads=True)
                b.run(A.BATCH_PERSISTENT)
                got.append(b.results())
        finally:
            b.close()
        for i, qq in enumerate(qs):
            want = ctx.group_agg(qq)
            assert all(same(g[i], want) for g in got), i
        for j, (bb, e, c, g) in enumerate(spec):
            nm, pr, pa, ag = W.q1_spec(date_cutoff=c, group=g)
            cnt, sums, oor = B.group_agg(B.make_query([wide[x] for x in nm], pr, pa, ag, bb, e))
            r = got[0][j]
            assert [int(x) for x in r.count] == [int(x) for x in cnt] and r.sums() == sums, j
    finally:
        for c in dev2.values():
            c.free()


def test_mixed_shape_batch_has_no_fallback(ctx, table):
    """A 4096-query batch of every query form: the six compiled shapes, Q6 with a fifth
    predicate, MIN / MAX, an input selection bitmap, a price column of another width,
    group counts beyond the LDS table (straight into the result) and empty ranges.
    All run in the batch launch (no query runs alone) and every result equals its
    one-launch-per-query result, bit for bit; a sample equals the oracle."""
    from turbogp_amd import DeviceBitmap
    host, dev = table
    rng = np.random.default_rng(5)
    price32 = DeviceColumn.from_numpy(ctx, host["price"].astype(np.int32))
    words = np.zeros((N + 63) // 64, np.uint64)
    bits = rng.random(N) < 0.3
    idx = np.flatnonzero(bits)
    np.bitwise_or.at(words, idx >> 6, np.left_shift(np.uint64(1), (idx & 63).astype(np.uint64)))
    names1, preds1, parts1, aggs1 = W.q1_spec()
    names6, preds6, parts6, aggs6 = W.q6_spec()
    specs, hostq = [], []
    for k in range(4096):
        L = int(np.exp(rng.uniform(0, np.log(N))))
        b0 = int(rng.integers(0, N - L + 1))
        e0 = b0 + L if k % 97 else b0          # some empty ranges
        form = k % 8
        if form in (0, 1):
            c, g = int(rng.integers(0, 2526)), ["ab", "a", "b", "month", ""][k % 5]
            nm, pr, pa, ag = W.q1_spec(date_cutoff=c, group=g)
            cols, sel = [dev[x] for x in nm], None
        elif form == 2:
            nm, pr, pa, ag = names6, preds6, parts6, aggs6
            cols, sel = [dev[x] for x in nm], None
        elif form == 3:                        # five predicates
            nm, pr, pa, ag = names6, preds6 + [(2, A.NE, int(rng.integers(1, 50)), True)], \
                parts6, aggs6
            cols, sel = [dev[x] for x in nm], None
        elif form == 4:                        # MIN / MAX next to sums
            nm, pr, pa = names1, preds1, parts1
            ag = [("min", 2), ("max_u64", 1), ("sum", 1, [(2, 0, 1)]),
                  ("max", 1, [(1, 0, 1), (3, 100, -1)])]
            cols, sel = [dev[x] for x in nm], None
        elif form == 5:                        # an input selection bitmap over [0, N)
            nm, pr, pa, ag = names1, preds1, parts1, aggs1
            cols, sel, b0, e0 = [dev[x] for x in nm], "bitmap", 0, N
        elif form == 6:                        # price as int32: another column layout
            nm, pr, pa, ag = names1, preds1, parts1, aggs1
            cols = [dev[x] if x != "price" else price32 for x in nm]
            sel = None
        else:                                  # 2526 x 2 groups: beyond the LDS table
            nm, pr, ag = names1, preds1, aggs1
            pa = [(0, 1, 0, 2526), (6, 1, 0, 2)]
            cols, sel = [dev[x] for x in nm], None
        specs.append((cols, pr, pa, ag, b0, e0, sel))
        hostq.append((nm, pr, pa, ag, b0, e0, sel, form))
    bm = DeviceBitmap(ctx, N)
    bm.buf.upload(words)
    try:
        qs = [make_agg_query(c, pr, pa, ag, b0, e0, bm if sel else None)
              for c, pr, pa, ag, b0, e0, sel in specs]
        b = Batch(ctx, len(qs))
        try:
            b.set_queries(qs)
            info = b.info()
            assert info["batched"] == len(qs) and info["fallback"] == 0, info
            b.set_tuning(0, use_graph=True, chunk_slices=64, order=3, cached_loads=True)
            b.run(A.BATCH_PERSISTENT)
            batched = b.results()
            b.run(A.BATCH_PERSISTENT)
            replay = b.results()
            b.run(A.BATCH_PER_QUERY)
            single = b.results()
        finally:
            b.close()
        for i in range(len(qs)):
            assert same(batched[i], single[i]) and same(replay[i], single[i]), (i, hostq[i][7])
        # the same batch with the specialised shapes compiled at planning: those queries
        # run their compiled kernels inside the batch's graph
        ctx.tune(A.TUNE_JIT, 2)
        b = Batch(ctx, len(qs))
        try:
            b.set_queries(qs)
            ctx.tune(A.TUNE_JIT, 0)
            assert b.info()["fallback"] == 0
            b.set_tuning(0, use_graph=True, chunk_slices=64, order=3, cached_loads=True)
            b.run(A.BATCH_PERSISTENT)
            compiled = b.results()
            b.run(A.BATCH_PERSISTENT)
            compiled_replay = b.results()
        finally:
            ctx.tune(A.TUNE_JIT, 0)
            b.close()
        for i in range(len(qs)):
            assert same(compiled[i], single[i]) and same(compiled_replay[i], single[i]), \
                (i, hostq[i][7])
        for i in list(range(0, 64, 3)) + [8 * j + 7 for j in range(3)]:
            nm, pr, pa, ag, b0, e0, sel, form = hostq[i]
            arrays = [host[x].astype(np.int32) if (form == 6 and x == "price") else host[x]
                      for x in nm]
            cnt, sums, oor = B.group_agg(B.make_query(arrays, pr, pa, ag, b0, e0,
                                                      words if sel else None))
            r = batched[i]
            assert [int(x) for x in r.count] == [int(x) for x in cnt], (i, form)
            assert r.sums() == sums and r.n_out_of_range == oor, (i, form)
    finally:
        bm.free()
        price32.free()


def test_shared_scan_many_members_per_chunk_and_wide_bounds(ctx, table):
    """150 family members over the same few chunks (more than one per lane of a wave:
    the planner splits each chunk's members into items of at most 64), key parts
    shifted so some keys fall out of range, and date bounds beyond int32, against one
    launch per query, a graph replay and the oracle."""
    host, dev = table
    rng = np.random.default_rng(11)
    specs = []
    for i in range(150):
        names, preds, parts, aggs = W.q1_spec(int(rng.integers(0, 2526)),
                                              ("a", "b", "ab", "month")[i % 4])
        if i % 5 == 1:
            preds = [(preds[0][0], A.LE, 2 ** 40, False)]
        if i % 5 == 2:
            preds = preds + [(preds[0][0], A.GE, -2 ** 40, False)]
        if i % 7 == 3:
            parts = [(c, d, base + 1, card) for (c, d, base, card) in parts]
        b0 = int(rng.integers(0, 40_000))
        specs.append((names, preds, parts, aggs, b0, b0 + int(rng.integers(100_000, 300_000))))
    qs = [make_agg_query([dev[c] for c in nm], pr, pa, ag, b0, e0)
          for (nm, pr, pa, ag, b0, e0) in specs]
    b = Batch(ctx, len(qs))
    try:
        b.set_queries(qs)
        assert b.info()["fallback"] == 0
        b.set_tuning(0, use_graph=True, chunk_slices=128, order=3, cached_loads=True)
        b.run(A.BATCH_PERSISTENT)
        got = b.results()
        b.run(A.BATCH_PERSISTENT)  # graph replay
        replay = b.results()
        b.run(A.BATCH_PER_QUERY)
        ref = b.results()
    finally:
        b.close()
    for i in range(len(qs)):
        assert same(got[i], ref[i]) and same(got[i], replay[i]), i
    assert any(r.n_out_of_range for r in got)
    for i in range(0, len(qs), 7):
        nm, pr, pa, ag, b0, e0 = specs[i]
        cnt, sums, oor = B.group_agg(B.make_query([host[x] for x in nm], pr, pa, ag, b0, e0))
        assert [int(x) for x in got[i].count] == [int(x) for x in cnt], i
        assert got[i].sums() == sums, i
        assert got[i].n_out_of_range == oor, i
