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


def _batch_results(ctx, qs, tunings):
    from turbogp_amd.batch import Batch
    b = Batch(ctx, len(qs))
    out = []
    try:
        b.set_queries(qs)
        for wg, chunk, order in tunings:
            b.set_tuning(wg, use_graph=False, chunk_slices=chunk, order=order, cached_loads=True)
            b.run(A.BATCH_PERSISTENT)
            out.append(b.results())
    finally:
        b.close()
    return out


def test_big_value_boundary_2_43(ctx):
    """Values on both sides of the 2^43 bound of the 64-bit LDS partials: slices of
    -2^43 and 2^43 - 1 only (the LDS path), slices holding 2^43 or -2^43 - 1 (the exact
    int128 path), and slices mixing them, in qty (summed alone) and in price (summed
    alone and in products that reach 2^50), through every single-query form, the
    batch kernels and the oracle."""
    rng = np.random.default_rng(43)
    n = 100_003
    host = W.host_fact_columns(n, B.synth_uniform)
    small = np.array([-(2 ** 43), 2 ** 43 - 1], np.int64)
    big = np.array([2 ** 43, -(2 ** 43) - 1], np.int64)
    for col in ("qty", "price"):
        v = host[col].copy()
        for s0 in range(0, n, 256):
            kind = (s0 // 256) % 4
            m = min(256, n - s0)
            if kind == 0:
                v[s0:s0 + m] = rng.choice(small, m)
            elif kind == 1:
                v[s0:s0 + m] = rng.choice(big, m)
            elif kind == 2:               # one exact-path value among LDS-path ones
                v[s0:s0 + m] = rng.choice(small, m)
                v[s0 + int(rng.integers(0, m))] = rng.choice(big)
        host[col] = v
    dev = {k: DeviceColumn.from_numpy(ctx, v) for k, v in host.items()}
    try:
        specs = [W.q1_spec(group=g) for g in ("ab", "month", "")]
        for spec in specs:
            both_paths(ctx, host, dev, spec)
        qs = [W.build_query(dev, s, b, e) for s in specs for b, e in ((0, n), (300, n - 77))]
        for got in _batch_results(ctx, qs, [(0, 1, 3), (0, 64, 3), (0, 8, 2), (0, 8, 0)]):
            for i, s in enumerate(s for s in specs for _ in range(2)):
                b, e = ((0, n), (300, n - 77))[i % 2]
                cnt, sums, oor = B.group_agg(B.make_query([host[c] for c in s[0]], s[1], s[2],
                                                          s[3], b, e))
                assert [int(x) for x in got[i].count] == [int(x) for x in cnt], i
                assert got[i].sums() == sums, i
    finally:
        for c in dev.values():
            c.free()


def test_flush_interval_with_values_just_below_2_43(ctx):
    """4.5M rows of qty = 2^43 - 1: one wave's 64-bit partial of 2^20 such rows is
    2^63 - 2^20, the last value below 2^63, so a flush later than every 2^20 rows would
    wrap.  One workgroup (static slices, then 4096-slice chunks) and the batch kernels
    with 2^20-row chunks must equal the oracle."""
    n = 4_500_000
    host = W.host_fact_columns(n, B.synth_uniform)
    host["qty"] = np.full(n, 2 ** 43 - 1, np.int64)
    dev = {k: DeviceColumn.from_numpy(ctx, v) for k, v in host.items()}
    try:
        spec = W.q1_spec(date_cutoff=2525, group="")
        names, preds, parts, aggs = spec
        cnt, sums, oor = B.group_agg(B.make_query([host[c] for c in names], preds, parts, aggs))
        q = W.build_query(dev, spec)
        got = []
        try:
            ctx.tune(A.TUNE_SPEC_KERNEL, 2)
            ctx.tune(A.TUNE_GRID_MAX, 1)
            got.append(ctx.group_agg(q))
            ctx.tune(A.TUNE_CHUNK, 4096)
            got.append(ctx.group_agg(q))
        finally:
            ctx.tune(A.TUNE_GRID_MAX, 0)
            ctx.tune(A.TUNE_CHUNK, 0)
            ctx.tune(A.TUNE_SPEC_KERNEL, 0)
        got += [r[0] for r in _batch_results(ctx, [q], [(1, 4096, 3), (1, 4096, 2), (1, 4096, 0)])]
        for r in got:
            assert [int(x) for x in r.count] == [int(x) for x in cnt]
            assert r.sums() == sums
        assert sums[0][0] == int(cnt[0]) * (2 ** 43 - 1)
    finally:
        for c in dev.values():
            c.free()
