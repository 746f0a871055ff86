"""The workloads at their full sizes.

- Q1 and Q6 over 60M rows against the oracle; Q1 through every kernel form, and
  as the int128 sum of two half ranges (range additivity).
- The join of 2M build rows and 60M Zipf(1.1) probe rows: every foreign key
  matches exactly one build row, so GROUP BY the build payload equals a plain
  grouping of the probe rows by the payload of their key's build row, GROUP BY
  the 1.5M-value probe key equals a plain grouping of the probe rows, and both
  add up to the probe count and the column's total (checked with numpy, whose
  float64 group sums are exact here: every partial sum is an integer < 2^53).
- The 4096-query batch over the 256M-row table: the persistent launch against one
  launch per query, and a sample of its queries against the oracle on slices of
  the table generated on the host.
"""
import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import HashTable, make_agg_query, _capi as A
from turbogp_amd import workloads as W
from turbogp_amd.batch import Batch
from turbogp_amd.joinload import make_join_tables

pytestmark = pytest.mark.gpu
N60 = 60_000_000
N256 = 256_000_000


def same(r1, r2):
    return np.array_equal(r1.count, r2.count) and np.array_equal(r1.sums_words, r2.sums_words) \
        and r1.n_out_of_range == r2.n_out_of_range


@pytest.fixture(scope="module")
def fact60(ctx):
    table = W.make_fact_table(ctx, N60)
    yield table
    for c in table.values():
        c.free()


def test_q1_q6_60m_vs_oracle(ctx, fact60):
    host = W.host_fact_columns(N60, B.synth_uniform)
    counts = []
    for spec in (W.q1_spec(), W.q6_spec()):
        names, preds, parts, aggs = spec
        got = ctx.group_agg(W.build_query(fact60, spec))
        cnt, sums, oor = B.group_agg(B.make_query([host[c] for c in names], preds, parts, aggs))
        assert [int(x) for x in got.count] == [int(x) for x in cnt]
        assert got.sums() == sums
        assert got.n_out_of_range == oor
        counts.append(int(got.count.sum()))
    print(f"[full] 60M rows equal the oracle: Q1 {counts[0]} rows, Q6 {counts[1]} rows", flush=True)


def test_q1_60m_kernel_forms_and_halves(ctx, fact60):
    spec = W.q1_spec()
    q = W.build_query(fact60, spec)
    forms = []
    try:
        for specialize, form in ((1, 0), (1, 1), (1, 2), (0, 0)):
            ctx.tune(A.TUNE_SPECIALIZE, specialize)
            ctx.tune(A.TUNE_SPEC_KERNEL, form)
            forms.append(ctx.group_agg(q))
    finally:
        ctx.tune(A.TUNE_SPECIALIZE, 1)
        ctx.tune(A.TUNE_SPEC_KERNEL, 0)
    for f in forms[1:]:
        assert same(forms[0], f)
    h = N60 // 2 + 77                       # not a slice boundary
    lo = ctx.group_agg(W.build_query(fact60, spec, 0, h))
    hi = ctx.group_agg(W.build_query(fact60, spec, h, N60))
    assert np.array_equal(lo.count + hi.count, forms[0].count)
    assert [[a + b for a, b in zip(x, y)] for x, y in zip(lo.sums(), hi.sums())] == \
        forms[0].sums()


def test_join_2m_x_60m_totals(ctx):
    build, probe = make_join_tables(ctx)
    t = HashTable(ctx, build["bkey"], build["pay"])
    try:
        n = W.JOIN_PROBE_ROWS
        matches = t.probe_count(probe["fk"])
        by_pay = t.probe_agg(probe["fk"], probe["val"], A.JOIN_GROUP_BUILD_PAYLOAD,
                             W.JOIN_PAY_CARD)
        by_key = t.probe_agg(probe["fk"], probe["val"], A.JOIN_GROUP_PROBE_COL, W.JOIN_GKEY_CARD,
                             0, probe["gkey"])
        total = ctx.group_agg(make_agg_query([probe["val"]], [], [], [("sum", 1, [(0, 0, 1)])]))
        bkey = build["bkey"].to_numpy()
        pay = build["pay"].to_numpy()
        fk = probe["fk"].to_numpy()
        val = probe["val"].to_numpy()
        gkey = probe["gkey"].to_numpy()
    finally:
        t.free()
        for c in list(build.values()) + list(probe.values()):
            c.free()
    assert matches == n
    assert int(by_pay.count.sum()) == n and int(by_key.count.sum()) == n
    assert by_pay.n_out_of_range == 0 and by_key.n_out_of_range == 0
    want_total = int(val.sum())              # < 2^63: 60M values below 10.5M
    assert total.sums()[0][0] == want_total
    assert sum(s[0] for s in by_pay.sums()) == want_total
    assert not by_key.sums_words[:, 0, 1].any()          # every group sum is below 2^64
    assert int(by_key.sums_words[:, 0, 0].astype(object).sum()) == want_total
    # each foreign key's build row: the join's payload grouping in closed form
    order = np.argsort(bkey)
    at = order[np.searchsorted(bkey[order], fk)]
    assert np.array_equal(bkey[at], fk)                  # every foreign key is a build key
    pg = pay[at]
    assert np.array_equal(by_pay.count, np.bincount(pg, minlength=W.JOIN_PAY_CARD))
    want_pay = np.bincount(pg, weights=val, minlength=W.JOIN_PAY_CARD)
    assert [s[0] for s in by_pay.sums()] == [int(x) for x in want_pay]
    assert np.array_equal(by_key.count, np.bincount(gkey, minlength=W.JOIN_GKEY_CARD))
    want_key = np.bincount(gkey, weights=val, minlength=W.JOIN_GKEY_CARD).astype(np.uint64)
    assert np.array_equal(by_key.sums_words[:, 0, 0], want_key)
    print(f"[full] join 2M x 60M: {matches} matches, total {want_total}", flush=True)


def test_batch_4096_queries_256m_rows(ctx):
    table = W.make_fact_table(ctx, N256)
    p = W.batch_query_params(4096, N256)
    b = Batch(ctx, len(p))
    try:
        b.set_queries(W.batch_queries(table, p))
        assert b.info()["batched"] == len(p)
        b.run(A.BATCH_PERSISTENT)
        batched = b.results()
        b.run(A.BATCH_PER_QUERY)
        per_query = b.results()
    finally:
        b.close()
        for c in table.values():
            c.free()
    for i in range(len(p)):
        assert same(batched[i], per_query[i]), i
    # the oracle on the shortest queries of each group key and the longest query,
    # over slices of the table generated on the host
    short = sorted(range(len(p)), key=lambda i: p[i][1] - p[i][0])
    sample = [short[-1]]
    for g in ("ab", "a", "b", "month"):
        sample += [i for i in short if p[i][3] == g][:3]
    for i in sample:
        bb, e, c, g = p[i]
        names, preds, parts, aggs = W.q1_spec(date_cutoff=c, group=g)
        cols = []
        for name in names:
            cid, dt, lo, hi = W.FACT_SCHEMA[name]
            cols.append(B.synth_uniform(dt, W.SEED, cid, lo, hi, bb, e))
        cnt, sums, oor = B.group_agg(B.make_query(cols, preds, parts, aggs, 0, e - bb))
        assert [int(x) for x in batched[i].count] == [int(x) for x in cnt], i
        assert batched[i].sums() == sums, i
        assert batched[i].n_out_of_range == oor, i
    print(f"[full] batch 4096 x 256M: identical to one launch per query; "
          f"{len(sample)} queries equal the oracle", flush=True)
