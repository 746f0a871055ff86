"""The serving mix at full size: a 256M-row fact table and the 2M x 60M join tables.
Every request's result is the same from the batched server (shared scan, signature
groups, shared build tables, batched join probes) and from the one-launch-per-query
server, and a sample of scan and join requests equals the oracle."""
import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import _capi as A
from turbogp_amd import workloads as W
from turbogp_amd.joinload import make_join_tables
from turbogp_amd.serve import Server

pytestmark = pytest.mark.gpu


def same(a, b):
    return (np.array_equal(a.count, b.count) and np.array_equal(a.sums_words, b.sums_words)
            and a.n_out_of_range == b.n_out_of_range)


def host_slice(names, b, e):
    """Rows [b, e) of fact columns, from the generator the device table uses."""
    out = []
    for n in names:
        cid, dt, lo, hi = W.FACT_SCHEMA[n]
        out.append(B.synth_uniform(dt, W.SEED, cid, lo, hi, b, e))
    return out


def test_full_size_mix_batched_equals_per_query_and_oracle(ctx):
    table = W.make_fact_table(ctx, W.BATCH_TABLE_ROWS)
    build, probe = make_join_tables(ctx)
    try:
        mix = W.serve_mix(400, W.BATCH_TABLE_ROWS, W.JOIN_BUILD_ROWS, W.JOIN_PROBE_ROWS)
        reqs = W.serve_requests(table, build, probe, mix)
        got = {}
        for per_query in (False, True):
            srv = Server(ctx, max_batch_queries=1024, batch_slots=3, join_slots=4,
                         heavy_rows=50_000_000, per_query_launch=per_query)
            try:
                ids, _ = srv.replay(reqs, np.zeros(len(reqs), np.int64))
                res = [srv.result(i) for i in ids]
            finally:
                srv.close()
            assert all(t["status"] == A.OK for _, t in res)
            got[per_query] = [r for r, _ in res]
        assert all(same(a, b) for a, b in zip(got[False], got[True]))
        print("mix", {k: sum(1 for m in mix if m[0] == k) for k in ("q1", "q6", "join", "full")},
              flush=True)
        # oracle: the first scan requests of each shape and the first join requests
        scans = [i for i, m in enumerate(mix) if m[0] in ("q1", "q6")][:8]
        for i in scans:
            kind, b, e, _ = mix[i]
            names, preds, parts, aggs = W.q1_spec() if kind == "q1" else W.q6_spec()
            cnt, sums, oor = B.group_agg(B.make_query(host_slice(names, b, e), preds, parts,
                                                      aggs, 0, e - b))
            r = got[False][i]
            assert [int(x) for x in r.count] == [int(x) for x in cnt], i
            assert r.sums() == sums and r.n_out_of_range == oor, i
        bkey, _ = W.join_build_side(W.JOIN_BUILD_ROWS)
        pay = B.synth_uniform(A.I32, W.SEED, W.COL_PAY, 0, W.JOIN_PAY_CARD - 1, 0,
                              W.JOIN_BUILD_ROWS)
        cdf = W.zipf_cdf(W.JOIN_BUILD_ROWS)
        ktab = W.zipf_key_table(bkey)
        joins = [i for i, m in enumerate(mix) if m[0] == "join"][:8]
        for i in joins:
            _, b, e, nb = mix[i]
            fk = B.synth_cdf(A.I64, W.SEED, W.COL_FK, cdf, ktab, b, e)
            val = B.synth_uniform(A.I64, W.SEED, W.COL_VAL, 90_000, 10_500_000, b, e)
            cnt, sums, oor = B.join_agg(bkey[:nb], pay[:nb], fk, val, None,
                                        A.JOIN_GROUP_BUILD_PAYLOAD, 0, W.JOIN_PAY_CARD)
            r = got[False][i]
            assert np.array_equal(r.count, cnt), i
            assert [x[0] for x in r.sums()] == sums and r.n_out_of_range == oor, i
    finally:
        for t in (table, build, probe):
            for c in t.values():
                c.free()
