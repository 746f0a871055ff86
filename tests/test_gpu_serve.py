"""Continuous-batching server: every request's result is the same alone and in
the mix (bit for bit), equals the direct single-call result and the oracle, and
hash tables are admitted by the HBM budget."""
import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import HashTable, _capi as A
from turbogp_amd import workloads as W
from turbogp_amd.joinload import make_join_tables
from turbogp_amd.serve import Server

pytestmark = pytest.mark.gpu
N_FACT, N_BUILD, N_PROBE = 3_000_000, 200_000, 3_000_000


@pytest.fixture(scope="module")
def tables(ctx):
    fact = W.make_fact_table(ctx, N_FACT)
    build, probe = make_join_tables(ctx, N_BUILD, N_PROBE, 1_000)
    yield fact, build, probe
    for t in (fact, build, probe):
        for c in t.values():
            c.free()


def same(a, b):
    return (a is not None and b is not None and np.array_equal(a.count, b.count)
            and np.array_equal(a.sums_words, b.sums_words)
            and a.n_out_of_range == b.n_out_of_range)


def direct(ctx, tables, desc):
    """The request run by the single-query entry points."""
    fact, build, probe = tables
    kind, b, e, nb = desc
    if kind in ("q1", "full"):
        return ctx.group_agg(W.build_query(fact, W.q1_spec(), b, e))
    if kind == "q6":
        return ctx.group_agg(W.build_query(fact, W.q6_spec(), b, e))
    keys = build["bkey"]
    pay = build["pay"]

    class Prefix:  # the first nb build rows
        def __init__(self, col):
            self.col, self.n = col, nb

        def desc(self):
            d = self.col.desc()
            d.len = nb
            return d

    t = HashTable(ctx, Prefix(keys), Prefix(pay))
    try:
        return t.probe_agg(probe["fk"], probe["val"], A.JOIN_GROUP_BUILD_PAYLOAD,
                           ngroups=W.JOIN_PAY_CARD, row_begin=b, row_end=e)
    finally:
        t.free()


def test_alone_equals_in_mix_equals_direct(ctx, tables):
    fact, build, probe = tables
    mix = W.serve_mix(160, N_FACT, N_BUILD, N_PROBE, lmax=1_000_000, build_max=N_BUILD)
    reqs = W.serve_requests(fact, build, probe, mix)
    # alone: one request in the server at a time
    srv = Server(ctx, max_batch_queries=256, batch_slots=2, join_slots=2)
    try:
        alone = []
        for r in reqs:
            rid = srv.submit(r)
            srv.wait(rid)
            res, t = srv.result(rid)
            assert t["status"] == A.OK
            alone.append(res)
        # in the mix: an open-loop burst at 50K requests/s, so batches form
        ids, _ = srv.replay(reqs, W.poisson_offsets(len(reqs), 50_000))
        mixed = [srv.result(i)[0] for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    assert st["batches"] < st["batched_requests"], st  # requests really were batched
    # the one-launch-per-query serving form, same burst
    srv = Server(ctx, batch_slots=2, per_query_launch=True)
    try:
        ids, _ = srv.replay(reqs, W.poisson_offsets(len(reqs), 50_000))
        single = [srv.result(i)[0] for i in ids]
    finally:
        srv.close()
    for i, desc in enumerate(mix):
        assert same(alone[i], mixed[i]), (i, desc)
        assert same(alone[i], single[i]), (i, desc)
        assert same(alone[i], direct(ctx, tables, desc)), (i, desc)


def test_scan_requests_match_the_oracle(ctx, tables):
    fact, build, probe = tables
    mix = [("q1", 0, N_FACT, 0), ("q6", 123, 2_000_001, 0), ("q1", 5, 5, 0), ("q6", 77, 78, 0)]
    reqs = W.serve_requests(fact, build, probe, mix)
    host = W.host_fact_columns(N_FACT, B.synth_uniform)
    srv = Server(ctx)
    try:
        ids, _ = srv.replay(reqs, [0] * len(reqs))
        got = [srv.result(i)[0] for i in ids]
    finally:
        srv.close()
    for (kind, b, e, _), r in zip(mix, got):
        names, preds, parts, aggs = W.q1_spec() if kind == "q1" else W.q6_spec()
        cnt, sums, oor = B.group_agg(B.make_query([host[c] for c in names], preds, parts, aggs,
                                                  b, e))
        assert [int(x) for x in r.count] == [int(x) for x in cnt]
        assert r.sums() == sums and r.n_out_of_range == oor


def test_joins_wait_for_hbm_budget_and_oversized_ones_are_rejected(ctx, tables):
    fact, build, probe = tables
    one = A.lib().tgp_hash_table_bytes_for(N_BUILD)
    mix = [("join", 0, 500_000, N_BUILD), ("join", 1_000, 900_000, N_BUILD),
           ("join", 7, 300_000, N_BUILD // 2), ("q1", 0, 1_000_000, 0)]
    reqs = W.serve_requests(fact, build, probe, mix)
    # room for one full-size table at a time: the second join waits for the first
    # (every join builds its own table here)
    srv = Server(ctx, join_slots=4, hash_budget_bytes=one + one // 4, share_build_tables=False)
    try:
        ids, _ = srv.replay(reqs, [0] * len(reqs))
        got = [srv.result(i) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    assert st["deferred"] >= 1 and st["rejected"] == 0, st
    for (res, t), desc in zip(got, mix):
        assert t["status"] == A.OK
        assert same(res, direct(ctx, tables, desc)), desc
    # a join whose table can never fit is rejected, not queued forever (either mode)
    for share in (False, True):
        srv = Server(ctx, hash_budget_bytes=one // 2, share_build_tables=share)
        try:
            rid = srv.submit(reqs[0])
            srv.wait(rid)
            res, t = srv.result(rid)
            st = srv.stats()
        finally:
            srv.close()
        assert res is None and t["status"] == A.ERR_BUDGET and st["rejected"] == 1, share


def test_joins_share_one_table_per_build_side(ctx, tables):
    """Joins over prefixes of one build side probe one resident table with their prefix
    bound (built again only when a longer prefix is asked for); joins over another
    build side (a copy of the key and payload columns) get a table of their own.  Every
    result equals the direct build-and-probe of its own prefix, in both serving forms
    and in a mix with scans."""
    from turbogp_amd import DeviceColumn
    fact, build, probe = tables
    rng = np.random.default_rng(77)
    nbs = [int(x) for x in rng.integers(1, N_BUILD + 1, 60)] + [N_BUILD, 5, N_BUILD // 3]
    mix = [("join", int(b), int(b) + 200_000, nb)
           for b, nb in zip(rng.integers(0, N_PROBE - 200_000, len(nbs)), nbs)]
    mix += [("q6", 100, 900_000, 0), ("q1", 0, 2_000_000, 0)]
    reqs = W.serve_requests(fact, build, probe, mix)
    copy = {k: DeviceColumn.from_numpy(ctx, build[k].to_numpy()) for k in ("bkey", "pay")}
    try:
        mix2 = [("join", 0, 1_000_000, 150_000), ("join", 3, 2_000_000, N_BUILD)]
        reqs += W.serve_requests(fact, copy, probe, mix2)
        got = {}
        for share in (True, False):
            for per_query in (False, True):
                srv = Server(ctx, join_slots=4, batch_slots=2, share_build_tables=share,
                             per_query_launch=per_query)
                try:
                    ids, _ = srv.replay(reqs, W.poisson_offsets(len(reqs), 20_000))
                    got[(share, per_query)] = [srv.result(i)[0] for i in ids]
                    st = srv.stats()
                finally:
                    srv.close()
                njoins = len(nbs) + len(mix2)
                if share:
                    assert 2 <= st["table_builds"] < njoins // 3, st
                else:
                    assert st["table_builds"] == njoins, st
        for i, desc in enumerate(mix):
            want = direct(ctx, tables, desc)
            assert all(same(g[i], want) for g in got.values()), (i, desc)
        for j, desc in enumerate(mix2):
            want = direct(ctx, (fact, copy, probe), desc)
            assert all(same(g[len(mix) + j], want) for g in got.values()), desc
    finally:
        for c in copy.values():
            c.free()


def test_heavy_requests_batch_on_their_own(ctx, tables):
    fact, build, probe = tables
    mix = [("full", 0, N_FACT, 0)] * 3 + [("q6", 10, 20_000, 0)] * 20
    reqs = W.serve_requests(fact, build, probe, mix)
    srv = Server(ctx, heavy_rows=N_FACT // 2, batch_slots=3)
    try:
        ids, _ = srv.replay(reqs, [0] * len(reqs))
        got = [srv.result(i)[0] for i in ids]
    finally:
        srv.close()
    want_full = direct(ctx, tables, mix[0])
    want_q6 = direct(ctx, tables, mix[3])
    assert all(same(g, want_full) for g in got[:3])
    assert all(same(g, want_q6) for g in got[3:])
