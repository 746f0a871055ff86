"""Continuous-batching server (tgp_serve_*): requests over device-resident columns."""
import ctypes as C

import numpy as np

from . import _capi as A
from .device import GroupAggResult


def agg_request(q):
    """A filter + GROUP BY request from an AggQuery (make_agg_query / build_query)."""
    r = A.ServeRequest()
    r.kind = A.SERVE_AGG
    r.agg = q
    return r


def join_request(build_keys, probe_keys, value, group_by, ngroups=1, key_base=0,
                 build_payload=None, group_col=None, build_rows=None, row_begin=0,
                 row_end=None):
    """A hash-join aggregate request: build over the first `build_rows` rows of
    build_keys (and build_payload), probe rows [row_begin, row_end), GROUP BY
    the build payload / a probe column / nothing, SUM(value) and COUNT(*)."""
    r = A.ServeRequest()
    r.kind = A.SERVE_JOIN_AGG
    bk = build_keys.desc()
    if build_rows is not None:
        bk.len = build_rows
    r.join.build_keys = bk
    if build_payload is not None:
        bp = build_payload.desc()
        if build_rows is not None:
            bp.len = build_rows
        r.join.build_payload = bp
    q = r.join.probe
    q.probe_keys = probe_keys.desc()
    q.value = value.desc()
    if group_col is not None:
        q.group_col = group_col.desc()
    q.group_by = group_by
    q.key_base = key_base
    q.ngroups = ngroups
    q.row_begin = row_begin
    q.row_end = probe_keys.n if row_end is None else row_end
    return r


def result_shape(req):
    """(ngroups, naggs) of a request's result."""
    if req.kind == A.SERVE_AGG:
        g = 1
        for k in range(req.agg.nparts):
            g *= req.agg.parts[k].card
        return g, req.agg.naggs
    q = req.join.probe
    return (1 if q.group_by == A.JOIN_GROUP_NONE else q.ngroups), 1


def parse_words(words, ngroups, naggs):
    cnt = words[:ngroups].copy()
    sums = words[ngroups: ngroups + ngroups * naggs * 2].reshape(ngroups, naggs, 2).copy()
    return GroupAggResult(cnt, sums, int(words[ngroups + ngroups * naggs * 2]))


class Server:
    def __init__(self, ctx, max_batch_queries=1024, batch_slots=2, join_slots=2, heavy_rows=0,
                 hash_budget_bytes=0, wg_per_cu=0, chunk_slices=64, order=3, cached_loads=True,
                 per_query_launch=False, batch_wait_us=250, share_build_tables=True):
        self.ctx = ctx
        cfg = A.ServeConfig(max_batch_queries, batch_slots, join_slots, batch_wait_us, heavy_rows,
                            hash_budget_bytes, wg_per_cu, chunk_slices, order, int(cached_loads),
                            int(per_query_launch), int(share_build_tables))
        h = C.c_void_p()
        A.check(A.lib().tgp_serve_create(ctx.handle, C.byref(cfg), C.byref(h)), "tgp_serve_create")
        self.handle = h.value
        self.shapes = {}

    @staticmethod
    def now_ns():
        return A.lib().tgp_serve_now_ns()

    def submit(self, req, arrival_ns=0):
        i = C.c_uint64()
        A.check(A.lib().tgp_serve_submit(self.handle, C.byref(req), arrival_ns, C.byref(i)),
                "tgp_serve_submit")
        self.shapes[i.value] = result_shape(req)
        return i.value

    def wait(self, rid, timeout_ns=-1):
        d = C.c_int32()
        A.check(A.lib().tgp_serve_wait(self.handle, rid, timeout_ns, C.byref(d)), "tgp_serve_wait")
        return bool(d.value)

    def result(self, rid, release=True):
        """-> (GroupAggResult or None on failure, timing dict)"""
        t = A.ServeTiming()
        rc = A.lib().tgp_serve_result(self.handle, rid, None, 0, C.byref(t))
        timing = {"arrival_ns": t.arrival_ns, "start_ns": t.start_ns, "done_ns": t.done_ns,
                  "status": t.status, "deferred": bool(t.deferred)}
        res = None
        if rc == A.OK:
            words = np.zeros(t.result_words, np.uint64)
            A.check(A.lib().tgp_serve_result(self.handle, rid, words.ctypes.data_as(
                C.POINTER(C.c_uint64)), t.result_words, None), "tgp_serve_result")
            ng, na = self.shapes[rid]
            res = parse_words(words, ng, na)
        if release:
            A.check(A.lib().tgp_serve_release(self.handle, rid), "tgp_serve_release")
            self.shapes.pop(rid, None)
        return res, timing

    def replay(self, reqs, offsets_ns):
        """Open-loop replay: request i arrives at start + offsets_ns[i].
        -> (ids, start_ns)"""
        n = len(reqs)
        arr = (A.ServeRequest * max(1, n))(*reqs)
        off = np.ascontiguousarray(offsets_ns, dtype=np.int64)
        ids = np.zeros(max(1, n), np.uint64)
        start = C.c_int64()
        A.check(A.lib().tgp_serve_replay(self.handle, arr, off.ctypes.data_as(
            C.POINTER(C.c_int64)), n, ids.ctypes.data_as(C.POINTER(C.c_uint64)),
            C.byref(start)), "tgp_serve_replay")
        ids = [int(x) for x in ids[:n]]
        for rid, req in zip(ids, reqs):
            self.shapes[rid] = result_shape(req)
        return ids, start.value

    def stats(self):
        st = A.ServeStats()
        A.check(A.lib().tgp_serve_get_stats(self.handle, C.byref(st)), "tgp_serve_get_stats")
        return {name: getattr(st, name) for name, _ in A.ServeStats._fields_}

    def close(self):
        if self.handle:
            A.check(A.lib().tgp_serve_destroy(self.handle), "tgp_serve_destroy")
            self.handle = None
