"""Batched executor (tgp_batch_*): many filter + GROUP BY queries per launch."""
import ctypes as C

import numpy as np

from . import _capi as A
from .device import GroupAggResult


class Batch:
    def __init__(self, ctx, max_queries):
        self.ctx = ctx
        h = C.c_void_p()
        A.check(A.lib().tgp_batch_create(ctx.handle, max_queries, C.byref(h)), "tgp_batch_create")
        self.handle = h.value
        self.n = 0

    def set_queries(self, queries, stream=None):
        arr = (A.AggQuery * max(1, len(queries)))(*queries)
        A.check(A.lib().tgp_batch_set_queries(self.handle, arr, len(queries), stream),
                "tgp_batch_set_queries")
        self.n = len(queries)
        self.naggs = [q.naggs for q in queries]

    def set_tuning(self, wg_per_cu=0, use_graph=True, chunk_slices=0, order=0,
                   cached_loads=False, items_per_grab=1):
        t = A.BatchTuning(wg_per_cu, chunk_slices, order, int(cached_loads), int(use_graph),
                          items_per_grab)
        A.check(A.lib().tgp_batch_set_tuning(self.handle, C.byref(t)), "tgp_batch_set_tuning")

    @staticmethod
    def _tuning_dict(t):
        return {"wg_per_cu": t.wg_per_cu, "chunk_slices": t.chunk_slices, "order": t.order,
                "cached_loads": bool(t.cached_loads), "use_graph": bool(t.use_graph),
                "items_per_grab": t.items_per_grab}

    def tuning(self):
        t = A.BatchTuning()
        A.check(A.lib().tgp_batch_get_tuning(self.handle, C.byref(t)), "tgp_batch_get_tuning")
        return self._tuning_dict(t)

    def autotune(self, stream=None):
        """Times the launch settings on the current queries and keeps the fastest.
        -> (settings dict, best ms)"""
        t, ms = A.BatchTuning(), C.c_float()
        A.check(A.lib().tgp_batch_autotune(self.handle, stream, C.byref(t), C.byref(ms)),
                "tgp_batch_autotune")
        return self._tuning_dict(t), ms.value

    def run(self, mode=A.BATCH_PERSISTENT, stream=None):
        A.check(A.lib().tgp_batch_run(self.handle, mode, stream), "tgp_batch_run")

    def run_timed(self, mode=A.BATCH_PERSISTENT, stream=None):
        q = (C.c_float * max(1, self.n))()
        tot = C.c_float()
        A.check(A.lib().tgp_batch_run_timed(self.handle, mode, stream, q, C.byref(tot)),
                "tgp_batch_run_timed")
        return tot.value, np.frombuffer(q, dtype=np.float32)[: self.n].copy()

    def info(self):
        w, s, nb, nf = C.c_uint64(), C.c_uint64(), C.c_uint32(), C.c_uint32()
        A.check(A.lib().tgp_batch_info(self.handle, C.byref(w), C.byref(s), C.byref(nb),
                                       C.byref(nf)), "tgp_batch_info")
        return {"result_words": w.value, "total_slices": s.value, "batched": nb.value,
                "fallback": nf.value}

    def results(self, stream=None):
        """List of GroupAggResult, one per query."""
        words = self.info()["result_words"]
        arena = np.zeros(max(1, words), np.uint64)
        A.check(A.lib().tgp_batch_results_sync(self.handle, arena.ctypes.data_as(
            C.POINTER(C.c_uint64)), words, stream), "tgp_batch_results_sync")
        out = []
        for i in range(self.n):
            co, so, oo, ng = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
            A.check(A.lib().tgp_batch_result_layout(self.handle, i, C.byref(co), C.byref(so),
                                                    C.byref(oo), C.byref(ng)),
                    "tgp_batch_result_layout")
            g, na = ng.value, self.naggs[i]
            cnt = arena[co.value: co.value + g]
            sums = arena[so.value: so.value + g * na * 2].reshape(g, na, 2) if na else \
                np.zeros((g, 0, 2), np.uint64)
            out.append(GroupAggResult(cnt.copy(), sums.copy(), int(arena[oo.value])))
        return out

    def close(self):
        if self.handle:
            A.lib().tgp_batch_destroy(self.handle)
            self.handle = None
