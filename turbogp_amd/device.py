"""Device context, HBM-tier columns and the operator calls of the scan /
group-aggregate path, over the C ABI (include/turbogp_gfx950.h).

This is the host-side mirror of the reference's operator surface:
  - `DeviceColumn` / `DeviceTable`  <- datasource/table.rs Table (columns at
    native width, resident in the HBM tier, memory/tier.rs MemoryTier::Hbm)
  - `Context.scan_filter`           <- exec/bitmap.rs filter_* + and_inplace
  - `DeviceBitmap`                  <- exec/bitmap.rs Bitmap / SetBitIter
  - `Context.group_agg`             <- query_interpreter/aggregate.rs execute_grouped
"""
import ctypes as C

import numpy as np

from . import _capi as A

_NP_DTYPE = {A.I8: np.int8, A.U8: np.uint8, A.I16: np.int16, A.U16: np.uint16,
             A.I32: np.int32, A.U32: np.uint32, A.I64: np.int64, A.U64: np.uint64}
_FROM_NP = {np.dtype(v): k for k, v in _NP_DTYPE.items()}


def dtype_of(arr):
    return _FROM_NP[np.asarray(arr).dtype]


def numeric_text(lo, hi, scale=0):
    """NUMERIC text of the int128 (lo, hi) scaled by 10^-scale (tgp_int128_to_numeric)."""
    buf = C.create_string_buffer(64)
    ln = C.c_uint64()
    A.check(A.lib().tgp_int128_to_numeric(int(lo), int(hi), scale, buf, 64, C.byref(ln)),
            "tgp_int128_to_numeric")
    return buf.value.decode()


def int128(lo, hi):
    """Python int from the (low, high) 64-bit words of a two's complement int128."""
    v = (int(hi) << 64) | int(lo)
    return v - (1 << 128) if v >> 127 else v


class DeviceBuffer:
    """A tier-accounted HBM allocation."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        A.check(A.lib().tgp_hbm_alloc(ctx.handle, self.nbytes, C.byref(p)), "tgp_hbm_alloc")
        self.ptr = p.value

    def free(self):
        if self.ptr is not None and self.ctx.handle:
            A.check(A.lib().tgp_hbm_free(self.ctx.handle, self.ptr), "tgp_hbm_free")
        self.ptr = None

    def upload(self, arr, stream=None):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        A.check(A.lib().tgp_memcpy_h2d(self.ctx.handle, self.ptr, arr.ctypes.data, arr.nbytes,
                                       stream), "tgp_memcpy_h2d")

    def download(self, dtype, count, stream=None, offset_bytes=0):
        out = np.empty(int(count), dtype=dtype)
        if out.nbytes:
            A.check(A.lib().tgp_memcpy_d2h(self.ctx.handle, out.ctypes.data,
                                           self.ptr + offset_bytes, out.nbytes, stream),
                    "tgp_memcpy_d2h")
        return out

    def fill_u64(self, value=0, nwords=None, stream=None):
        n = self.nbytes // 8 if nwords is None else nwords
        A.check(A.lib().tgp_fill_u64(self.ctx.handle, self.ptr, n, value & (2**64 - 1), stream),
                "tgp_fill_u64")


class DeviceColumn:
    """A column resident in the HBM tier at its native width."""

    def __init__(self, ctx, n, dtype, name=None):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = dtype
        self.width = A.DTYPE_WIDTH[dtype]
        self.name = name
        self.buf = DeviceBuffer(ctx, max(16, self.n * self.width))

    @classmethod
    def from_numpy(cls, ctx, arr, name=None):
        arr = np.ascontiguousarray(arr)
        c = cls(ctx, arr.shape[0], dtype_of(arr), name)
        if arr.nbytes:
            c.buf.upload(arr)
        return c

    @classmethod
    def from_cells(cls, ctx, cells, dtype, name=None):
        """A column of `dtype` from the reference's u64 cells (tgp_column_from_cells);
        a cell that does not fit the type raises."""
        cells = np.ascontiguousarray(cells, np.uint64)
        c = cls(ctx, cells.shape[0], dtype, name)
        bad = C.c_uint64()
        try:
            A.check(A.lib().tgp_column_from_cells(ctx.handle, cells.ctypes.data, c.n, dtype,
                                                  c.buf.ptr, C.byref(bad), None),
                    "tgp_column_from_cells")
        except Exception:
            c.free()
            raise
        return c

    def desc(self):
        return A.Column(self.buf.ptr, self.n, self.dtype, 0)

    def to_numpy(self):
        return self.buf.download(_NP_DTYPE[self.dtype], self.n)

    def minmax(self):
        lo, hi = C.c_int64(), C.c_int64()
        d = self.desc()
        A.check(A.lib().tgp_column_minmax_sync(self.ctx.handle, C.byref(d), C.byref(lo),
                                               C.byref(hi)), "tgp_column_minmax_sync")
        return lo.value, hi.value

    def free(self):
        self.buf.free()


class DeviceBitmap:
    """Selection bitmap on the device (bit i <-> row i, LSB-first words)."""

    def __init__(self, ctx, nbits):
        self.ctx = ctx
        self.nbits = int(nbits)
        self.nwords = int(A.lib().tgp_bitmap_words(self.nbits))
        self.buf = DeviceBuffer(ctx, max(8, self.nwords * 8))

    def words(self):
        return self.buf.download(np.uint64, self.nwords)

    def to_bytes(self):
        """The reference's Vec<u8> layout: ceil(nbits/8) LSB-first bytes."""
        return self.words().view(np.uint8)[: (self.nbits + 7) // 8].copy()

    def count_ones_range(self, begin, end):
        cnt = DeviceBuffer(self.ctx, 8)
        try:
            cnt.fill_u64(0)
            A.check(A.lib().tgp_bitmap_count_range(self.ctx.handle, self.buf.ptr, self.nbits,
                                                   begin, end, cnt.ptr, None),
                    "tgp_bitmap_count_range")
            return int(cnt.download(np.uint64, 1)[0])
        finally:
            cnt.free()

    def count_ones(self):
        return self.count_ones_range(0, self.nbits)

    @classmethod
    def from_null_flags(cls, ctx, flags):
        """NullBitmap flags (one bool per row, True = NULL) as a device bitmap."""
        flags = np.ascontiguousarray(flags, np.uint8)
        bm = cls(ctx, len(flags))
        A.check(A.lib().tgp_null_bitmap_from_flags(ctx.handle, flags.ctypes.data, len(flags),
                                                   bm.buf.ptr, None), "tgp_null_bitmap_from_flags")
        return bm

    def combine(self, other, op):
        out = DeviceBitmap(self.ctx, min(self.nbits, other.nbits))
        A.check(A.lib().tgp_bitmap_combine(self.ctx.handle, op, self.buf.ptr, other.buf.ptr,
                                           out.buf.ptr, out.nbits, None), "tgp_bitmap_combine")
        return out

    def not_(self):
        out = DeviceBitmap(self.ctx, self.nbits)
        A.check(A.lib().tgp_bitmap_not(self.ctx.handle, self.buf.ptr, out.buf.ptr, self.nbits,
                                       None), "tgp_bitmap_not")
        return out

    def iter_set_bits(self, base=0, index_bytes=8):
        """Ascending indices of the set bits (SetBitIter) computed on the device."""
        cap = max(1, self.count_ones())
        out = DeviceBuffer(self.ctx, cap * index_bytes)
        try:
            n = C.c_uint64()
            A.check(A.lib().tgp_bitmap_to_indices_sync(self.ctx.handle, self.buf.ptr, self.nbits,
                                                       base, index_bytes, out.ptr, cap,
                                                       C.byref(n), None),
                    "tgp_bitmap_to_indices_sync")
            return out.download(np.uint64 if index_bytes == 8 else np.uint32, n.value)
        finally:
            out.free()

    def free(self):
        self.buf.free()


class Context:
    """tgp_ctx: one gfx950 device, its stream and its HBM tier."""

    def __init__(self, device=0, hbm_budget_bytes=0):
        h = C.c_void_p()
        A.check(A.lib().tgp_ctx_create(device, hbm_budget_bytes, C.byref(h)), "tgp_ctx_create")
        self.handle = h.value
        info = A.DeviceInfo()
        A.check(A.lib().tgp_device_info_get(self.handle, C.byref(info)), "tgp_device_info_get")
        self.info = info

    def close(self):
        if self.handle:
            A.lib().tgp_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self, stream=None):
        A.check(A.lib().tgp_stream_sync(self.handle, stream), "tgp_stream_sync")

    def tune(self, key, value):
        A.check(A.lib().tgp_tune_set(self.handle, key, value), "tgp_tune_set")

    def hbm_usage(self):
        u, b, p = C.c_uint64(), C.c_uint64(), C.c_uint64()
        A.check(A.lib().tgp_hbm_usage(self.handle, C.byref(u), C.byref(b), C.byref(p)),
                "tgp_hbm_usage")
        return u.value, b.value, p.value

    # ---- scan -------------------------------------------------------------------------
    def scan_filter(self, cols, preds, row_begin=0, row_end=None, bitmap=True, count=True,
                    stream=None):
        """AND of predicate leaves (col_index, op, lit, is_signed) over cols.
        Returns (DeviceBitmap or None, count or None)."""
        n = cols[0].n if cols else 0
        row_end = n if row_end is None else row_end
        carr = (A.Column * max(1, len(cols)))(*[c.desc() for c in cols])
        parr = (A.Pred * max(1, len(preds)))(
            *[A.Pred(ci, op, int(bool(sg)), 0, lit) for (ci, op, lit, sg) in preds])
        bm = DeviceBitmap(self, row_end - row_begin) if bitmap else None
        cnt = DeviceBuffer(self, 8) if count else None
        try:
            if cnt:
                cnt.fill_u64(0, stream=stream)
            A.check(A.lib().tgp_scan_filter(self.handle, carr, len(cols), parr, len(preds),
                                            row_begin, row_end, bm.buf.ptr if bm else None,
                                            cnt.ptr if cnt else None, stream),
                    "tgp_scan_filter")
            c = int(cnt.download(np.uint64, 1, stream)[0]) if cnt else None
        finally:
            if cnt:
                cnt.free()
        return bm, c

    def scan_filter_f64(self, col, op, lit, eq_epsilon=False, row_begin=0, row_end=None,
                        stream=None):
        """FLOAT column (f64 bits in 8-byte cells) op literal (tgp_scan_filter_f64) ->
        (DeviceBitmap over [row_begin, row_end), count of set bits)."""
        row_end = col.n if row_end is None else row_end
        bm = DeviceBitmap(self, row_end - row_begin)
        cnt = DeviceBuffer(self, 8)
        try:
            cnt.fill_u64(0, stream=stream)
            d = col.desc()
            A.check(A.lib().tgp_scan_filter_f64(self.handle, C.byref(d), op, int(bool(eq_epsilon)),
                                                float(lit), row_begin, row_end, bm.buf.ptr,
                                                cnt.ptr, stream), "tgp_scan_filter_f64")
            return bm, int(cnt.download(np.uint64, 1, stream)[0])
        except Exception:
            bm.free()
            raise
        finally:
            cnt.free()

    def scan_expr(self, cols, prog, row_begin=0, row_end=None, bitmap=None, stream=None):
        """Rows of [row_begin, row_end) whose expression (a postfix program of (code, arg,
        imm), see turbogp_amd.expr) is truthy (tgp_scan_expr): narrows `bitmap` in place
        (bit i <-> row_begin + i) or, without one, returns a new DeviceBitmap."""
        row_end = (cols[0].n if cols else 0) if row_end is None else row_end
        own = bitmap is None
        bm = DeviceBitmap(self, row_end - row_begin) if own else bitmap
        try:
            if own:
                bm.buf.fill_u64(~0, nwords=bm.nwords, stream=stream)
            arr = (A.Column * max(1, len(cols)))(*[c.desc() for c in cols])
            ops = (A.ExprOp * len(prog))(*[A.ExprOp(c, a, im) for c, a, im in prog])
            A.check(A.lib().tgp_scan_expr(self.handle, arr, len(cols), ops, len(prog), row_begin,
                                          row_end, bm.buf.ptr, stream), "tgp_scan_expr")
            return bm
        except Exception:
            if own:
                bm.free()
            raise

    def join_residual(self, left_cols, right_cols, prog, probe_idx, build_row, left=False,
                      row_begin=0, row_end=None, stream=None):
        """Equi-join pairs (host arrays: probe rows, build rows / JOIN_NO_MATCH) filtered
        by a residual program over left_cols (read at the probe row) then right_cols (at
        the build row), LEFT keeping a probe row whose matches all fail
        (tgp_join_residual_sync).  Returns the kept (probe rows u64, build rows u32)."""
        probe_idx = np.ascontiguousarray(probe_idx, np.uint64)
        build_row = np.ascontiguousarray(build_row, np.uint32)
        n = len(probe_idx)
        if row_end is None:
            row_end = left_cols[0].n if left_cols else (int(probe_idx.max()) + 1 if n else 0)
        bufs = [DeviceBuffer(self, max(8, 8 * n)), DeviceBuffer(self, max(8, 4 * n)),
                DeviceBuffer(self, max(8, 8 * n)), DeviceBuffer(self, max(8, 4 * n))]
        try:
            if n:
                bufs[0].upload(probe_idx)
                bufs[1].upload(build_row)
            larr = (A.Column * max(1, len(left_cols)))(*[c.desc() for c in left_cols])
            rarr = (A.Column * max(1, len(right_cols)))(*[c.desc() for c in right_cols])
            ops = (A.ExprOp * len(prog))(*[A.ExprOp(c, a, im) for c, a, im in prog])
            m = C.c_uint64()
            A.check(A.lib().tgp_join_residual_sync(
                self.handle, larr, len(left_cols), rarr, len(right_cols), ops, len(prog),
                A.JOIN_LEFT if left else A.JOIN_INNER,
                row_begin, row_end, bufs[0].ptr, bufs[1].ptr, n, bufs[2].ptr, bufs[3].ptr,
                C.byref(m), stream), "tgp_join_residual_sync")
            k = m.value
            if k == 0:
                return np.zeros(0, np.uint64), np.zeros(0, np.uint32)
            return bufs[2].download(np.uint64, k), bufs[3].download(np.uint32, k)
        finally:
            for b in bufs:
                b.free()

    def compare_columns(self, a, op, b, is_signed, row_begin=0, row_end=None, bitmap=None,
                        stream=None):
        """Rows of [row_begin, row_end) where a op b (tgp_scan_compare_columns): narrows
        `bitmap` in place (bit i <-> row_begin + i) or, without one, returns a new
        DeviceBitmap of the comparison alone."""
        row_end = min(a.n, b.n) if row_end is None else row_end
        own = bitmap is None
        bm = DeviceBitmap(self, row_end - row_begin) if own else bitmap
        try:
            if own:
                bm.buf.fill_u64(~0, nwords=bm.nwords, stream=stream)
            da, db = a.desc(), b.desc()
            A.check(A.lib().tgp_scan_compare_columns(self.handle, C.byref(da), op, C.byref(db),
                                                     int(bool(is_signed)), row_begin, row_end,
                                                     bm.buf.ptr, stream),
                    "tgp_scan_compare_columns")
        except Exception:
            if own:
                bm.free()
            raise
        return bm

    # ---- COUNT(DISTINCT col) ------------------------------------------------------------
    def count_distinct(self, col, row_begin=0, row_end=None, sel_bitmap=None,
                       mode=A.DISTINCT_AUTO, stream=None):
        """COUNT(DISTINCT col) of rows [row_begin, row_end) selected by the optional
        DeviceBitmap (bit i <-> row_begin + i); mode AUTO / EXACT / HLL."""
        row_end = col.n if row_end is None else row_end
        d = col.desc()
        out = C.c_uint64()
        A.check(A.lib().tgp_count_distinct(self.handle, C.byref(d), row_begin, row_end,
                                           sel_bitmap.buf.ptr if sel_bitmap is not None else None,
                                           mode, C.byref(out), stream), "tgp_count_distinct")
        return out.value

    def hll_registers(self, col, row_begin=0, row_end=None, sel_bitmap=None, stream=None):
        """The 16384 HyperLogLog registers (uint8) of the same selection."""
        row_end = col.n if row_end is None else row_end
        d = col.desc()
        regs = np.zeros(1 << 14, np.uint8)
        A.check(A.lib().tgp_hll_sketch(self.handle, C.byref(d), row_begin, row_end,
                                       sel_bitmap.buf.ptr if sel_bitmap is not None else None,
                                       regs.ctypes.data, stream), "tgp_hll_sketch")
        return regs

    # ---- SELECT rows ---------------------------------------------------------------------
    def select_rows(self, row_begin, row_end, sel_bitmap=None, order_col=None, descending=False,
                    limit=None, stream=None):
        """Row indices of SELECT ... [ORDER BY order_col [DESC]] [LIMIT limit] over the rows
        [row_begin, row_end) selected by the optional DeviceBitmap -> (DeviceBuffer of u64
        rows, count)."""
        n = row_end - row_begin
        limit = n if limit is None else min(limit, n)
        out = DeviceBuffer(self, max(8, limit * 8))
        cnt = C.c_uint64()
        od = order_col.desc() if order_col is not None else None
        try:
            A.check(A.lib().tgp_select_rows(self.handle, row_begin, row_end,
                                            sel_bitmap.buf.ptr if sel_bitmap is not None else None,
                                            C.byref(od) if od is not None else None,
                                            int(bool(descending)), limit, out.ptr, limit,
                                            C.byref(cnt), stream), "tgp_select_rows")
        except Exception:
            out.free()
            raise
        return out, cnt.value

    def gather_column(self, col, rows, n, stream=None):
        """A new DeviceColumn of col's dtype holding col at the n row indices in the
        DeviceBuffer `rows` (GATHER_NO_ROW gives 0), kept on the device."""
        out = DeviceColumn(self, n, col.dtype)
        try:
            d = col.desc()
            A.check(A.lib().tgp_gather(self.handle, C.byref(d), rows.ptr, n, 0, out.buf.ptr,
                                       stream), "tgp_gather")
        except Exception:
            out.free()
            raise
        return out

    def gather(self, col, rows, n, as_cells=True, stream=None):
        """col at the n row indices in the DeviceBuffer `rows` -> numpy (u64 cells, or the
        column's dtype)."""
        dtype = np.uint64 if as_cells else _NP_DTYPE[col.dtype]
        out = DeviceBuffer(self, max(8, n * np.dtype(dtype).itemsize))
        try:
            d = col.desc()
            A.check(A.lib().tgp_gather(self.handle, C.byref(d), rows.ptr, n, int(as_cells),
                                       out.ptr, stream), "tgp_gather")
            return out.download(dtype, n)
        finally:
            out.free()

    # ---- aggregation --------------------------------------------------------------------
    def group_agg(self, q, stream=None, result=None):
        """Runs an AggQuery; returns GroupAggResult (downloaded, exact Python ints)."""
        ng = C.c_uint64()
        A.check(A.lib().tgp_group_agg_ngroups(C.byref(q), C.byref(ng)), "tgp_group_agg_ngroups")
        res = result or GroupAggBuffers(self, ng.value, q.naggs)
        A.check(A.lib().tgp_group_agg(self.handle, C.byref(q), C.byref(res.desc), stream),
                "tgp_group_agg")
        out = res.fetch(stream)
        if result is None:
            res.free()
        return out

    def sum_f64(self, cells, n=None, sel_bitmap=None, neg_zero_start=True, stream=None):
        """Exact, once-rounded SUM of f64 cells (tgp_sum_f64_sync); `cells` is a
        DeviceBuffer / DeviceColumn holding the doubles' bits."""
        buf = cells.buf if hasattr(cells, "buf") else cells
        n = (cells.n if hasattr(cells, "n") else buf.nbytes // 8) if n is None else n
        out = C.c_double()
        A.check(A.lib().tgp_sum_f64_sync(self.handle, buf.ptr, n,
                                         sel_bitmap.buf.ptr if sel_bitmap is not None else None,
                                         int(bool(neg_zero_start)), C.byref(out), stream),
                "tgp_sum_f64_sync")
        return out.value

    def hash_group_agg(self, q, key_cols, key_nulls=None, capacity=None, stream=None):
        """GROUP BY the key tuple of columns q.cols[key_cols] (any values; NULL keys
        from the optional per-key NULL bitmaps) through tgp_hash_group_agg_sync.
        Returns HashGroups (groups in no particular order).  `capacity` is the first
        guess of the group count; a larger table is taken as the call reports."""
        n = q.row_end - q.row_begin
        nk = len(key_cols)
        hq = A.HashGroupQuery()
        hq.agg = q
        hq.nkeys = nk
        for k, c in enumerate(key_cols):
            hq.key_col[k] = c
            nb = (key_nulls or [None] * nk)[k]
            hq.d_key_nulls[k] = nb.buf.ptr if nb is not None else None
        cap = max(1, min(n, 1 << 16) if capacity is None else capacity)
        while True:
            keys = DeviceBuffer(self, max(8, cap * nk * 8))
            nulls = DeviceBuffer(self, max(8, cap * 4))
            res = GroupAggBuffers(self, cap, q.naggs)
            try:
                out = A.HashGroupResult(keys.ptr, nulls.ptr, res.count.ptr, res.sums.ptr, cap)
                ng = C.c_uint64()
                rc = A.lib().tgp_hash_group_agg_sync(self.handle, C.byref(hq), C.byref(out),
                                                     C.byref(ng), stream)
                if rc == A.ERR_CAPACITY and ng.value > cap:
                    cap = max(int(ng.value), min(max(n, 1), 4 * cap))
                    continue
                A.check(rc, "tgp_hash_group_agg_sync")
                g = int(ng.value)
                kv = keys.download(np.uint64, g * nk).reshape(g, nk)
                nf = nulls.download(np.uint32, g)
                cnt = res.count.download(np.uint64, g)
                sums = res.sums.download(np.uint64, g * q.naggs * 2).reshape(g, q.naggs, 2)
                return HashGroups(kv, nf, GroupAggResult(cnt, sums, 0))
            finally:
                keys.free()
                nulls.free()
                res.free()

    def probe_agg_batch(self, queries, stream=None):
        """n probe aggregations (tgp_join_agg_query, e.g. HashTable.agg_query) in one
        tgp_hash_probe_agg_batch call; returns their GroupAggResults in order."""
        n = len(queries)
        bufs = [GroupAggBuffers(self, q.ngroups, 1) for q in queries]
        try:
            qa = (A.JoinAggQuery * max(1, n))(*queries)
            ra = (A.GroupResult * max(1, n))(*[b.desc for b in bufs])
            A.check(A.lib().tgp_hash_probe_agg_batch(self.handle, qa, n, ra, stream),
                    "tgp_hash_probe_agg_batch")
            return [b.fetch(stream) for b in bufs]
        finally:
            for b in bufs:
                b.free()

    def group_agg_async(self, q, res, stream=None):
        """Queues tgp_group_agg into preallocated GroupAggBuffers (no sync, no copy)."""
        A.check(A.lib().tgp_group_agg(self.handle, C.byref(q), C.byref(res.desc), stream),
                "tgp_group_agg")


class Event:
    """hipEvent on the context (timing of work queued on a stream)."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = C.c_void_p()
        A.check(A.lib().tgp_event_create(ctx.handle, C.byref(h)), "tgp_event_create")
        self.h = h.value

    def record(self, stream=None):
        A.check(A.lib().tgp_event_record(self.ctx.handle, self.h, stream), "tgp_event_record")

    def elapsed_ms(self, end):
        ms = C.c_float()
        A.check(A.lib().tgp_event_elapsed_ms(self.ctx.handle, self.h, end.h, C.byref(ms)),
                "tgp_event_elapsed_ms")
        return ms.value

    def free(self):
        if self.h:
            A.lib().tgp_event_destroy(self.ctx.handle, self.h)
            self.h = None


class GroupAggBuffers:
    """Device result buffers of tgp_group_agg."""

    def __init__(self, ctx, ngroups, naggs):
        self.ctx = ctx
        self.ngroups, self.naggs = int(ngroups), int(naggs)
        self.count = DeviceBuffer(ctx, max(8, self.ngroups * 8))
        self.sums = DeviceBuffer(ctx, max(16, self.ngroups * self.naggs * 16))
        self.oor = DeviceBuffer(ctx, 8)
        self.desc = A.GroupResult(self.count.ptr, self.sums.ptr, self.oor.ptr)

    def fetch(self, stream=None):
        cnt = self.count.download(np.uint64, self.ngroups, stream)
        sums = self.sums.download(np.uint64, self.ngroups * self.naggs * 2, stream)
        oor = int(self.oor.download(np.uint64, 1, stream)[0])
        return GroupAggResult(cnt, sums.reshape(self.ngroups, max(1, self.naggs), 2)
                              if self.naggs else np.zeros((self.ngroups, 0, 2), np.uint64), oor)

    def free(self):
        for b in (self.count, self.sums, self.oor):
            b.free()


class GroupAggResult:
    """Dense per-group COUNT(*) and exact int128 sums (as Python ints)."""

    def __init__(self, count, sums_words, oor):
        self.count = count
        self.sums_words = sums_words
        self.n_out_of_range = oor

    def sums(self):
        """list over groups of [int per aggregate]."""
        return [[int128(self.sums_words[g, a, 0], self.sums_words[g, a, 1])
                 for a in range(self.sums_words.shape[1])] for g in range(len(self.count))]

    def avg(self, a, form="exact"):
        """AVG of aggregate `a` per group from its exact sum and COUNT(*) (None
        for a group without rows):
          'exact': fractions.Fraction(sum, count);
          'f64':   the sum as f64 over the count as f64, as eval_agg_expr's Avg
                   (query_interpreter/aggregate.rs:1416-1424) for exact integer sums;
          'u64':   (sum mod 2^64) // count, the accumulator path's unsigned
                   truncating AVG (query_interpreter/accum_grouped.rs:380-390)."""
        from fractions import Fraction
        out = []
        for g in range(len(self.count)):
            n = int(self.count[g])
            if n == 0:
                out.append(None)
                continue
            s = int128(self.sums_words[g, a, 0], self.sums_words[g, a, 1])
            if form == "exact":
                out.append(Fraction(s, n))
            elif form == "f64":
                out.append(float(s) / float(n))
            elif form == "u64":
                out.append((s % (1 << 64)) // n)
            else:
                raise ValueError(f"avg form {form!r}: 'exact', 'f64' or 'u64'")
        return out

    def numeric(self, a, scale=0):
        """Aggregate `a` per group as NUMERIC text (pgwire type 1700) of the exact int128
        sum scaled by 10^-scale (tgp_int128_to_numeric)."""
        return [numeric_text(self.sums_words[g, a, 0], self.sums_words[g, a, 1], scale)
                for g in range(len(self.count))]

    def avg_numeric(self, a, scale=0, frac_digits=None):
        """AVG of aggregate `a` per group as NUMERIC text, rounded half away from zero to
        frac_digits fraction digits (default scale + 4); None for a group without rows."""
        frac_digits = scale + 4 if frac_digits is None else frac_digits
        out = []
        for g in range(len(self.count)):
            n = int(self.count[g])
            if n == 0:
                out.append(None)
                continue
            buf = C.create_string_buffer(96)
            ln = C.c_uint64()
            A.check(A.lib().tgp_int128_avg_to_numeric(int(self.sums_words[g, a, 0]),
                                                      int(self.sums_words[g, a, 1]), n, scale,
                                                      frac_digits, buf, 96, C.byref(ln)),
                    "tgp_int128_avg_to_numeric")
            out.append(buf.value.decode())
        return out

    def rows(self):
        """Present groups as (gid, count, [sums]) sorted by gid (compare after sorting)."""
        s = self.sums()
        return [(g, int(self.count[g]), s[g]) for g in range(len(self.count)) if self.count[g]]


class HashGroups:
    """Groups of tgp_hash_group_agg_sync: keys[g] (u64 cells, one per key column),
    null_flags[g] (bit k = key k is NULL) and the aggregates as a GroupAggResult whose
    group g is the g-th key tuple."""

    def __init__(self, keys, null_flags, result):
        self.keys = keys
        self.null_flags = null_flags
        self.result = result

    def __len__(self):
        return len(self.null_flags)

    def key_tuples(self, signed=None):
        """Key tuples as Python ints (None for NULL parts); signed[k] reads key k's
        cell as int64."""
        nk = self.keys.shape[1] if self.keys.ndim == 2 else 0
        signed = signed or [True] * nk
        out = []
        for g in range(len(self)):
            t = []
            for k in range(nk):
                if (int(self.null_flags[g]) >> k) & 1:
                    t.append(None)
                    continue
                v = int(self.keys[g, k])
                t.append(v - (1 << 64) if signed[k] and v >= (1 << 63) else v)
            out.append(tuple(t))
        return out

    def as_dict(self, signed=None):
        """{key tuple: (count, [exact sums])}."""
        s = self.result.sums()
        return {t: (int(self.result.count[g]), s[g])
                for g, t in enumerate(self.key_tuples(signed))}


_MINMAX = ("min", "max", "min_u64", "max_u64")


def make_agg_query(cols, preds=(), parts=(), aggs=(), row_begin=0, row_end=None, sel_bitmap=None):
    """Builds an AggQuery.
    cols:  list of DeviceColumn (or objects with .desc())
    preds: (col, op, lit, is_signed)
    parts: (col, divisor, base, card)
    aggs:  ('sum', coef, [(col, fa, fb), ...]), ('count_nonzero', col), or
           ('min' | 'max' | 'min_u64' | 'max_u64', col) / (..., coef, factors):
           MIN/MAX in signed int64 or raw u64 order
    """
    q = A.AggQuery()
    n = cols[0].n if cols else 0
    q.row_begin = row_begin
    q.row_end = n if row_end is None else row_end
    assert len(cols) <= A.MAX_COLS and len(preds) <= A.MAX_PREDS
    for i, c in enumerate(cols):
        q.cols[i] = c.desc()
    q.ncols = len(cols)
    for i, (ci, op, lit, sg) in enumerate(preds):
        q.preds[i] = A.Pred(ci, op, int(bool(sg)), 0, lit)
    q.npreds = len(preds)
    for i, (ci, div, base, card) in enumerate(parts):
        q.parts[i] = A.GroupPart(ci, div, base, card)
    q.nparts = len(parts)
    for i, ag in enumerate(aggs):
        a = A.Agg()
        if ag[0] == "count_nonzero":
            a.kind, a.nfactors = A.AGG_COUNT_NONZERO, 1
            a.fcol[0] = ag[1]
            a.coef = 1
        elif ag[0] in _MINMAX:
            # ('min' | 'max' | 'min_u64' | 'max_u64', col) or (..., coef, [(col, fa, fb), ...])
            a.kind = A.AGG_MINMAX[ag[0]]
            if len(ag) == 2:
                a.nfactors, a.coef = 1, 1
                a.fcol[0], a.fa[0], a.fb[0] = ag[1], 0, 1
            else:
                a.nfactors, a.coef = len(ag[2]), ag[1]
                for k, (ci, fa, fb) in enumerate(ag[2]):
                    a.fcol[k], a.fa[k], a.fb[k] = ci, fa, fb
        else:
            _, coef, factors = ag
            a.kind, a.nfactors, a.coef = A.AGG_SUM, len(factors), coef
            for k, (ci, fa, fb) in enumerate(factors):
                a.fcol[k], a.fa[k], a.fb[k] = ci, fa, fb
        q.aggs[i] = a
    q.naggs = len(aggs)
    q.d_sel_bitmap = sel_bitmap.buf.ptr if sel_bitmap is not None else None
    return q


def _key_array(keys):
    """One DeviceColumn or a list of them -> (ctypes array of tgp_column, count)."""
    cols = list(keys) if isinstance(keys, (list, tuple)) else [keys]
    arr = (A.Column * len(cols))(*[c.desc() for c in cols])
    return arr, len(cols)


class HashTable:
    """Device hash table of a join's build side (tgp_hash_build).
    Mirrors JoinHashTable (exec/join_hash_table.rs) / HashBuildScalar (kernel/hash.rs).
    `keys` may be a list of 2..4 columns: a multi-column table (tgp_hash_build_multi,
    exec/join.rs pack_key), probed with join_pairs / join_mark; those key columns must
    stay allocated while the table is used."""

    def __init__(self, ctx, keys, payload=None, stream=None, memory=None):
        """memory: a DeviceBuffer to hold the slots (tgp_hash_build_in), reusable
        for another table once this one is freed."""
        self.ctx = ctx
        h = C.c_void_p()
        pd = payload.desc() if payload is not None else None
        pdp = C.byref(pd) if pd else None
        if isinstance(keys, (list, tuple)):
            if memory is not None:
                raise ValueError("multi-column tables are built in their own memory")
            arr, nk = _key_array(keys)
            self.nkeys = nk
            A.check(A.lib().tgp_hash_build_multi(ctx.handle, arr, nk, pdp, C.byref(h), stream),
                    "tgp_hash_build_multi")
            self.handle = h.value
            return
        self.nkeys = 1
        kd = keys.desc()
        if memory is None:
            A.check(A.lib().tgp_hash_build(ctx.handle, C.byref(kd), pdp, C.byref(h), stream),
                    "tgp_hash_build")
        else:
            A.check(A.lib().tgp_hash_build_in(ctx.handle, C.byref(kd), pdp, memory.ptr,
                                              memory.nbytes, C.byref(h), stream),
                    "tgp_hash_build_in")
        self.handle = h.value

    def footprint(self):
        b = C.c_uint64()
        A.check(A.lib().tgp_hash_table_footprint(self.handle, C.byref(b)),
                "tgp_hash_table_footprint")
        return b.value

    def info(self):
        n, cap, dup = C.c_uint64(), C.c_uint64(), C.c_int32()
        A.check(A.lib().tgp_hash_table_info(self.handle, C.byref(n), C.byref(cap), C.byref(dup)),
                "tgp_hash_table_info")
        return n.value, cap.value, bool(dup.value)

    def probe_count(self, keys, row_begin=0, row_end=None, stream=None, result=None,
                    fetch=True):
        """Number of matches.  `result`: a DeviceBuffer of one word to count into
        (zeroed first); with fetch=False nothing is copied back."""
        kd = keys.desc()
        row_end = keys.n if row_end is None else row_end
        cnt = result or DeviceBuffer(self.ctx, 8)
        try:
            cnt.fill_u64(0, nwords=1, stream=stream)
            A.check(A.lib().tgp_hash_probe_count(self.ctx.handle, self.handle, C.byref(kd),
                                                 row_begin, row_end, cnt.ptr, stream),
                    "tgp_hash_probe_count")
            if fetch:
                return int(cnt.download(np.uint64, 1, stream)[0])
            return None
        finally:
            if result is None:
                cnt.free()

    def probe_mark(self, keys, row_begin=0, row_end=None, stream=None):
        kd = keys.desc()
        row_end = keys.n if row_end is None else row_end
        bm = DeviceBitmap(self.ctx, row_end - row_begin)
        A.check(A.lib().tgp_hash_probe_mark(self.ctx.handle, self.handle, C.byref(kd), row_begin,
                                            row_end, bm.buf.ptr, stream), "tgp_hash_probe_mark")
        return bm

    def probe_pairs(self, keys, row_begin=0, row_end=None, capacity=None, stream=None):
        kd = keys.desc()
        row_end = keys.n if row_end is None else row_end
        npairs = C.c_uint64()
        if capacity is None:
            rc = A.lib().tgp_hash_probe_pairs_sync(self.ctx.handle, self.handle, C.byref(kd),
                                                   row_begin, row_end, None, None, 0,
                                                   C.byref(npairs), stream)
            if rc not in (A.OK, A.ERR_CAPACITY):
                A.check(rc, "tgp_hash_probe_pairs_sync")
            capacity = npairs.value
        pi = DeviceBuffer(self.ctx, max(8, capacity * 8))
        pp = DeviceBuffer(self.ctx, max(8, capacity * 4))
        try:
            A.check(A.lib().tgp_hash_probe_pairs_sync(self.ctx.handle, self.handle, C.byref(kd),
                                                      row_begin, row_end, pi.ptr, pp.ptr,
                                                      capacity, C.byref(npairs), stream),
                    "tgp_hash_probe_pairs_sync")
            n = npairs.value
            return pi.download(np.uint64, n), pp.download(np.uint32, n)
        finally:
            pi.free()
            pp.free()

    def join_pairs(self, keys, join_type=None, row_begin=0, row_end=None, stream=None):
        """(probe rows, build payloads) of an INNER or LEFT join over the probe key
        column(s) (tgp_hash_join_pairs_sync); a LEFT row without a match has payload
        JOIN_NO_MATCH."""
        arr, nk = _key_array(keys)
        join_type = A.JOIN_INNER if join_type is None else join_type
        row_end = arr[0].len if row_end is None else row_end
        npairs = C.c_uint64()
        rc = A.lib().tgp_hash_join_pairs_sync(self.ctx.handle, self.handle, arr, nk, join_type,
                                              row_begin, row_end, None, None, 0,
                                              C.byref(npairs), stream)
        if rc not in (A.OK, A.ERR_CAPACITY):
            A.check(rc, "tgp_hash_join_pairs_sync")
        capacity = npairs.value
        pi = DeviceBuffer(self.ctx, max(8, capacity * 8))
        pp = DeviceBuffer(self.ctx, max(8, capacity * 4))
        try:
            A.check(A.lib().tgp_hash_join_pairs_sync(self.ctx.handle, self.handle, arr, nk,
                                                     join_type, row_begin, row_end, pi.ptr,
                                                     pp.ptr, capacity, C.byref(npairs), stream),
                    "tgp_hash_join_pairs_sync")
            n = npairs.value
            return pi.download(np.uint64, n), pp.download(np.uint32, n)
        finally:
            pi.free()
            pp.free()

    def join_mark(self, keys, row_begin=0, row_end=None, stream=None):
        """Semi-join marks over the probe key column(s) (tgp_hash_join_mark)."""
        arr, nk = _key_array(keys)
        row_end = arr[0].len if row_end is None else row_end
        bm = DeviceBitmap(self.ctx, row_end - row_begin)
        try:
            A.check(A.lib().tgp_hash_join_mark(self.ctx.handle, self.handle, arr, nk, row_begin,
                                               row_end, bm.buf.ptr, stream), "tgp_hash_join_mark")
        except Exception:
            bm.free()
            raise
        return bm

    def add_bloom(self, bits_per_key=8, stream=None):
        """Blocked bloom filter in front of the table (tgp_hash_table_add_bloom)."""
        A.check(A.lib().tgp_hash_table_add_bloom(self.ctx.handle, self.handle, bits_per_key,
                                                 stream), "tgp_hash_table_add_bloom")

    def agg_query(self, keys, value, group_by, ngroups=1, key_base=0, group_col=None,
                  row_begin=0, row_end=None, build_rows=0):
        """The tgp_join_agg_query of a probe aggregation over this table."""
        q = A.JoinAggQuery()
        q.table = self.handle
        q.build_rows = build_rows
        q.probe_keys = keys.desc()
        q.value = value.desc()
        if group_col is not None:
            q.group_col = group_col.desc()
        q.group_by = group_by
        q.key_base = key_base
        q.ngroups = ngroups
        q.row_begin = row_begin
        q.row_end = keys.n if row_end is None else row_end
        return q

    def probe_agg(self, keys, value, group_by, ngroups=1, key_base=0, group_col=None,
                  row_begin=0, row_end=None, result=None, stream=None, fetch=True,
                  build_rows=0):
        """Fused probe -> GROUP BY -> SUM(value), COUNT(*); build_rows n > 0 joins
        with the first n build rows only (a table built over more of the build side)."""
        q = self.agg_query(keys, value, group_by, ngroups, key_base, group_col, row_begin,
                           row_end, build_rows)
        res = result or GroupAggBuffers(self.ctx, ngroups, 1)
        A.check(A.lib().tgp_hash_probe_agg(self.ctx.handle, C.byref(q), C.byref(res.desc),
                                           stream), "tgp_hash_probe_agg")
        if not fetch:
            return res
        out = res.fetch(stream)
        if result is None:
            res.free()
        return out

    def free(self):
        if self.handle and self.ctx.handle:
            A.check(A.lib().tgp_hash_table_destroy(self.ctx.handle, self.handle),
                    "tgp_hash_table_destroy")
        self.handle = None
