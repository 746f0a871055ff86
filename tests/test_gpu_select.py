"""GPU SELECT rows (compaction, stable ORDER BY on u64 cells, LIMIT) and gather
against the oracle and the reference's ORDER BY tests."""
import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import DeviceBitmap, DeviceColumn, _capi as A
from turbogp_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu
DTYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64]


def dev(ctx, arr):
    return DeviceColumn.from_numpy(ctx, np.ascontiguousarray(arr))


def run_select(ctx, b, e, words, order, desc, limit):
    bm = None
    if words is not None:
        bm = DeviceBitmap(ctx, e - b)
        bm.buf.upload(words)
    try:
        rows, cnt = ctx.select_rows(b, e, bm, order, desc, limit)
        try:
            got = rows.download(np.uint64, cnt) if cnt else np.zeros(0, np.uint64)
        finally:
            rows.free()
        return got
    finally:
        if bm:
            bm.free()


def test_reference_order_by_kats(ctx):  # tests/joins_groupby_orderby.rs:112-155
    for a_vals, b_vals, desc, limit, want_a in (([3, 1, 2], [30, 10, 20], False, None, [1, 2, 3]),
                                               ([1, 3, 2], [10, 30, 20], True, None, [3, 2, 1]),
                                               ([5, 1, 3, 2, 4], [50, 10, 30, 20, 40], False, 2,
                                                [1, 2])):
        a, b = dev(ctx, np.array(a_vals, np.int64)), dev(ctx, np.array(b_vals, np.int64))
        try:
            rows, cnt = ctx.select_rows(0, len(a_vals), None, a, desc, limit)
            try:
                assert list(ctx.gather(a, rows, cnt)) == want_a
                assert list(ctx.gather(b, rows, cnt)) == [10 * x for x in want_a]
            finally:
                rows.free()
        finally:
            a.free()
            b.free()


@pytest.mark.parametrize("dtype", DTYPES)
def test_select_rows_vs_oracle(ctx, dtype):
    rng = np.random.default_rng(np.dtype(dtype).itemsize * 13 + (np.iinfo(dtype).min < 0))
    n = 200_003
    info = np.iinfo(dtype)
    host = rng.integers(max(info.min, -50), min(info.max, 50), n, dtype=dtype, endpoint=True)
    host[::1000] = info.min
    host[1::1000] = info.max
    col = dev(ctx, host)
    try:
        for b, e, use_bm, limit in ((0, n, False, None), (17, 150_017, True, None),
                                    (5, 199_000, True, 1000), (3, 3, True, 10), (100, 101, False, 5)):
            words = None
            if use_bm and e > b:
                words = rng.integers(0, 2**63, (e - b + 63) // 64).astype(np.uint64)
            for order, desc in ((None, False), (col, False), (col, True)):
                got = run_select(ctx, b, e, words, order, desc, limit)
                want = B.select_rows(b, e, words, host if order is not None else None, desc, limit)
                assert np.array_equal(got, want), (b, e, use_bm, limit, order is not None, desc)
        # gather at those rows: cells and native width
        rows, cnt = ctx.select_rows(10, 180_010, None, col, True, 50_000)
        try:
            idx = rows.download(np.uint64, cnt).astype(np.int64)
            assert np.array_equal(ctx.gather(col, rows, cnt), B.cells(host)[idx])
            assert np.array_equal(ctx.gather(col, rows, cnt, as_cells=False), host[idx])
        finally:
            rows.free()
    finally:
        col.free()


def test_gather_rejects_rows_outside_the_column(ctx):
    col = dev(ctx, np.arange(100, dtype=np.int64))
    rows = DeviceBuffer(ctx, 3 * 8)
    try:
        rows.upload(np.array([1, 99, 100], np.uint64))  # 100 is past the end: checked in-kernel
        with pytest.raises(RuntimeError, match="outside the column"):
            ctx.gather(col, rows, 3)
        rows.upload(np.array([1, 99, 0], np.uint64))
        assert list(ctx.gather(col, rows, 3)) == [1, 99, 0]
        rows.upload(np.array([7, A.GATHER_NO_ROW, 5], np.uint64))  # no row: a 0 cell
        assert list(ctx.gather(col, rows, 3)) == [7, 0, 5]
        g = ctx.gather_column(col, rows, 3)
        try:
            assert list(g.to_numpy()) == [7, 0, 5]
        finally:
            g.free()
    finally:
        rows.free()
        col.free()


@pytest.mark.parametrize("kind", ["full_range", "all_equal", "few_values"])
def test_order_by_limit_top_k_vs_oracle(ctx, kind):
    """LIMIT up to 4096 takes the radix-select path (LIMIT 4097 the full sort):
    the first k rows of the stable (key, scan order) sort, ASC and DESC, with
    every key digit in play, with all keys equal (scan order decides) and with
    heavy ties."""
    rng = np.random.default_rng({"full_range": 1, "all_equal": 2, "few_values": 3}[kind])
    n = 1_500_007
    if kind == "full_range":
        host = rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64, endpoint=True)
        host[::99_991] = np.iinfo(np.int64).min
        host[1::99_991] = np.iinfo(np.int64).max
    elif kind == "all_equal":
        host = np.full(n, -7, np.int64)
    else:
        host = rng.integers(0, 5, n).astype(np.int64)
    col = dev(ctx, host)
    try:
        for b, e, use_bm in ((0, n, False), (11, n - 5, True)):
            words = None
            if use_bm:
                words = rng.integers(0, 2**63, (e - b + 63) // 64).astype(np.uint64)
            for limit in (1, 37, 4095, 4096, 4097):
                for desc in (False, True):
                    got = run_select(ctx, b, e, words, col, desc, limit)
                    want = B.select_rows(b, e, words, host, desc, limit)
                    assert np.array_equal(got, want), (kind, b, e, use_bm, limit, desc)
    finally:
        col.free()


def test_top_k_unsigned_and_narrow_columns(ctx):
    """u64 cells above 2^63 and narrow signed columns (sign-extended cells) in
    the top-k path."""
    rng = np.random.default_rng(4)
    n = 300_001
    for host in (rng.integers(0, 2**64 - 1, n, dtype=np.uint64, endpoint=True),
                 rng.integers(-128, 127, n, dtype=np.int8, endpoint=True),
                 rng.integers(0, 65535, n, dtype=np.uint16, endpoint=True)):
        col = dev(ctx, host)
        try:
            for desc in (False, True):
                for limit in (3, 2000):
                    got = run_select(ctx, 0, n, None, col, desc, limit)
                    want = B.select_rows(0, n, None, host, desc, limit)
                    assert np.array_equal(got, want), (host.dtype, desc, limit)
        finally:
            col.free()
