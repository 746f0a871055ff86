"""GPU parity of the predicate scan, bitmap algebra and compaction against the
oracle (bit-exact), on the reference's known answers and seeded columns."""
import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import DeviceColumn, _capi as A

pytestmark = pytest.mark.gpu
EQ, NE, LT, LE, GT, GE = range(6)


def oracle_bitmap(host_cols, preds, b, e):
    q = B.make_query(host_cols, preds, row_begin=b, row_end=e)
    return B.query_filter(q)


def gpu_bitmap(ctx, host_cols, preds, b, e):
    cols = [DeviceColumn.from_numpy(ctx, c) for c in host_cols]
    try:
        bm, cnt = ctx.scan_filter(cols, preds, b, e)
        out = bm.to_bytes()
        bm.free()
        return out, cnt
    finally:
        for c in cols:
            c.free()


def test_filter_eq_u64_kat(ctx):  # bitmap.rs test_filter_eq_u64
    col = np.array([1, 5, 5, 3, 5, 7, 8, 5, 9, 10], np.uint64)
    bits, cnt = gpu_bitmap(ctx, [col], [(0, EQ, 5, False)], 0, 10)
    assert [bool((bits[i >> 3] >> (i & 7)) & 1) for i in range(10)] == \
        [False, True, True, False, True, False, False, True, False, False]
    assert cnt == 4


def test_filter_signed_vs_unsigned(ctx):  # bitmap.rs test_filter_lt_i64_signed
    col = np.array([-5, -3, -1, 0, 1, 3, 5], np.int64)
    b1, c1 = gpu_bitmap(ctx, [col], [(0, LT, 0, True)], 0, 7)
    b2, c2 = gpu_bitmap(ctx, [col], [(0, LT, 0, False)], 0, 7)
    assert c1 == 3 and b1[0] == 0b111 and c2 == 0


def test_large_filter_kats(ctx):  # bitmap.rs test_large_filter_u64 / and_into_bool
    col = np.arange(1_000_000, dtype=np.uint64) % 100
    _, c = gpu_bitmap(ctx, [col], [(0, EQ, 50, False)], 0, len(col))
    assert c == 10_000
    _, c = gpu_bitmap(ctx, [col], [(0, GE, 10, False), (0, LE, 50, False)], 0, len(col))
    assert c == 410_000


@pytest.mark.parametrize("dtype", [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32,
                                   np.int64, np.uint64])
@pytest.mark.parametrize("n,b,e", [(0, 0, 0), (1, 0, 1), (63, 0, 63), (64, 0, 64), (65, 0, 65),
                                   (1023, 0, 1023), (1025, 0, 1025), (5000, 3, 4999),
                                   (70001, 17, 70001), (100000, 1024, 99000)])
def test_scan_all_dtypes_ragged(ctx, dtype, n, b, e):
    rng = np.random.default_rng(n * 7 + np.dtype(dtype).itemsize)
    info = np.iinfo(dtype)
    col = rng.integers(info.min, info.max, n, dtype=dtype, endpoint=True) if n else \
        np.zeros(0, dtype)
    col2 = rng.integers(0, 100, n).astype(np.int32)
    lit = int(col[n // 2]) if n else 0
    for op in range(6):
        for sg in (True, False):
            preds = [(0, op, lit, sg), (1, LT, 60, True)]
            if n == 0:
                continue
            want = oracle_bitmap([col, col2], preds, b, e)
            got, cnt = gpu_bitmap(ctx, [col, col2], preds, b, e)
            assert np.array_equal(got, want), (op, sg)
            assert cnt == int(np.unpackbits(want, bitorder="little")[: e - b].sum())


def test_bitmap_algebra_and_compaction(ctx):
    rng = np.random.default_rng(5)
    n = 200_003
    a = rng.integers(0, 1000, n).astype(np.int64)
    ca = DeviceColumn.from_numpy(ctx, a)
    bm1, _ = ctx.scan_filter([ca], [(0, LT, 300, True)])
    bm2, _ = ctx.scan_filter([ca], [(0, GT, 900, True)])
    o1, o2 = B.filter_cells(a.view(np.uint64), LT, 1, 300), B.filter_cells(a.view(np.uint64), GT,
                                                                            1, 900)
    orr = bm1.combine(bm2, A.BIT_OR)
    assert np.array_equal(orr.to_bytes(), o1 | o2)
    nt = orr.not_()
    assert np.array_equal(nt.to_bytes(), B.bitmap_not(o1 | o2, n))
    andn = bm1.combine(bm2, A.BIT_AND)
    assert andn.count_ones() == 0
    for (s, e) in [(0, n), (5, 77), (64, 128), (1000, 150_001), (n - 3, n + 10), (9, 9)]:
        assert nt.count_ones_range(s, e) == B.count_range(B.bitmap_not(o1 | o2, n), n, s, e)
    idx = nt.iter_set_bits()
    assert np.array_equal(idx, B.set_bits(B.bitmap_not(o1 | o2, n), n))
    idx32 = nt.iter_set_bits(base=10, index_bytes=4)
    assert np.array_equal(idx32.astype(np.uint64), idx + 10)
    for x in (bm1, bm2, orr, nt, andn):
        x.free()
    ca.free()


def test_compaction_kats(ctx):  # bitmap.rs test_iter_set_bits_sparse / skips_zero_bytes
    col = np.zeros(100, np.int64)
    col[[5, 31, 63, 99]] = 1
    c = DeviceColumn.from_numpy(ctx, col)
    bm, cnt = ctx.scan_filter([c], [(0, EQ, 1, True)])
    assert cnt == 4 and list(bm.iter_set_bits()) == [5, 31, 63, 99]
    bm.free()
    c.free()


def test_scan_rejects_bad_arguments(ctx):
    col = DeviceColumn.from_numpy(ctx, np.arange(10, dtype=np.int64))
    with pytest.raises(A.TgpError):
        ctx.scan_filter([col], [(3, EQ, 1, True)])          # predicate column out of range
    with pytest.raises(A.TgpError):
        ctx.scan_filter([col], [(0, EQ, 1, True)], 0, 11)   # row_end beyond the column
    col.free()


def _cells(a):
    a = np.asarray(a)
    return a.astype(np.int64).view(np.uint64) if a.dtype.kind == "i" else a.astype(np.uint64)


@pytest.mark.parametrize("da,db", [(np.int64, np.int64), (np.int32, np.int64), (np.uint64, np.uint64),
                                   (np.int8, np.uint8), (np.uint32, np.int16)])
def test_compare_columns_vs_closed_form(ctx, da, db):
    """tgp_scan_compare_columns (expr.rs:1382-1451 Col-op-Col): bit i stays set only
    where a op b, signed or u64 compare of the widened cells; narrows a selection;
    ragged and unaligned ranges, extremes of both widths."""
    from turbogp_amd import _capi as A
    rng = np.random.default_rng(np.dtype(da).itemsize * 10 + np.dtype(db).itemsize)
    n = 100_003
    ia, ib = np.iinfo(da), np.iinfo(db)
    a = rng.integers(max(ia.min, -40), min(ia.max, 40), n, endpoint=True).astype(da)
    b = rng.integers(max(ib.min, -40), min(ib.max, 40), n, endpoint=True).astype(db)
    a[::17], b[::19] = ia.max, ib.min
    a[5::23], b[7::29] = ia.min, ib.max
    b[::31] = a[::31].astype(db) if np.can_cast(da, db) else b[::31]     # some ties
    ca, cb = _cells(a), _cells(b)
    dev_a, dev_b = DeviceColumn.from_numpy(ctx, a), DeviceColumn.from_numpy(ctx, b)
    ops = {A.EQ: np.equal, A.NE: np.not_equal, A.LT: np.less, A.LE: np.less_equal,
           A.GT: np.greater, A.GE: np.greater_equal}
    try:
        for signed in (True, False):
            x = ca.view(np.int64) if signed else ca
            y = cb.view(np.int64) if signed else cb
            for op, f in ops.items():
                for b0, e0 in ((0, n), (37, n - 11), (64, 64 + 640), (5, 5)):
                    bm = ctx.compare_columns(dev_a, op, dev_b, signed, b0, e0)
                    try:
                        want = f(x[b0:e0], y[b0:e0])
                        got = np.unpackbits(bm.words().view(np.uint8), bitorder="little")
                        assert np.array_equal(got[:e0 - b0].astype(bool), want), (op, signed, b0)
                        assert not got[e0 - b0:].any()          # tail bits cleared
                    finally:
                        bm.free()
            # narrowing an existing selection
            sel, _ = ctx.scan_filter([dev_a], [(0, A.GE, 0, True)], 0, n)
            try:
                ctx.compare_columns(dev_a, A.LT, dev_b, signed, 0, n, sel)
                got = np.unpackbits(sel.words().view(np.uint8), bitorder="little")[:n].astype(bool)
                want = (ca.view(np.int64) >= 0) & ((x < y))
                assert np.array_equal(got, want), signed
            finally:
                sel.free()
    finally:
        dev_a.free()
        dev_b.free()
