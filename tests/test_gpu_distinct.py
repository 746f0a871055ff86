"""GPU COUNT(DISTINCT): HyperLogLog registers bit-exact against the oracle,
estimates equal, exact counts equal, the AUTO switch at 100 000 selected rows."""
import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import DeviceBitmap, DeviceColumn, _capi as A

pytestmark = pytest.mark.gpu
DTYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64]


def bitmap(ctx, words, nbits):
    bm = DeviceBitmap(ctx, nbits)
    bm.buf.upload(words)
    return bm


def rand_col(rng, dtype, n):
    info = np.iinfo(dtype)
    col = rng.integers(info.min, info.max, n, dtype=dtype, endpoint=True)
    # the all-ones u64 cell (the device set's empty marker): -1 sign-extended, or u64 max
    col[::97] = -1 if info.min < 0 else info.max
    return col


@pytest.mark.parametrize("dtype", DTYPES)
def test_registers_and_counts_vs_oracle(ctx, dtype):
    rng = np.random.default_rng(np.dtype(dtype).itemsize * 7 + np.iinfo(dtype).min % 5)
    n = 300_000
    host = rand_col(rng, dtype, n)
    col = DeviceColumn.from_numpy(ctx, host)
    try:
        for b, e in ((0, n), (12_345, 250_001), (77, 78), (5, 5)):
            sub = host[b:e]
            assert np.array_equal(ctx.hll_registers(col, b, e), B.hll_registers(sub)), (b, e)
            for mode in (A.DISTINCT_AUTO, A.DISTINCT_EXACT, A.DISTINCT_HLL):
                assert ctx.count_distinct(col, b, e, mode=mode) == B.count_distinct(sub, None, mode)
        # with a selection bitmap over [b, e)
        b, e = 1_000, 290_000
        words = rng.integers(0, 2**63, (e - b + 63) // 64).astype(np.uint64)
        words[-1] &= np.uint64((1 << ((e - b) % 64)) - 1)  # (e - b) % 64 == 40
        bm = bitmap(ctx, words, e - b)
        try:
            sub = host[b:e]
            assert np.array_equal(ctx.hll_registers(col, b, e, bm), B.hll_registers(sub, words))
            for mode in (A.DISTINCT_AUTO, A.DISTINCT_EXACT, A.DISTINCT_HLL):
                assert ctx.count_distinct(col, b, e, bm, mode) == B.count_distinct(sub, words, mode)
        finally:
            bm.free()
    finally:
        col.free()


def test_auto_switch_and_estimate(ctx):
    host = (np.arange(400_000, dtype=np.int64) * 7919) % 150_001
    col = DeviceColumn.from_numpy(ctx, host)
    try:
        assert ctx.count_distinct(col, 0, 99_999) == len(np.unique(host[:99_999]))
        regs = ctx.hll_registers(col, 0, 100_000)
        assert ctx.count_distinct(col, 0, 100_000) == A.lib().tgp_hll_estimate(regs.ctypes.data)
        assert A.lib().tgp_hll_estimate(regs.ctypes.data) == B.hll_estimate(regs)
        # the selected count decides, not the range: 150 000 rows, 75 000 selected -> exact
        words = np.full((150_000 + 63) // 64, 0x5555555555555555, np.uint64)
        words[-1] &= (1 << (150_000 % 64)) - 1
        bm = bitmap(ctx, words, 150_000)
        try:
            got = ctx.count_distinct(col, 0, 150_000, bm)
            assert got == len(np.unique(host[0:150_000:2])) == B.count_distinct(host[:150_000], words)
        finally:
            bm.free()
        full = ctx.count_distinct(col)
        assert full == B.count_distinct(host) and abs(full - 150_001) / 150_001 < 0.02
    finally:
        col.free()


def test_hll_kat_large(ctx):  # exec/hll.rs test_hll_large_approximate / with_duplicates
    a = np.arange(1_000_000, dtype=np.uint64)
    dup = np.concatenate([np.arange(500_000), np.arange(500_000)]).astype(np.uint64)
    for host, want in ((a, 1_000_000), (dup, 500_000)):
        col = DeviceColumn.from_numpy(ctx, host)
        try:
            got = ctx.count_distinct(col)
            assert got == B.count_distinct(host) and abs(got - want) / want < 0.02
        finally:
            col.free()


@pytest.mark.parametrize("shrink", [0, 12])
def test_exact_count_with_sketch_sized_set(ctx, shrink):
    """From 2^20 rows the exact set is sized from the HyperLogLog estimate; with the
    size cut by 2^12 it overflows (bounded probes) and the count is redone in a set
    sized by the row count.  Exact either way, with and without a bitmap, for many and
    for few distinct values, including the set's marker cell."""
    rng = np.random.default_rng(40 + shrink)
    n = 3_000_001
    many = rng.integers(-2**40, 2**40, n).astype(np.int64) // 1024     # ~2.6M distinct
    many[::1001] = -1                                                  # the marker cell
    few = rng.integers(1, 51, n).astype(np.int64)                      # 50 distinct
    words = rng.integers(0, 2**63, (n + 63) // 64).astype(np.uint64)
    words[-1] &= np.uint64((1 << (n % 64)) - 1)
    ctx.tune(A.TUNE_DISTINCT_SHRINK, shrink)
    try:
        for host in (many, few):
            col = DeviceColumn.from_numpy(ctx, host)
            bm = bitmap(ctx, words, n)
            try:
                assert ctx.count_distinct(col, mode=A.DISTINCT_EXACT) == len(np.unique(host))
                got = ctx.count_distinct(col, 0, n, bm, A.DISTINCT_EXACT)
                assert got == B.count_distinct(host, words, A.DISTINCT_EXACT)
            finally:
                bm.free()
                col.free()
    finally:
        ctx.tune(A.TUNE_DISTINCT_SHRINK, 0)
