"""GPU parity of the exact f64 SUM (tgp_sum_f64_sync) and of the KernelTable's
aggregate operators (AggregateSumF64, AggregateCountDistinct) with the oracle:
the reference's known answers (kernel/aggregate.rs:187-239), math.fsum on millions of
doubles of every magnitude, the closed forms a sequential sum gets wrong, NaN,
infinities, subnormals and the signs of zero, and a selection bitmap."""
import math

import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import DeviceBitmap, DeviceColumn
from turbogp_amd.kernel import KernelParams, KernelTable, MemoryTier, Operator

pytestmark = pytest.mark.gpu


def bits(xs):
    return np.ascontiguousarray(np.array(xs, np.float64).view(np.uint64))


def dev_sum(ctx, xs, sel=None, neg_zero_start=True):
    b = bits(xs)
    col = DeviceColumn.from_numpy(ctx, b) if len(b) else DeviceColumn(ctx, 0, 7)
    bm = None
    try:
        if sel is not None:
            bm = DeviceBitmap(ctx, len(b))
            bm.buf.upload(sel)
        return ctx.sum_f64(col, len(b), bm, neg_zero_start)
    finally:
        col.free()
        if bm:
            bm.free()


def same(a, b):
    if math.isnan(a) or math.isnan(b):
        return math.isnan(a) and math.isnan(b)
    return a == b and math.copysign(1, a) == math.copysign(1, b)


def test_closed_forms_and_specials(ctx):
    cases = [[1.0, 2.0, 3.0, 4.0, 5.0], [float(i) for i in range(100)], [],
             [2.0 ** 53, 1.0, 1.0], [1e308, 1e308, -1e308], [1e308, 1e308], [-1e308, -1e308],
             [0.1] * 10, [5e-324] * 3, [1e16, 1.0, -1e16], [2.0 ** 53, 1.0],
             [2.0 ** 53 + 2, 1.0], [1.0, math.nan], [math.inf, -math.inf],
             [math.inf, 1e308, 1e308], [-math.inf, 3.0], [-0.0, -0.0], [-0.0, 0.0],
             [1.5, -1.5], [-0.0], [2.0 ** -1074, -(2.0 ** -1074), 2.0 ** 1023]]
    for xs in cases:
        for nz in (True, False):
            want = B.sum_f64(bits(xs), neg_zero_start=nz)
            got = dev_sum(ctx, xs, neg_zero_start=nz)
            assert same(got, want), (xs, nz, got, want)


def test_random_doubles_equal_fsum(ctx):
    rng = np.random.default_rng(31)
    for n, lo, hi in ((1000, -1070, 1000), (3_000_000, -30, 30), (5_000_000, -1000, 1000),
                      (2_000_000, 0, 60)):
        xs = np.ldexp(rng.uniform(-1, 1, n), rng.integers(lo, hi, n))
        assert dev_sum(ctx, xs) == math.fsum(xs), n


def test_selection_bitmap(ctx):
    rng = np.random.default_rng(32)
    n = 100_003
    xs = np.ldexp(rng.uniform(-1, 1, n), rng.integers(-60, 60, n))
    keep = rng.random(n) < 0.3
    w = np.zeros((n + 63) // 64, np.uint64)
    idx = np.flatnonzero(keep)
    np.bitwise_or.at(w, idx >> 6, np.left_shift(np.uint64(1), (idx & 63).astype(np.uint64)))
    assert dev_sum(ctx, xs, w) == math.fsum(xs[keep])


def test_kernel_table_aggregate_operators(ctx):
    t = KernelTable(ctx)
    try:
        for tier in (MemoryTier.L3, MemoryTier.Hbm):
            k = t.select(Operator.AggregateSumF64, tier)
            assert k is not None and k.operator() == Operator.AggregateSumF64
            # aggregate.rs:196-201 scalar_sum_f64_works; :203-226 0..99 -> 4950
            for xs, want in (([1.0, 2.0, 3.0, 4.0, 5.0], 15.0),
                             ([float(i) for i in range(100)], 4950.0)):
                c = bits(xs)
                r = k.execute(c, np.zeros(64, np.uint8), KernelParams(cell_count=len(c)))
                assert r.sum == want and r.count == len(c)
        k = t.select(Operator.AggregateCountDistinct, MemoryTier.Ddr5)
        assert k is not None and k.operator() == Operator.AggregateCountDistinct
        cells = np.array([1, 2, 3, 2, 4, 1, 5], np.uint64)  # aggregate.rs:229-238 -> 5
        assert k.execute(cells, np.zeros(64, np.uint8),
                         KernelParams(cell_count=len(cells))).count == 5
        rng = np.random.default_rng(33)
        cells = rng.integers(0, 2 ** 63, 50_000, dtype=np.int64).view(np.uint64)[
            rng.integers(0, 50_000, 300_000)]
        assert k.execute(cells, None, KernelParams(cell_count=len(cells))).count == \
            len(np.unique(cells))
        assert k.execute(None, None, KernelParams(cell_count=0)).count == 0
        # device-resident input (tier Hbm: used in place)
        col = DeviceColumn.from_numpy(ctx, bits(np.arange(1000, dtype=np.float64)))
        try:
            ks = t.select(Operator.AggregateSumF64, MemoryTier.Hbm)
            assert ks.execute(col.buf.ptr, None, KernelParams(cell_count=1000)).sum == 499500.0
        finally:
            col.free()
    finally:
        t.close()
