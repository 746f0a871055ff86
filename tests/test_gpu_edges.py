"""Edge cases of the selection, distinct and join paths: limits next to the row
count, one-row inputs, a column made only of the distinct set's marker cell, and
empty probe ranges, each against the oracle or its closed form."""
import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import DeviceColumn, HashTable, _capi as A

pytestmark = pytest.mark.gpu


def dev(ctx, arr):
    return DeviceColumn.from_numpy(ctx, np.ascontiguousarray(arr))


def select(ctx, col, n, desc, limit):
    rows, cnt = ctx.select_rows(0, n, None, col, desc, limit)
    try:
        return rows.download(np.uint64, cnt) if cnt else np.zeros(0, np.uint64)
    finally:
        rows.free()


def test_top_k_limits_next_to_the_row_count(ctx):
    rng = np.random.default_rng(1)
    for n in (1, 2, 63, 64, 65, 4097):
        host = rng.integers(-5, 5, n).astype(np.int64)
        col = dev(ctx, host)
        try:
            for limit in sorted({1, max(1, n - 1), n, n + 1}):
                for desc in (False, True):
                    got = select(ctx, col, n, desc, limit)
                    assert np.array_equal(got, B.select_rows(0, n, None, host, desc, limit)), \
                        (n, limit, desc)
        finally:
            col.free()


def test_exact_distinct_of_marker_cells_only(ctx):
    """Every cell is the set's empty marker (-1 sign-extended: all ones): one distinct
    value, counted through the marker flag, also above the sketch-sizing threshold."""
    for n in (1, 1000, (1 << 20) + 7):
        col = dev(ctx, np.full(n, -1, np.int64))
        try:
            assert ctx.count_distinct(col, mode=A.DISTINCT_EXACT) == 1
        finally:
            col.free()


def test_joins_against_one_build_row_and_empty_ranges(ctx):
    one = dev(ctx, np.array([7], np.int64))
    probe = dev(ctx, np.array([7, 8, 7, 9], np.int64))
    k2a, k2b = dev(ctx, np.array([7], np.int64)), dev(ctx, np.array([1], np.int32))
    p2a, p2b = dev(ctx, np.array([7, 7, 8], np.int64)), dev(ctx, np.array([1, 2, 1], np.int32))
    t1 = HashTable(ctx, one)
    t2 = HashTable(ctx, [k2a, k2b])
    try:
        p, b = t1.join_pairs(probe, A.JOIN_LEFT)
        assert list(p) == [0, 1, 2, 3] and list(b) == [0, A.JOIN_NO_MATCH, 0, A.JOIN_NO_MATCH]
        p, b = t1.join_pairs(probe, A.JOIN_INNER, 1, 2)                # one row, no match
        assert len(p) == 0
        p, b = t1.join_pairs(probe, A.JOIN_LEFT, 2, 2)                 # empty range
        assert len(p) == 0
        assert t1.probe_count(probe, 3, 3) == 0
        p, b = t2.join_pairs([p2a, p2b], A.JOIN_LEFT)
        assert list(p) == [0, 1, 2] and list(b) == [0, A.JOIN_NO_MATCH, A.JOIN_NO_MATCH]
        bm = t2.join_mark([p2a, p2b])
        try:
            assert list(bm.iter_set_bits()) == [0]
        finally:
            bm.free()
    finally:
        t1.free()
        t2.free()
        for c in (one, probe, k2a, k2b, p2a, p2b):
            c.free()


def test_compare_columns_one_row(ctx):
    a, b = dev(ctx, np.array([3, 1], np.int64)), dev(ctx, np.array([2, 2], np.int64))
    try:
        bm = ctx.compare_columns(a, A.GT, b, True, 1, 2)   # row 1: 1 > 2 is false
        try:
            assert int(bm.words()[0]) == 0
        finally:
            bm.free()
        bm = ctx.compare_columns(a, A.GT, b, True, 0, 1)   # row 0: 3 > 2
        try:
            assert int(bm.words()[0]) == 1
        finally:
            bm.free()
    finally:
        a.free()
        b.free()
