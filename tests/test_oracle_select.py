"""CPU checks of the SELECT-rows oracle: the reference's ORDER BY / LIMIT tests and
brute force against Python's stable sort of (u64 cell, scan order)."""
import numpy as np
import pytest

from oracle import binding as B


def test_reference_order_by_kats():  # tests/joins_groupby_orderby.rs:112-155
    a, b = np.array([3, 1, 2], np.int64), np.array([30, 10, 20], np.int64)
    rows = B.select_rows(0, 3, order_col=a)
    assert list(B.cells(a)[rows.astype(np.int64)]) == [1, 2, 3]
    assert list(B.cells(b)[rows.astype(np.int64)]) == [10, 20, 30]
    a = np.array([1, 3, 2], np.int64)
    assert list(a[B.select_rows(0, 3, order_col=a, descending=True).astype(np.int64)]) == [3, 2, 1]
    a = np.array([5, 1, 3, 2, 4], np.int64)
    assert list(a[B.select_rows(0, 5, order_col=a, limit=2).astype(np.int64)]) == [1, 2]


@pytest.mark.parametrize("seed", range(4))
def test_select_rows_vs_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 400))
    col = rng.integers(-6, 6, n).astype(np.int32)  # many ties; negatives sort last as u64 cells
    b, e = sorted(int(x) for x in rng.integers(0, n + 1, 2))
    words = rng.integers(0, 2**63, (max(1, e - b) + 63) // 64).astype(np.uint64)
    limit = int(rng.integers(0, n + 2))
    for desc in (False, True):
        sel = [r for r in range(b, e) if (int(words[(r - b) >> 6]) >> ((r - b) & 63)) & 1]
        key = lambda r: int(col[r]) % 2**64
        want = sorted(sel, key=key, reverse=False)
        if desc:  # stable descending: larger cells first, ties in scan order
            want = sorted(sel, key=lambda r: -key(r))
        got = B.select_rows(b, e, words, col, desc, limit)
        assert [int(x) for x in got] == want[:limit]
        assert [int(x) for x in B.select_rows(b, e, words, None, desc, limit)] == sel[:limit]
