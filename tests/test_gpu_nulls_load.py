"""Loading the reference's u64 cells into native-width columns, and NULL
semantics on the dispatch path (tests/null_semantics.rs)."""
import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import DeviceBitmap, DeviceColumn, _capi as A
from turbogp_amd.dispatch import Aggregate, Column, SelectQuery, Star, Table, execute_dispatched

pytestmark = pytest.mark.gpu
DT = {np.int8: A.I8, np.uint8: A.U8, np.int16: A.I16, np.uint16: A.U16, np.int32: A.I32,
      np.uint32: A.U32, np.int64: A.I64, np.uint64: A.U64}


@pytest.mark.parametrize("dtype", list(DT))
def test_from_cells_narrows_and_checks(ctx, dtype):
    rng = np.random.default_rng(np.dtype(dtype).itemsize)
    info = np.iinfo(dtype)
    vals = rng.integers(info.min, info.max, 70_001, dtype=dtype, endpoint=True)
    vals[:2] = [info.min, info.max]
    cells = B.cells(vals)  # as the loader stores them: i64 as `v as u64`
    col = DeviceColumn.from_cells(ctx, cells, DT[dtype])
    try:
        assert col.dtype == DT[dtype] and np.array_equal(col.to_numpy(), vals)
    finally:
        col.free()
    if np.dtype(dtype).itemsize < 8:
        bad = cells.copy()
        bad[12_345] = np.uint64(info.max) + np.uint64(1)  # one past the type's range
        bad[50_000] = np.uint64(info.max) + np.uint64(7)
        with pytest.raises(RuntimeError, match="row 12345 does not fit"):
            DeviceColumn.from_cells(ctx, bad, DT[dtype])


def null_table(ctx, made):
    # tests/null_semantics.rs: (id, val) = (1,10) (2,NULL) (3,30) (4,NULL) (5,50)
    ids = DeviceColumn.from_cells(ctx, np.arange(1, 6, dtype=np.uint64), A.I64)
    val = DeviceColumn.from_cells(ctx, np.array([10, 0, 30, 0, 50], np.uint64), A.I64)
    nb = DeviceBitmap.from_null_flags(ctx, [0, 1, 0, 1, 0])
    made += [ids, val, nb]
    return Table(ctx, {"id": ids, "val": val}, {"val": nb})


def test_reference_null_semantics(ctx):
    made = []
    try:
        t = null_table(ctx, made)
        run = lambda *sel, **kw: execute_dispatched(SelectQuery(list(sel), **kw), t).scalar()
        assert run(Aggregate("count", "*")) == 5               # count_star_counts_all_rows
        assert run(Aggregate("count", "val")) == 3             # count_col_excludes_nulls
        assert run(Aggregate("sum", "val")) == 90              # sum_ignores_nulls
        assert run(Aggregate("avg", "val")) == "30.0000"       # avg_ignores_nulls (not 18)
        assert run(Aggregate("min", "val")) == 10              # min_ignores_nulls (not 0)
        assert run(Aggregate("max", "val")) == 50              # max_ignores_nulls
        assert run(Aggregate("count_distinct", "val")) == 3    # count_distinct_ignores_nulls
        assert run(Aggregate("count", "val"), where=[("id", ">", 2)]) == 2  # count_col_with_where
        # WHERE on the nullable column: NULL cells fail every comparison
        assert run(Aggregate("count", "*"), where=[("val", "<", 40)]) == 2
        r = execute_dispatched(SelectQuery([Column("id"), Column("val")], where=[("id", ">=", 2)]), t)
        assert r.columns[0].values == [2, 3, 4, 5] and r.columns[1].null_mask == [True, False, True, False]
        assert r.columns[0].null_mask is None
    finally:
        for x in made:
            x.free()


def test_grouped_aggregates_skip_nulls_per_column(ctx):
    rng = np.random.default_rng(21)
    n = 300_000
    grp = rng.integers(0, 7, n).astype(np.int64)
    a = rng.integers(-1000, 1000, n).astype(np.int64)
    b = rng.integers(0, 5000, n).astype(np.int64)
    na, nbm = rng.random(n) < 0.2, rng.random(n) < 0.5
    cols = {k: DeviceColumn.from_numpy(ctx, v) for k, v in (("grp", grp), ("a", a), ("b", b))}
    nulls = {"a": DeviceBitmap.from_null_flags(ctx, na), "b": DeviceBitmap.from_null_flags(ctx, nbm)}
    try:
        t = Table(ctx, cols, nulls)
        where = b > 100
        r = execute_dispatched(SelectQuery(
            [Column("grp"), Aggregate("count", "*"), Aggregate("sum", "a"), Aggregate("count", "a"),
             Aggregate("max", "b"), Aggregate("sum", (1, [("a", 0, 1), ("b", 0, 1)]))],
            where=[("b", ">", 100)], group_by=["grp"]), t)
        base = where & ~nbm  # WHERE on b: NULL b fails
        want = [[], [], [], [], [], []]
        for g in range(7):
            m = base & (grp == g)
            want[0].append(g)
            want[1].append(int(m.sum()))
            want[2].append(int(a[m & ~na].sum()))
            want[3].append(int((m & ~na).sum()))
            want[4].append(int(b[m].max()))
            want[5].append(int((a[m & ~na].astype(object) * b[m & ~na].astype(object)).sum()) % 2**64)
        got = [c.values for c in r.columns]
        got[2] = [v - 2**64 if v >= 2**63 else v for v in got[2]]
        assert got == want
    finally:
        for x in list(cols.values()) + list(nulls.values()):
            x.free()
