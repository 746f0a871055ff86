"""Dispatch of SELECT shapes onto the device operators: the reference's
dispatch tests (engine/dispatch.rs:1776-1860) and its engine tests for these
shapes, then parity with the oracle on a 1M-row fact table."""
import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import DeviceColumn, _capi as A
from turbogp_amd import workloads as W
from turbogp_amd.dispatch import (Aggregate, Column, Join, QueryShape, SelectQuery, Star, Table,
                                  classify_query, execute_dispatched, execute_with_join)

pytestmark = pytest.mark.gpu


@pytest.fixture()
def make(ctx):
    made = []

    def build(**cols):
        dev = {k: DeviceColumn.from_numpy(ctx, np.ascontiguousarray(v)) for k, v in cols.items()}
        made.extend(dev.values())
        return Table(ctx, dev)
    yield build
    for c in made:
        c.free()


def ref_table(make, n=100):  # dispatch.rs make_table: id = i, val = i % 20, grp = i % 5
    i = np.arange(n, dtype=np.uint64)
    return make(id=i, val=i % 20, grp=i % 5)


def test_reference_dispatch_tests(make):
    t = ref_table(make)
    run = lambda q: execute_dispatched(q, t)
    assert run(SelectQuery([Aggregate("count", "*")])).scalar() == 100
    assert run(SelectQuery([Aggregate("count", "*")], where=[("val", "=", 5)])).scalar() == 5
    assert run(SelectQuery([Aggregate("sum", "val")], where=[("val", ">", 15)])).scalar() == 350
    assert run(SelectQuery([Column("grp"), Aggregate("count", "*")], group_by=["grp"])) \
        .row_count == 5
    r = run(SelectQuery([Column("grp"), Aggregate("count", "*")], group_by=["grp"],
                        order_by=[("grp", True)], limit=3))
    assert r.row_count == 3 and r.columns[0].values == [0, 1, 2] and r.columns[1].values == [20] * 3
    assert run(SelectQuery([Aggregate("max", "val")])).scalar() == 19
    assert run(SelectQuery([Aggregate("count_distinct", "val")])).scalar() == 20
    assert run(SelectQuery([Column("grp")], group_by=["grp"])) is None  # Complex: not dispatched


def test_reference_engine_tests(make):
    # tests/dispatch_path_features.rs:14-35: SUM(price * 2) = 600, SUM(a * b) = 1400
    t = make(price=np.array([100, 200], np.int64), discount=np.array([10, 20], np.int64))
    assert execute_dispatched(SelectQuery([Aggregate("sum", (2, [("price", 0, 1)]))]), t) \
        .scalar() == 600
    t = make(a=np.array([10, 30], np.int64), b=np.array([20, 40], np.int64))
    assert execute_dispatched(SelectQuery([Aggregate("sum", (1, [("a", 0, 1), ("b", 0, 1)]))]),
                              t).scalar() == 1400
    # tests/interpreter_fallback.rs sales table: avg 200, min 100, max 300, 3 regions
    s = make(id=np.arange(1, 6, dtype=np.int64), region=np.array([1, 1, 2, 2, 3], np.int64),
             amount=np.array([100, 200, 150, 300, 250], np.int64),
             qty=np.array([2, 3, 1, 5, 4], np.int64))
    assert execute_dispatched(SelectQuery([Aggregate("avg", "amount")]), s).scalar() == "200.0000"
    assert execute_dispatched(SelectQuery([Aggregate("min", "amount")]), s).scalar() == 100
    assert execute_dispatched(SelectQuery([Aggregate("max", "amount")]), s).scalar() == 300
    assert execute_dispatched(SelectQuery([Aggregate("count_distinct", "region")]), s).scalar() == 3
    # tests/joins_groupby_orderby.rs:82-106: grp 1 -> sum 60, count 3; grp 2 -> 90, 2
    g = make(grp=np.array([1, 1, 1, 2, 2], np.int64), a=np.array([10, 20, 30, 40, 50], np.int64))
    r = execute_dispatched(SelectQuery([Column("grp"), Aggregate("sum", "a"),
                                        Aggregate("count", "*")], group_by=["grp"],
                                       order_by=[("grp", True)]), g)
    assert [c.values for c in r.columns] == [[1, 2], [60, 90], [3, 2]]
    # tests/joins_groupby_orderby.rs:112-155: ORDER BY, DESC, LIMIT
    m = make(a=np.array([5, 1, 3, 2, 4], np.int64), b=np.array([50, 10, 30, 20, 40], np.int64))
    r = execute_dispatched(SelectQuery([Column("a"), Column("b")], order_by=[("a", True)],
                                       limit=2), m)
    assert r.columns[0].values == [1, 2] and r.columns[1].values == [10, 20]
    r = execute_dispatched(SelectQuery([Column("a"), Column("b")], order_by=[("a", False)]), m)
    assert r.columns[0].values == [5, 4, 3, 2, 1]
    r = execute_dispatched(SelectQuery([Star()], where=[("a", ">", 2)], limit=2), m)
    assert [c.values for c in r.columns] == [[5, 3], [50, 30]]
    # an aggregate over a column the table lacks is an error, not column 0
    with pytest.raises(KeyError):
        execute_dispatched(SelectQuery([Aggregate("sum", "nope")]), m)


def test_fact_table_parity_with_the_oracle(ctx, make):
    n = 1_000_000
    host = W.host_fact_columns(n, B.synth_uniform)
    t = make(**host)
    where = [("dt", ">=", 731), ("dt", "<", 1096), ("disc", ">=", 5), ("disc", "<=", 7),
             ("qty", "<", 24)]
    sel = ((host["dt"] >= 731) & (host["dt"] < 1096) & (host["disc"] >= 5) & (host["disc"] <= 7)
           & (host["qty"] < 24))
    run = lambda q: execute_dispatched(q, t)
    assert run(SelectQuery([Aggregate("count", "*")], where=where)).scalar() == int(sel.sum())
    rev = (1, [("price", 0, 1), ("disc", 0, 1)])
    want = int((host["price"][sel].astype(object) * host["disc"][sel].astype(object)).sum())
    assert run(SelectQuery([Aggregate("sum", rev)], where=where)).scalar() == want
    assert run(SelectQuery([Aggregate("min", "price")], where=where)).scalar() == \
        int(host["price"][sel].min())
    assert run(SelectQuery([Aggregate("count_distinct", "price")], where=where)).scalar() == \
        len(np.unique(host["price"][sel]))
    r = run(SelectQuery([Column("flaga"), Aggregate("sum", "qty"), Aggregate("count", "*")],
                        where=[("dt", "<=", 2435)], group_by=["flaga"]))
    m = host["dt"] <= 2435
    assert r.columns[0].values == [0, 1, 2]
    assert r.columns[1].values == [int(host["qty"][m & (host["flaga"] == k)].sum()) for k in range(3)]
    assert r.columns[2].values == [int((m & (host["flaga"] == k)).sum()) for k in range(3)]
    r = run(SelectQuery([Column("price"), Column("qty")], where=where,
                        order_by=[("price", False)], limit=50))
    idx = np.flatnonzero(sel)
    order = idx[np.argsort(~B.cells(host["price"])[idx], kind="stable")][:50]
    assert r.columns[0].values == [int(v) for v in host["price"][order]]
    assert r.columns[1].values == [int(v) for v in host["qty"][order]]


def test_reference_join_engine_tests(make):
    """tests/joins_groupby_orderby.rs:19-59 through execute_with_join (executor.rs:1747):
    LEFT keeps unmatched left rows, INNER drops them, COUNT(*) over a LEFT join with
    no matches counts every left row."""
    orders = make(id=np.array([1, 2, 3], np.int64), cust=np.array([10, 20, 30], np.int64))
    returns = make(oid=np.array([1], np.int64), reason=np.array([99], np.int64))
    cat = {"orders": orders, "returns": returns}
    star = SelectQuery([Star()])
    r = execute_with_join(star, "orders", [Join("returns", [("id", "oid")], "LEFT")], cat)
    assert r.row_count == 3
    by_name = {c.name: c.values for c in r.columns}
    assert by_name["orders.id"] == [1, 2, 3] and by_name["returns.reason"] == [99, 0, 0]
    r = execute_with_join(star, "orders", [Join("returns", [("orders.id", "returns.oid")],
                                                 "INNER")], cat)
    assert r.row_count == 1
    o = make(id=np.array([1, 2, 3, 4], np.int64))
    rr = make(oid=np.zeros(0, np.int64))
    r = execute_with_join(SelectQuery([Aggregate("count", "*")]), "o",
                          [Join("r", [("id", "oid")], "LEFT")], {"o": o, "r": rr})
    assert r.scalar() == 4


def _joined_rows(left, right, on, jt):
    """Python reference of hash_join (exec/join.rs:68-160) and cross_join_into
    (executor.rs:1886-1948): (left row or None, right row or None) in output order."""
    if jt == "CROSS":
        return [(i, j) for i in range(len(next(iter(left.values()))))
                for j in range(len(next(iter(right.values()))))]
    bcols = [right[b] for _, b in on]
    pcols = [left[a] for a, _ in on]
    p, b = B.hash_join(bcols, pcols, 1 if jt in ("LEFT", "FULL") else 0)
    rows = [(int(i), None if int(x) == 0xFFFFFFFF else int(x)) for i, x in zip(p, b)]
    if jt in ("RIGHT", "FULL"):
        seen = {x for _, x in rows if x is not None}
        rows += [(None, j) for j in range(len(bcols[0])) if j not in seen]
    return rows


@pytest.mark.parametrize("jt", ["INNER", "LEFT", "RIGHT", "FULL", "CROSS"])
def test_join_then_dispatch_vs_reference(make, jt):
    """Every join type over a one- and a two-column key, then SELECT * and a GROUP BY
    SUM over the joined table, against the Python restatement of hash_join."""
    rng = np.random.default_rng(len(jt))
    nl, nr = (60, 25) if jt == "CROSS" else (3000, 800)
    left = {"k1": rng.integers(0, 40, nl).astype(np.int64),
            "k2": rng.integers(0, 3, nl).astype(np.int32),
            "v": rng.integers(-1000, 1000, nl).astype(np.int64)}
    right = {"rk1": rng.integers(0, 50, nr).astype(np.int64),
             "rk2": rng.integers(0, 3, nr).astype(np.int64),
             "g": rng.integers(0, 7, nr).astype(np.uint8)}
    cat = {"l": make(**left), "r": make(**right)}
    for on in ([("k1", "rk1")], [("k1", "rk1"), ("k2", "rk2")]):
        want = _joined_rows(left, right, on, jt)
        r = execute_with_join(SelectQuery([Star()]), "l", [Join("r", on, jt)], cat)
        assert r.row_count == len(want)
        got = {c.name: c.values for c in r.columns}
        cell = lambda a, i: 0 if i is None else int(B.cells(a)[i])
        for name, arr in left.items():
            assert got[f"l.{name}"] == [cell(arr, i) for i, _ in want], (jt, on, name)
        for name, arr in right.items():
            assert got[f"r.{name}"] == [cell(arr, j) for _, j in want], (jt, on, name)
        # GROUP BY a right column, SUM of a left column, over the joined rows
        q = SelectQuery([Column("r.g"), Aggregate("sum", "l.v")], group_by=["r.g"])
        r = execute_with_join(q, "l", [Join("r", on, jt)], cat)
        sums = {}
        for i, j in want:
            g = 0 if j is None else int(right["g"][j])
            sums[g] = sums.get(g, 0) + (0 if i is None else int(left["v"][i]))
        keys = [int(x) for x in r.columns[0].values]
        vals = [int(np.int64(np.uint64(x))) for x in r.columns[1].values]
        assert dict(zip(keys, vals)) == sums, (jt, on)


def test_where_column_vs_column(make):
    """A Q21-style column comparison (l_receiptdate > l_commitdate) in WHERE, with a
    literal predicate beside it, through the dispatch mirror against numpy."""
    rng = np.random.default_rng(21)
    n = 200_000
    commit = rng.integers(0, 2500, n).astype(np.int32)
    receipt = (commit + rng.integers(-30, 31, n)).astype(np.int32)
    qty = rng.integers(1, 51, n).astype(np.int64)
    t = make(l_commitdate=commit, l_receiptdate=receipt, qty=qty)
    late = receipt > commit
    q = SelectQuery([Aggregate("count", "*")], where=[("l_receiptdate", ">", Column("l_commitdate"))])
    assert execute_dispatched(q, t).scalar() == int(late.sum())
    q = SelectQuery([Aggregate("sum", "qty")],
                    where=[("qty", "<", 24), ("l_receiptdate", ">", Column("l_commitdate"))])
    assert execute_dispatched(q, t).scalar() == int(qty[late & (qty < 24)].sum())
    q = SelectQuery([Aggregate("count", "*")], where=[("l_receiptdate", "<=", Column("l_commitdate")),
                                                     ("l_commitdate", ">=", 100)])
    assert execute_dispatched(q, t).scalar() == int(((receipt <= commit) & (commit >= 100)).sum())


def test_reference_having_tests(make):
    """basic_parser_features.rs:76-98 and interpreter_fallback.rs:58-67 through the
    accumulator path mirror (execute_grouped): HAVING count(*) and HAVING sum(...) when
    the sum is not selected."""
    from turbogp_amd.dispatch import execute_grouped
    t = make(dept=np.array([1, 1, 1, 2, 2, 3], np.int64),
             salary=np.array([100, 200, 300, 150, 250, 350], np.int64))
    r = execute_grouped(SelectQuery([Column("dept"), Aggregate("count", "*")], group_by=["dept"],
                                    having=(Aggregate("count", "*"), ">", 2)), t)
    assert r.row_count == 1 and r.columns[0].values == [1]
    r = execute_grouped(SelectQuery([Column("dept")], group_by=["dept"],
                                    having=(Aggregate("sum", "salary"), ">", 400)), t)
    assert r.row_count == 1 and [c.name for c in r.columns] == ["dept"] and r.columns[0].values == [1]
    sales = make(region=np.array([1, 1, 2, 2, 3], np.int64), amount=np.array([100, 200, 150, 300, 250], np.int64))
    r = execute_grouped(SelectQuery([Column("region"), Aggregate("count", "*")], group_by=["region"],
                                    having=(Aggregate("count", "*"), ">", 1)), sales)
    assert r.row_count == 2


def test_having_order_limit_vs_numpy(make):
    """HAVING before ORDER BY and LIMIT, on exact signed sums (negative ones too)."""
    from turbogp_amd.dispatch import execute_grouped
    rng = np.random.default_rng(7)
    n = 100_000
    g = rng.integers(0, 300, n).astype(np.int64)
    v = rng.integers(-1000, 900, n).astype(np.int64)
    t = make(g=g, v=v)
    sums = {k: int(v[g == k].sum()) for k in np.unique(g)}
    q = SelectQuery([Column("g"), Aggregate("sum", "v", "s")], group_by=["g"],
                    having=("s", "<", -1000), order_by=[("s", True)], limit=10)
    r = execute_grouped(q, t)
    keys = [int(x) for x in r.columns[0].values]
    got = [int(np.uint64(x).astype(np.int64)) for x in r.columns[1].values]  # INT8 cells
    # ascending sums, ties in key order (the groups come in key order, the sort is stable)
    want = sorted((s, k) for k, s in sums.items() if s < -1000)[:10]
    assert list(zip(got, keys)) == want


def test_group_by_several_nullable_and_sparse_keys(ctx):
    """GROUP BY two columns, a nullable column and a sparse int64 column (the hash
    table path) against Python dicts over the same rows; NULL keys form their own
    group and carry a null_mask; groups come out in key order, NULL last."""
    from turbogp_amd import DeviceBitmap
    rng = np.random.default_rng(21)
    n = 200_000
    host = {"fa": rng.integers(0, 3, n).astype(np.uint8),
            "fb": rng.integers(0, 2, n).astype(np.uint8),
            "sid": rng.integers(-(2 ** 62), 2 ** 62, 3000, dtype=np.int64)[
                rng.integers(0, 3000, n)],
            "v": rng.integers(-(10 ** 12), 10 ** 12, n).astype(np.int64)}
    null_sid = rng.random(n) < 0.1
    null_v = rng.random(n) < 0.05
    dev = {k: DeviceColumn.from_numpy(ctx, v) for k, v in host.items()}
    nulls = {"sid": DeviceBitmap.from_null_flags(ctx, null_sid),
             "v": DeviceBitmap.from_null_flags(ctx, null_v)}
    try:
        t = Table(ctx, dev, nulls)
        sel = host["v"] > -(5 * 10 ** 11)

        def ref(keys):
            out = {}
            for i in np.flatnonzero(sel & ~null_v):
                k = tuple(None if (c == "sid" and null_sid[i]) else int(host[c][i]) for c in keys)
                c, s = out.get(k, (0, 0))
                out[k] = (c + 1, s + int(host["v"][i]))
            return dict(sorted(out.items(), key=lambda kv: [(1, 0) if x is None else (0, x)
                                                             for x in kv[0]]))

        for keys in (["fa", "fb"], ["sid"], ["fb", "sid"]):
            q = SelectQuery([Column(k) for k in keys] +
                            [Aggregate("count", "*"), Aggregate("sum", "v")],
                            where=[("v", ">", -(5 * 10 ** 11))], group_by=keys)
            r = execute_dispatched(q, t)
            want = ref(keys)
            got_keys = list(zip(*[[None if (c.null_mask and c.null_mask[i]) else
                                   (v - 2 ** 64 if v >= 2 ** 63 else v)
                                   for i, v in enumerate(c.values)]
                                  for c in r.columns[:len(keys)]]))
            assert got_keys == list(want), keys
            assert r.columns[len(keys)].values == [c for c, _ in want.values()]
            assert r.columns[len(keys) + 1].values == [s % 2 ** 64 for _, s in want.values()]
    finally:
        for c in list(dev.values()) + list(nulls.values()):
            c.free()


def test_where_expressions_and_float_columns(ctx):
    """WHERE terms the interpreter evaluates per row (arithmetic, OR, IN, CASE) run
    as device expressions; FLOAT columns filter through the f64 filters, `=` with
    apply_comparison's epsilon; the answers equal numpy over the same rows
    (tests/interpreter_fallback.rs shapes on a 1M-row table)."""
    from turbogp_amd import expr as X
    rng = np.random.default_rng(17)
    n = 1_000_003
    price = rng.integers(90_000, 10_500_000, n).astype(np.int64)
    disc = rng.integers(0, 11, n).astype(np.int8)
    region = rng.integers(1, 6, n).astype(np.int32)
    fprice = price.astype(np.float64) / 100.0
    fprice[::1001] = 2500.0 + 1e-7                          # within the epsilon of 2500.0
    dt = rng.integers(0, 2526, n).astype(np.int32)
    dev = {k: DeviceColumn.from_numpy(ctx, np.ascontiguousarray(v)) for k, v in
           {"price": price, "disc": disc, "region": region, "fprice": fprice.view(np.uint64),
            "dt": dt}.items()}
    t = Table(ctx, dev, types={"fprice": "float", "dt": "date"})
    try:
        P, D, R = X.Col("price"), X.Col("disc"), X.Col("region")
        prod = price * disc.astype(np.int64)
        q = SelectQuery([Aggregate("count", "*")],
                        where=[X.BinOp(">", X.BinOp("*", P, D), X.Lit(40_000_000)), ("dt", "<", 1500)])
        assert execute_dispatched(q, t).scalar() == int(((prod > 40_000_000) & (dt < 1500)).sum())
        e = X.BinOp("OR", X.BinOp("AND", X.BinOp("=", R, X.Lit(1)), X.BinOp(">", P, X.Lit(5_000_000))),
                    X.InList(R, (X.Lit(4), X.Lit(5))))
        want = ((region == 1) & (price > 5_000_000)) | np.isin(region, [4, 5])
        q = SelectQuery([Aggregate("sum", "price")], where=[e])
        assert execute_dispatched(q, t).scalar() == int(price[want].sum())
        case = X.Case(((X.BinOp(">", D, X.Lit(7)), X.Lit(1)),), X.Lit(0))
        q = SelectQuery([Aggregate("count", "*")], where=[X.BinOp("=", case, X.Lit(1))])
        assert execute_dispatched(q, t).scalar() == int((disc > 7).sum())
        q = SelectQuery([Aggregate("count", "*")],
                        where=[X.BinOp(">", X.BinOp("/", X.Col("fprice"), D), X.Lit(20_000.0))])
        with np.errstate(divide="ignore", invalid="ignore"):
            assert execute_dispatched(q, t).scalar() == int(((disc != 0) & (fprice / disc > 20_000.0)).sum())
        for op, lit, w in (("<", 1000.5, fprice < 1000.5), (">=", 50_000.25, fprice >= 50_000.25),
                           ("=", 2500.0, np.abs(fprice - 2500.0) <= 1e-6 * 2500.0),
                           ("<>", 2500.0, ~(np.abs(fprice - 2500.0) <= 1e-6 * 2500.0))):
            q = SelectQuery([Aggregate("count", "*")], where=[("fprice", op, lit)])
            assert execute_dispatched(q, t).scalar() == int(w.sum()), op
    finally:
        t.free()


def test_float_column_aggregates(ctx):
    """SUM / AVG of a FLOAT column sum doubles from +0.0 (sum_masked_f64 /
    avg_masked_f64, dispatch.rs:399-403, :428-432): the device's correctly rounded
    sum equals math.fsum of the selected values; MIN / MAX compare the raw cells as
    the reference's u64 min / max does; shapes whose integer engine would add the
    doubles' bit patterns (GROUP BY, products) go back to the interpreter."""
    import math
    import struct
    rng = np.random.default_rng(23)
    n = 300_007
    f = rng.normal(0, 1e3, n)
    f[::97] = -0.0
    f[::89] = 1e15 + rng.random(len(f[::89]))        # magnitudes that round in a plain sum
    k = rng.integers(0, 100, n).astype(np.int32)
    g = rng.integers(0, 7, n).astype(np.int32)
    dev = {"f": DeviceColumn.from_numpy(ctx, f.view(np.uint64)),
           "k": DeviceColumn.from_numpy(ctx, k), "g": DeviceColumn.from_numpy(ctx, g)}
    t = Table(ctx, dev, types={"f": "float"})
    bits = lambda v: struct.unpack("<Q", struct.pack("<d", v))[0]  # noqa: E731
    try:
        for where, m in (([], np.ones(n, bool)), ([("k", "<", 40)], k < 40)):
            r = execute_dispatched(SelectQuery([Aggregate("sum", "f")], where=where), t)
            assert r.columns[0].values == [bits(math.fsum(f[m]))]
            r = execute_dispatched(SelectQuery([Aggregate("avg", "f")], where=where), t)
            assert r.columns[0].values == [bits(math.fsum(f[m]) / int(m.sum()))]
        r = execute_dispatched(SelectQuery([Aggregate("avg", "f")], where=[("k", "<", 0)]), t)
        assert r.columns[0].values == [0]                    # no rows: 0
        cells = f.view(np.uint64)
        r = execute_dispatched(SelectQuery([Aggregate("max", "f")], where=[("k", "<", 40)]), t)
        assert r.scalar() == int(cells[k < 40].max())
        assert execute_dispatched(SelectQuery([Aggregate("sum", "f")], group_by=["g"]), t) is None
        prod = Aggregate("sum", (1, [("f", 0, 1), ("k", 0, 1)]))
        assert execute_dispatched(SelectQuery([prod]), t) is None
    finally:
        t.free()
