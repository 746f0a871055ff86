"""QueryShape classification (engine/dispatch.rs:225-299), host only: the
reference's classify_* tests (dispatch.rs:1740-1772) and the other shapes."""
from turbogp_amd.dispatch import Aggregate, Column, QueryShape, SelectQuery, Star, classify_query


def test_reference_classify_tests():
    assert classify_query(SelectQuery([Aggregate("count", "*")])) == QueryShape.CountAll
    assert classify_query(SelectQuery([Aggregate("count", "*")], where=[("id", "=", 5)])) == \
        QueryShape.CountFilter
    assert classify_query(SelectQuery([Aggregate("sum", "val")])) == QueryShape.SumCol
    assert classify_query(SelectQuery([Column("grp"), Aggregate("count", "*")],
                                      group_by=["grp"])) == QueryShape.GroupByCount


def test_other_shapes():
    agg = lambda f, a="val": SelectQuery([Aggregate(f, a)])
    assert classify_query(agg("count")) == QueryShape.CountFilter  # COUNT(col), no WHERE
    assert classify_query(agg("avg")) == QueryShape.AvgCol
    assert classify_query(agg("min")) == QueryShape.MinMax
    assert classify_query(agg("max")) == QueryShape.MinMax
    assert classify_query(agg("count_distinct")) == QueryShape.CountDistinct
    assert classify_query(agg("stddev")) == QueryShape.Complex
    g = [Column("grp"), Aggregate("count", "*")]
    assert classify_query(SelectQuery(g, group_by=["grp"], order_by=[("grp", True)], limit=3)) \
        == QueryShape.GroupByOrderByLimit
    assert classify_query(SelectQuery(g, group_by=["grp"], order_by=[("grp", True)])) == \
        QueryShape.GroupByCount
    assert classify_query(SelectQuery([Column("grp"), Aggregate("sum", "val")],
                                      group_by=["grp"])) == QueryShape.GroupBySum
    assert classify_query(SelectQuery([Column("grp")], group_by=["grp"])) == QueryShape.Complex
    assert classify_query(SelectQuery([Star()])) == QueryShape.SelectStar
    assert classify_query(SelectQuery([Column("a")])) == QueryShape.SelectColumn
    assert classify_query(SelectQuery([Column("a"), Column("b")])) == QueryShape.SelectMulti
    assert classify_query(SelectQuery([Column("a"), Aggregate("sum", "b")])) == QueryShape.Complex


def test_having_is_not_dispatched():
    """dispatch.rs:329-331: a query with HAVING goes to the interpreter (None)."""
    from turbogp_amd.dispatch import Aggregate, Column, SelectQuery, execute_dispatched
    q = SelectQuery([Column("dept"), Aggregate("count", "*")], group_by=["dept"],
                    having=(Aggregate("count", "*"), ">", 2))
    assert execute_dispatched(q, None) is None


def test_unsupported_details_return_none():
    """Shapes the device path does not take fall back (None), as dispatch.rs:304-333
    does, instead of raising: a non-grouped SELECT column, COUNT(DISTINCT) per
    group, five GROUP BY columns, a LIKE filter.  These are decided before any
    device work, so a table of column descriptors without a device suffices."""
    from turbogp_amd import _capi as A
    from turbogp_amd.dispatch import Table, execute_dispatched, execute_grouped

    class Col:
        def __init__(self, dtype):
            self.n, self.dtype = 8, dtype

    t = Table(None, {c: Col(A.I64) for c in "abcdef"})
    cnt = Aggregate("count", "*")
    for q in (SelectQuery([Column("a"), Column("b"), cnt], group_by=["a"]),
              SelectQuery([Column("a"), Aggregate("count_distinct", "b")], group_by=["a"]),
              SelectQuery([Column("a"), cnt], group_by=list("abcde")),
              SelectQuery([Column("a"), cnt], group_by=["a", "a"]),
              SelectQuery([Aggregate("sum", "a")], where=[("b", "LIKE", 3)])):
        assert execute_dispatched(q, t) is None, q
    assert execute_grouped(SelectQuery([Column("a"), Column("b"), cnt], group_by=["a"]),
                           t) is None
