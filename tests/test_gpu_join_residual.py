"""GPU parity of the non-equi ON conditions of a join (tgp_join_residual_sync,
query_interpreter/join.rs:1146-1195 with eval_non_equi_match :1233-1300) with the
oracle's restatement, over pairs of device joins with duplicate keys and misses,
columns of every width and kind on both sides, INNER and LEFT, probe row ranges;
pairs outside the tables rejected; the dispatch mirror end to end."""
import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import DeviceColumn, HashTable, expr as X, _capi as A
from turbogp_amd.dispatch import (Aggregate, Join, SelectQuery, Star, Table, execute_with_join,
                                  split_on)

pytestmark = pytest.mark.gpu


def dev(ctx, a):
    return DeviceColumn.from_numpy(ctx, np.ascontiguousarray(a))


def tables(rng, nl, nr, kmax=400):
    left = {"k": rng.integers(0, kmax, nl).astype(np.int64),
            "a8": rng.integers(-128, 128, nl).astype(np.int8),
            "a32": rng.integers(-2**31, 2**31, nl).astype(np.int32),
            "af": rng.normal(0, 50, nl).astype(np.float64).view(np.uint64),
            "ad": rng.integers(0, 2526, nl).astype(np.int32)}
    right = {"k": rng.integers(0, int(kmax * 1.2), nr).astype(np.int64),
             "b16": rng.integers(0, 2**16, nr).astype(np.uint16),
             "b64": rng.integers(-2**62, 2**62, nr).astype(np.int64),
             "bd": rng.integers(0, 2526, nr).astype(np.int32)}
    return left, right


KINDS = {"af": X.FLOAT, "ad": X.DATE, "bd": X.DATE}
L = lambda n: X.Col("l." + n)  # noqa: E731
R = lambda n: X.Col("r." + n)  # noqa: E731
RESIDUALS = [
    X.BinOp("<", L("a8"), X.BinOp("%", R("b16"), X.Lit(200))),
    X.BinOp("AND", X.BinOp(">=", L("ad"), R("bd")), X.BinOp("<", L("af"), X.Lit(10.5))),
    X.BinOp("OR", X.Between(R("b64"), L("a32"), X.BinOp("*", L("a32"), X.Lit(-3))),
            X.InList(L("a8"), (X.Lit(1), X.Lit(2), X.Lit(3)))),
    X.Not(X.BinOp("=", X.BinOp("+", L("ad"), X.Lit(30)), R("bd"))),
    X.BinOp(">", X.BinOp("/", L("a32"), X.BinOp("-", R("b16"), R("b16"))), X.Lit(0)),  # NULL
    X.Case(((X.BinOp(">", R("b16"), X.Lit(40000)), X.BinOp("<", L("a8"), X.Lit(0))),),
           X.BinOp(">", L("af"), X.Lit(0.0))),
]


def device_vs_oracle(ctx, left, right, e, jt, b=0, end=None):
    end = len(left["k"]) if end is None else end
    lnames = [n for n in X.columns(e) if n.startswith("l.")]
    rnames = [n for n in X.columns(e) if n.startswith("r.")]
    order = lnames + rnames
    prog = X.compile_expr(e, lambda nm: (order.index(nm), KINDS.get(nm[2:], X.INT)))
    bk, pk = dev(ctx, right["k"]), dev(ctx, left["k"])
    lc = [dev(ctx, left[n[2:]]) for n in lnames]
    rc = [dev(ctx, right[n[2:]]) for n in rnames]
    t = HashTable(ctx, bk)
    try:
        p, bb = t.join_pairs(pk, A.JOIN_LEFT if jt == "LEFT" else A.JOIN_INNER, b, end)
        got = ctx.join_residual(lc, rc, prog, p, bb, left=jt == "LEFT", row_begin=b, row_end=end)
    finally:
        t.free()
        for c in [bk, pk] + lc + rc:
            c.free()
    op, ob = B.hash_join(right["k"], left["k"], 1 if jt == "LEFT" else 0, b, end)
    assert np.array_equal(p, op) and np.array_equal(bb, ob)   # the equi pairs themselves
    want = B.join_residual([left[n[2:]] for n in lnames], [right[n[2:]] for n in rnames], prog,
                           op, ob, jt == "LEFT")
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), (e, jt)
    return got


@pytest.mark.parametrize("k", range(len(RESIDUALS)))
@pytest.mark.parametrize("jt", ["INNER", "LEFT"])
def test_residual_vs_oracle(ctx, k, jt):
    rng = np.random.default_rng(10 + k)
    left, right = tables(rng, 50_000, 30_000)
    got = device_vs_oracle(ctx, left, right, RESIDUALS[k], jt)
    if jt == "LEFT":  # every probe row keeps at least one pair
        assert len(np.unique(got[0])) == len(left["k"])
    device_vs_oracle(ctx, left, right, RESIDUALS[k], jt, 777, 40_001)


def test_large_join_and_edge_cases(ctx):
    rng = np.random.default_rng(3)
    left, right = tables(rng, 2_000_000, 300_000, kmax=250_000)
    e = X.BinOp(">", X.BinOp("+", L("a8"), R("b16")), X.Lit(30_000))
    for jt in ("INNER", "LEFT"):
        device_vs_oracle(ctx, left, right, e, jt)
    # no pair passes / every pair passes
    device_vs_oracle(ctx, left, right, X.BinOp("<", R("b16"), X.Lit(0)), "LEFT")
    device_vs_oracle(ctx, left, right, X.Lit(1), "INNER")
    # an empty probe range
    device_vs_oracle(ctx, left, right, e, "LEFT", 5, 5)


def test_pairs_outside_the_tables_are_rejected(ctx):
    lk, rv = dev(ctx, np.arange(10, dtype=np.int64)), dev(ctx, np.arange(5, dtype=np.int64))
    prog = [(A.EX_COL, 0, X.INT), (A.EX_COL, 1, X.INT), (A.EX_LT, 0, 0)]
    try:
        for p, b, rb, re_ in (([3, 12], [0, 1], 0, 10),     # probe row past row_end
                              ([3, 4], [0, 5], 0, 10),      # build row past the right columns
                              ([1, 4], [0, 1], 2, 10)):     # probe row before row_begin
            with pytest.raises(A.TgpError):
                ctx.join_residual([lk], [rv], prog, np.array(p), np.array(b), row_begin=rb,
                                  row_end=re_)
        with pytest.raises(A.TgpError):  # a column index past the two tables
            ctx.join_residual([lk], [rv], [(A.EX_COL, 2, X.INT)], np.array([1]), np.array([1]))
        got = ctx.join_residual([lk], [rv], prog, np.array([1, 1, 4]), np.array([0, 3, 2]))
        assert got[0].tolist() == [1] and got[1].tolist() == [3]
    finally:
        lk.free()
        rv.free()


def test_dispatch_join_with_residual(ctx):
    """execute_with_join with non-equi ON conjuncts, split from one ON expression as the
    interpreter splits it, against nested loops in Python."""
    rng = np.random.default_rng(4)
    nl, nr = 4000, 900
    host_l = {"id": rng.integers(0, 300, nl).astype(np.int64),
              "qty": rng.integers(0, 50, nl).astype(np.int32)}
    host_r = {"oid": rng.integers(0, 350, nr).astype(np.int64),
              "lim": rng.integers(0, 50, nr).astype(np.int16),
              "w": rng.integers(-100, 100, nr).astype(np.int64)}
    cat = {"o": Table(ctx, {k: dev(ctx, v) for k, v in host_l.items()}),
           "r": Table(ctx, {k: dev(ctx, v) for k, v in host_r.items()})}
    try:
        on = X.BinOp("AND", X.BinOp("=", X.Col("o.id"), X.Col("r.oid")),
                     X.BinOp("AND", X.BinOp("<", X.Col("qty"), X.Col("r.lim")),
                             X.BinOp("=", X.Col("qty"), X.Col("o.qty"))))  # one table: dropped
        keys, residual = split_on(on, cat["o"], cat["r"], "r")
        assert keys == [("id", "oid")]
        assert residual == X.BinOp("<", X.Col("qty"), X.Col("r.lim"))
        for jt in ("INNER", "LEFT"):
            want = []
            for i in range(nl):
                m = [j for j in range(nr) if host_r["oid"][j] == host_l["id"][i]
                     and host_l["qty"][i] < host_r["lim"][j]]
                want += [(i, j) for j in m] or ([(i, None)] if jt == "LEFT" else [])
            r = execute_with_join(SelectQuery([Star()]), "o", [Join("r", keys, jt, residual)], cat)
            got = {c.name: c.values for c in r.columns}
            assert r.row_count == len(want), jt
            assert got["o.qty"] == [int(host_l["qty"][i]) for i, _ in want]
            assert got["r.w"] == [0 if j is None else int(host_r["w"][j]) % 2**64
                                  for _, j in want]
            s = execute_with_join(SelectQuery([Aggregate("sum", "r.w")]), "o",
                                  [Join("r", keys, jt, residual)], cat)
            # an INT8 result column holds the sum's two's complement cell
            assert s.scalar() == sum(0 if j is None else int(host_r["w"][j])
                                     for _, j in want) % 2**64
        # RIGHT / FULL with non-equi conditions: handed back
        assert execute_with_join(SelectQuery([Star()]), "o",
                                 [Join("r", keys, "FULL", residual)], cat) is None
    finally:
        for t in cat.values():
            t.free()
