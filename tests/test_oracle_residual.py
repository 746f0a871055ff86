"""The oracle's residual (non-equi) join filter, orc_join_residual, pinned by
brute force: nested loops over small tables with each residual written as plain
Python, for INNER and LEFT joins (query_interpreter/join.rs:1146-1195: a LEFT row
whose matches all fail is kept once with no build row), and by the reference's
LEFT / INNER JOIN known answers (tests/joins_groupby_orderby.rs:18-59) under an
always-true residual."""
import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import expr as X

NO = B.NO_BUILD_ROW


def brute(lkey, rkey, lcols, rcols, pred, left):
    out = []
    for l in range(len(lkey)):
        matches = [r for r in range(len(rkey)) if rkey[r] == lkey[l]]
        passing = [r for r in matches if pred({k: int(v[l]) for k, v in lcols.items()},
                                              {k: int(v[r]) for k, v in rcols.items()})]
        out += [(l, r) for r in passing]
        if left and not passing:
            out.append((l, NO))
    return out


def run(lkey, rkey, lcols, rcols, e, left):
    names = list(lcols) + list(rcols)
    prog = X.compile_expr(e, lambda nm: (names.index(nm), X.INT))
    p, b = B.hash_join(rkey, lkey, 1 if left else 0)
    op, ob = B.join_residual([lcols[k] for k in lcols], [rcols[k] for k in rcols], prog, p, b,
                             left)
    return list(zip(op.tolist(), ob.tolist()))


A_, B_, C_ = X.Col("a"), X.Col("b"), X.Col("c")
RESIDUALS = [
    (X.BinOp("<", A_, B_), lambda L, R: L["a"] < R["b"]),
    (X.BinOp(">", X.BinOp("+", A_, B_), X.Lit(10)), lambda L, R: L["a"] + R["b"] > 10),
    (X.Between(B_, A_, X.BinOp("+", A_, X.Lit(5))), lambda L, R: L["a"] <= R["b"] <= L["a"] + 5),
    (X.BinOp("=", X.BinOp("%", X.BinOp("*", A_, C_), X.Lit(3)), X.Lit(0)),
     lambda L, R: (L["a"] * R["c"]) % 3 == 0),   # divisibility: either remainder sign
    (X.Not(X.BinOp("=", A_, C_)), lambda L, R: L["a"] != R["c"]),
    # a / (b - b) is NULL: never truthy, so every match fails
    (X.BinOp(">", X.BinOp("/", A_, X.BinOp("-", B_, B_)), X.Lit(0)), lambda L, R: False),
    (X.InList(C_, (A_, X.Lit(7))), lambda L, R: R["c"] in (L["a"], 7)),
]


@pytest.mark.parametrize("k", range(len(RESIDUALS)))
@pytest.mark.parametrize("left", [False, True])
def test_residual_vs_brute_force(k, left):
    e, pred = RESIDUALS[k]
    rng = np.random.default_rng(k)
    nl, nr = 300, 120
    lkey = rng.integers(0, 40, nl).astype(np.int64)
    rkey = rng.integers(0, 50, nr).astype(np.int64)        # duplicates and misses
    lcols = {"a": rng.integers(-20, 20, nl).astype(np.int32)}
    rcols = {"b": rng.integers(-20, 20, nr).astype(np.int64),
             "c": rng.integers(-10, 10, nr).astype(np.int16)}
    got = run(lkey, rkey, lcols, rcols, e, left)
    assert got == brute(lkey, rkey, lcols, rcols, pred, left)


def test_always_true_residual_keeps_the_reference_join_kats():
    true = X.Lit(1)
    # orders (id, cust) LEFT / INNER JOIN returns (oid, reason) ON o.id = r.oid
    oid, cust = np.array([1, 2, 3], np.int64), np.array([10, 20, 30], np.int64)
    roid, reason = np.array([1], np.int64), np.array([99], np.int64)
    lc, rc = {"id": oid, "cust": cust}, {"oid": roid, "reason": reason}
    assert run(oid, roid, lc, rc, true, True) == [(0, 0), (1, NO), (2, NO)]     # 3 rows
    assert run(oid, roid, lc, rc, true, False) == [(0, 0)]                      # 1 row
    o = np.array([1, 2, 3, 4], np.int64)
    r = np.zeros(0, np.int64)
    assert len(run(o, r, {"id": o}, {"oid": r}, true, True)) == 4               # count 4
    # a residual no match passes: LEFT keeps every order once, INNER none
    never = X.BinOp(">", X.Col("reason"), X.Lit(100))
    assert run(oid, roid, lc, rc, never, True) == [(0, NO), (1, NO), (2, NO)]
    assert run(oid, roid, lc, rc, never, False) == []
