"""The oracle's f64 filters and per-row expression evaluator, pinned by the
reference's known answers (exec/bitmap.rs f64 filter tests,
tests/interpreter_fallback.rs) and by numpy computing the same values another
way (wrapping int64 products, truncated remainders, f64 comparisons)."""
import math
import struct

import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import expr as X
from turbogp_amd import _capi as A


def bools(bits, n):
    return [bool((bits[i >> 3] >> (i & 7)) & 1) for i in range(n)]


def run(cols, e, kinds=None, b=0, e_=None):
    """Evaluates expression e over named host columns {name: array}."""
    names = list(cols)
    kinds = kinds or {}

    def ref(name):
        return names.index(name), kinds.get(name, X.INT)
    prog = X.compile_expr(e, ref)
    n = len(next(iter(cols.values())))
    end = n if e_ is None else e_
    return bools(B.scan_expr([cols[k] for k in names], prog, b, end), end - b)


# ---- exec/bitmap.rs:1168-1270 ----
def test_f64_filter_kats():
    five = [1.0, 2.5, 2.5, 3.0, 2.5]
    ten = [float(x) for x in range(1, 11)]
    assert bools(B.filter_f64(five, 1, 0, 2.5), 5) == [True, False, False, True, False]   # ne
    assert bools(B.filter_f64(five, 0, 0, 2.5), 5) == [False, True, True, False, True]    # eq
    assert bools(B.filter_f64(ten, 2, 0, 5.0), 10) == [x < 5 for x in ten]                # lt
    assert bools(B.filter_f64(ten, 3, 0, 5.0), 10) == [x <= 5 for x in ten]               # le
    assert bools(B.filter_f64(ten, 5, 0, 5.0), 10) == [x >= 5 for x in ten]               # ge
    assert bools(B.filter_f64(ten, 4, 0, 5.0), 10) == [x > 5 for x in ten]                # gt
    big = np.arange(1_000_000, dtype=np.float64) * 0.001                                  # large
    bits = B.filter_f64(big, 2, 0, 50.0)
    assert int(np.unpackbits(bits, bitorder="little").sum()) == 50_000


def test_f64_filters_vs_numpy_special_values():
    rng = np.random.default_rng(1)
    v = np.concatenate([rng.normal(0, 10, 5000), [0.0, -0.0, np.nan, np.inf, -np.inf, 5e-324,
                                                  2.5, 2.5 + 1e-7, 2.5 + 1e-5, 1e300]])
    for lit in (0.0, -0.0, 2.5, 1e-7, -3.25, np.inf):
        tol = 1e-6 * max(abs(lit), 1.0)
        with np.errstate(invalid="ignore"):
            want = {(0, 0): v == lit, (0, 1): np.abs(v - lit) <= tol,
                    (1, 0): ~(np.abs(v - lit) <= tol), (2, 0): v < lit, (3, 0): v <= lit,
                    (4, 0): v > lit, (5, 0): v >= lit}
        for (op, eps), w in want.items():
            assert bools(B.filter_f64(v, op, eps, lit), len(v)) == w.tolist(), (op, eps, lit)


# ---- tests/interpreter_fallback.rs (sales: id, region, amount, qty) ----
SALES = {"id": np.array([1, 2, 3, 4, 5], np.int64), "region": np.array([1, 1, 2, 2, 3], np.int64),
         "amount": np.array([100, 200, 150, 300, 250], np.int64),
         "qty": np.array([2, 3, 1, 5, 4], np.int64)}
R, AM = X.Col("region"), X.Col("amount")


def count(e):
    return sum(run(SALES, e))


def test_interpreter_fallback_kats():
    # :72-78 (ignored upstream: the reference's interpreter returned every row; 2 is the
    # answer the test asserts)
    e = X.BinOp("OR", X.BinOp("AND", X.BinOp("=", R, X.Lit(1)), X.BinOp(">", AM, X.Lit(150))),
                X.BinOp("AND", X.BinOp("=", R, X.Lit(2)), X.BinOp(">", AM, X.Lit(200))))
    assert count(e) == 2
    # :91-97 CASE WHEN amount > 200 THEN 1 ELSE 0 END = 1
    case = X.Case(((X.BinOp(">", AM, X.Lit(200)), X.Lit(1)),), X.Lit(0))
    assert count(X.BinOp("=", case, X.Lit(1))) == 2
    assert run(SALES, case) == [False, False, False, True, True]  # :107-116 as truthiness
    # :125-127 region IN (1, 2)
    assert count(X.InList(R, (X.Lit(1), X.Lit(2)))) == 4
    # :175-177 region = 1 OR region = 3
    assert count(X.BinOp("OR", X.BinOp("=", R, X.Lit(1)), X.BinOp("=", R, X.Lit(3)))) == 3
    # :184-187 amount BETWEEN 100 AND 250 AND region = 1
    assert count(X.BinOp("AND", X.Between(AM, X.Lit(100), X.Lit(250)),
                         X.BinOp("=", R, X.Lit(1)))) == 2
    # :51-54 sum(amount * qty) = 3450: every product in its row
    assert count(X.BinOp("=", X.BinOp("*", AM, X.Col("qty")), X.Lit(3450))) == 0
    prods = SALES["amount"] * SALES["qty"]
    for p in prods:
        assert run(SALES, X.BinOp("=", X.BinOp("*", AM, X.Col("qty")), X.Lit(int(p)))).count(True) \
            == int((prods == p).sum())


def test_arithmetic_vs_numpy():
    rng = np.random.default_rng(7)
    n = 20_000
    a = rng.integers(-2**62, 2**62, n).astype(np.int64)
    b = rng.integers(-2**40, 2**40, n).astype(np.int64)
    b[::17] = 0
    b[::31] = -1
    a[::29] = np.iinfo(np.int64).min
    c = rng.integers(-1000, 1000, n).astype(np.int64)
    cols = {"a": a, "b": b, "c": c}
    A_, B_, C_ = X.Col("a"), X.Col("b"), X.Col("c")
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        prod = (a * b).astype(np.float64)          # wrapping int64 product, then compared as f64
        k = 3.5e18
        assert run(cols, X.BinOp(">", X.BinOp("*", A_, B_), X.Lit(k))) == (prod > k).tolist()
        s = (a + c).astype(np.float64)
        assert run(cols, X.BinOp("<=", X.BinOp("+", A_, C_), X.Lit(-4611686018427387904))) == \
            (s <= -4611686018427387904.0).tolist()
        # / is f64 division; a zero divisor is NULL (false)
        q = a.astype(np.float64) / b.astype(np.float64)
        assert run(cols, X.BinOp(">", X.BinOp("/", A_, B_), X.Lit(1.5))) == \
            ((b != 0) & (q > 1.5)).tolist()
        # % of two Ints: truncated remainder (np.fmod on ints); zero divisor NULL
        r = np.where(b == -1, 0, np.fmod(a, np.where(b == 0, 1, b)))
        assert run(cols, X.BinOp("=", X.BinOp("%", A_, B_), X.Lit(0))) == \
            ((b != 0) & (r == 0)).tolist()
        assert run(cols, X.BinOp(">", X.BinOp("%", C_, X.Lit(7)), X.Lit(2))) == \
            (np.fmod(c, 7) > 2).tolist()
        # unary minus wraps
        neg = (0 - a.astype(np.uint64)).astype(np.int64).astype(np.float64)
        assert run(cols, X.BinOp("<", X.Neg(A_), X.Lit(0))) == (neg < 0).tolist()


def test_comparisons_are_f64_and_null_quirks():
    big = np.array([2**53, 2**53 + 1, 2**53 + 2, -(2**63)], np.int64)
    cols = {"x": big}
    x = X.Col("x")
    # 2^53 + 1 as f64 is 2^53: equal to both neighbours' rounding
    assert run(cols, X.BinOp("=", x, X.Lit(2**53 + 1))) == [True, True, False, False]
    assert run(cols, X.BinOp("=", x, X.Lit(None))) == [False] * 4
    assert run(cols, X.BinOp("<>", x, X.Lit(None))) == [True] * 4      # NOT of a false =
    assert run(cols, X.BinOp("<", x, X.Lit(None))) == [False] * 4
    assert run(cols, X.Not(X.BinOp("=", x, X.Lit(None)))) == [True] * 4
    assert run(cols, X.BinOp("/", x, X.Lit(0))) == [False] * 4         # NULL is not truthy
    assert run(cols, X.InList(x, (X.Lit(None), X.Lit(2**53)))) == [True, True, False, False]


def test_float_and_date_columns():
    f = np.array([0.0, -0.0, np.nan, 1.5, -2.0, np.inf], np.float64)
    d = np.array([0, 1, 2, 730, 731, 2525], np.int32)
    cols = {"f": f.view(np.uint64), "d": d}
    kinds = {"f": X.FLOAT, "d": X.DATE}
    fc, dc = X.Col("f"), X.Col("d")
    assert run(cols, fc, kinds) == [False, False, True, True, True, True]    # truthy Float
    assert run(cols, dc, kinds) == [False] * 6                               # a Date is never truthy
    assert run(cols, X.BinOp(">=", dc, X.Lit(731, X.DATE)), kinds) == (d >= 731).tolist()
    # Date + Int is Float arithmetic; Neg of a Date is NULL
    assert run(cols, X.BinOp(">", X.BinOp("+", dc, X.Lit(30)), X.Lit(760)), kinds) == \
        (d + 30 > 760).tolist()
    assert run(cols, X.Neg(dc), kinds) == [False] * 6
    with np.errstate(invalid="ignore"):
        assert run(cols, X.BinOp("<", X.BinOp("*", fc, X.Lit(2)), X.Lit(3.0)), kinds) == \
            (f * 2 < 3.0).tolist()
    assert run(cols, X.Between(fc, X.Lit(-1.0), X.Lit(1.5)), kinds) == \
        [True, True, False, True, False, False]


def test_row_range_and_existing_selection():
    v = np.arange(100, dtype=np.int64)
    prog = X.compile_expr(X.BinOp(">", X.Col("v"), X.Lit(50)), lambda n: (0, X.INT))
    sel = np.zeros(5, np.uint8)
    sel[0] = 0b10110101
    sel[4] = 0xFF
    got = bools(B.scan_expr([v], prog, 45, 85, sel), 40)
    want = [bool((sel[i >> 3] >> (i & 7)) & 1) and 45 + i > 50 for i in range(40)]
    assert got == want
