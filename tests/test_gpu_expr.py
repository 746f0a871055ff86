"""GPU parity of the f64 filters (tgp_scan_filter_f64) and the expression
predicates (tgp_scan_expr) with the oracle: the reference's known answers, and
random expression trees over columns of every width, FLOAT and DATE columns,
ragged row ranges and existing selections, bit for bit."""
import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import DeviceBitmap, DeviceColumn, expr as X, _capi as A

pytestmark = pytest.mark.gpu


def dev(ctx, a):
    return DeviceColumn.from_numpy(ctx, np.ascontiguousarray(a))


def device_bits(bm, n):
    w = bm.words()
    return np.unpackbits(w.view(np.uint8), bitorder="little")[:n].astype(bool)


def oracle_bits(bits, n):
    return np.unpackbits(bits, bitorder="little")[:n].astype(bool)


def test_f64_filter_kats_and_special_values(ctx):
    rng = np.random.default_rng(2)
    v = np.concatenate([[1.0, 2.5, 2.5, 3.0, 2.5], rng.normal(0, 10, 100_000),
                        [0.0, -0.0, np.nan, np.inf, -np.inf, 5e-324, 2.5 + 1e-7, 2.5 + 1e-5]])
    c = dev(ctx, v.view(np.uint64))
    try:
        for lit in (2.5, 0.0, -0.0, -3.25, np.inf, 1e-300):
            for op in range(6):
                for eps in (0, 1):
                    for b, e in ((0, len(v)), (3, len(v) - 70), (64, 129)):
                        bm, cnt = ctx.scan_filter_f64(c, op, lit, eps, b, e)
                        try:
                            got = device_bits(bm, e - b)
                        finally:
                            bm.free()
                        want = oracle_bits(B.filter_f64(v[b:e], op, eps, lit), e - b)
                        assert np.array_equal(got, want), (lit, op, eps, b, e)
                        assert cnt == int(want.sum())
        bm, _ = ctx.scan_filter_f64(c, A.NE, 2.5, False, 0, 5)   # bitmap.rs test_filter_ne_f64
        try:
            assert device_bits(bm, 5).tolist() == [True, False, False, True, False]
        finally:
            bm.free()
    finally:
        c.free()


SALES = {"id": np.array([1, 2, 3, 4, 5], np.int64), "region": np.array([1, 1, 2, 2, 3], np.int32),
         "amount": np.array([100, 200, 150, 300, 250], np.int64),
         "qty": np.array([2, 3, 1, 5, 4], np.int8)}


def run_both(ctx, host, e, kinds=None, b=0, end=None, sel=None):
    names = list(host)
    kinds = kinds or {}
    prog = X.compile_expr(e, lambda nm: (names.index(nm), kinds.get(nm, X.INT)))
    n = len(next(iter(host.values())))
    end = n if end is None else end
    cols = [dev(ctx, host[k]) for k in names]
    bm = None
    try:
        if sel is not None:
            bm = DeviceBitmap(ctx, end - b)
            words = np.zeros(bm.nwords, np.uint64)
            packed = np.packbits(sel, bitorder="little")
            words.view(np.uint8)[: len(packed)] = packed
            bm.buf.upload(words)
            ctx.scan_expr(cols, prog, b, end, bm)
        else:
            bm = ctx.scan_expr(cols, prog, b, end)
        got = device_bits(bm, end - b)
        tail = np.unpackbits(bm.words().view(np.uint8), bitorder="little")[end - b:]
        assert not tail[: bm.nwords * 64 - (end - b)].any(), "bits past the range must be clear"
    finally:
        if bm is not None:
            bm.free()
        for c in cols:
            c.free()
    obits = None if sel is None else np.packbits(sel, bitorder="little")
    want = oracle_bits(B.scan_expr([host[k] for k in names], prog, b, end, obits), end - b)
    assert np.array_equal(got, want), e
    return got


def test_interpreter_fallback_kats_on_device(ctx):
    R, AM = X.Col("region"), X.Col("amount")
    cases = [
        (X.BinOp("OR", X.BinOp("AND", X.BinOp("=", R, X.Lit(1)), X.BinOp(">", AM, X.Lit(150))),
                 X.BinOp("AND", X.BinOp("=", R, X.Lit(2)), X.BinOp(">", AM, X.Lit(200)))), 2),
        (X.BinOp("=", X.Case(((X.BinOp(">", AM, X.Lit(200)), X.Lit(1)),), X.Lit(0)), X.Lit(1)), 2),
        (X.InList(R, (X.Lit(1), X.Lit(2))), 4),
        (X.BinOp("OR", X.BinOp("=", R, X.Lit(1)), X.BinOp("=", R, X.Lit(3))), 3),
        (X.BinOp("AND", X.Between(AM, X.Lit(100), X.Lit(250)), X.BinOp("=", R, X.Lit(1))), 2),
    ]
    for e, want in cases:
        assert int(run_both(ctx, SALES, e).sum()) == want


def random_tree(rng, names, kinds, depth):
    """A random numeric or boolean expression over the named columns."""
    if depth == 0 or rng.random() < 0.25:
        r = rng.random()
        if r < 0.55:
            return X.Col(names[rng.integers(len(names))])
        if r < 0.75:
            return X.Lit(int(rng.integers(-300, 300)))
        if r < 0.9:
            return X.Lit(float(rng.normal(0, 100)))
        return X.Lit(None) if rng.random() < 0.5 else X.Lit(int(rng.integers(0, 3000)), X.DATE)
    kind = rng.integers(0, 9)
    sub = lambda: random_tree(rng, names, kinds, depth - 1)  # noqa: E731
    if kind <= 3:
        return X.BinOp(["+", "-", "*", "/", "%"][rng.integers(5)], sub(), sub())
    if kind <= 5:
        return X.BinOp(["=", "<>", "<", "<=", ">", ">="][rng.integers(6)], sub(), sub())
    if kind == 6:
        return X.BinOp(["AND", "OR"][rng.integers(2)], sub(), sub())
    if kind == 7:
        r = rng.integers(4)
        if r == 0:
            return X.Not(sub())
        if r == 1:
            return X.Neg(sub())
        if r == 2:
            return X.Between(sub(), sub(), sub(), bool(rng.integers(2)))
        return X.InList(sub(), (sub(), X.Lit(int(rng.integers(-5, 5)))), bool(rng.integers(2)))
    return X.Case(((sub(), sub()), (sub(), sub())), sub() if rng.random() < 0.7 else None)


@pytest.mark.parametrize("seed", range(4))
def test_random_expressions_vs_oracle(ctx, seed):
    rng = np.random.default_rng(100 + seed)
    n = 70_001
    host = {"i8": rng.integers(-128, 128, n).astype(np.int8),
            "u16": rng.integers(0, 2**16, n).astype(np.uint16),
            "i32": rng.integers(-2**31, 2**31, n).astype(np.int32),
            "i64": rng.integers(-2**62, 2**62, n).astype(np.int64),
            "u64": rng.integers(0, 2**63, n, dtype=np.uint64) * np.uint64(2) + np.uint64(1),
            "f": rng.normal(0, 50, n).astype(np.float64).view(np.uint64),
            "d": rng.integers(0, 2526, n).astype(np.int32),
            "small": rng.integers(-3, 4, n).astype(np.int64)}
    host["f"][::97] = np.array([np.nan], np.float64).view(np.uint64)[0]
    kinds = {"f": X.FLOAT, "d": X.DATE}
    names = list(host)
    done = 0
    while done < 12:
        e = random_tree(rng, names, kinds, 4)
        prog = X.compile_expr(e, lambda nm: (names.index(nm), kinds.get(nm, X.INT)))
        depth = mx = 0
        for code, arg, _ in prog:  # keep programs within the device's limits
            pops = {A.EX_COL: 0, A.EX_INT: 0, A.EX_FLOAT: 0, A.EX_DATE: 0, A.EX_NULL: 0,
                    A.EX_NOT: 1, A.EX_NEG: 1, A.EX_BETWEEN: 3, A.EX_NOT_BETWEEN: 3, A.EX_IF: 3,
                    A.EX_IN: arg + 1, A.EX_NOT_IN: arg + 1}.get(code, 2)
            depth += 1 - pops
            mx = max(mx, depth)
        if len(prog) > A.MAX_EXPR_OPS or mx > A.MAX_EXPR_DEPTH:
            continue
        b, end = int(rng.integers(0, 300)), n - int(rng.integers(0, 300))
        sel = rng.random(end - b) < 0.7 if done % 2 else None
        run_both(ctx, host, e, kinds, b, end, sel)
        done += 1


def test_malformed_programs_are_rejected(ctx):
    c = dev(ctx, np.arange(10, dtype=np.int64))
    try:
        for prog in ([(A.EX_INT, 0, 1), (A.EX_INT, 0, 2)],           # two values left
                     [(A.EX_ADD, 0, 0)],                             # pops an empty stack
                     [(A.EX_COL, 3, X.INT)],                         # no column 3
                     [(A.EX_COL, 0, 7)],                             # no such column kind
                     [(99, 0, 0)],
                     [(A.EX_INT, 0, 1)] * 9 + [(A.EX_ADD, 0, 0)] * 8):  # deeper than 8
            with pytest.raises(A.TgpError):
                ctx.scan_expr([c], prog)
    finally:
        c.free()
