"""Run-time compiled shape-specialised kernels (TGP_TUNE_JIT): queries whose
shape has no kernel compiled ahead of time (other column dtypes and orders,
other predicate sets, divided keys, selection bitmaps) run the same kernel
bodies compiled with hipRTC; every kernel form must equal the generic engine
and the oracle bit for bit, and a shape compiles once per process."""
import ctypes

import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import DeviceBitmap, DeviceColumn, make_agg_query, _capi as A
from turbogp_amd import workloads as W

pytestmark = pytest.mark.gpu
EQ, NE, LT, LE, GT, GE = range(6)
N = 300_007


def jit_info():
    c, f, l, ms = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_double()
    A.check(A.lib().tgp_jit_info(ctypes.byref(c), ctypes.byref(f), ctypes.byref(l),
                                 ctypes.byref(ms)), "tgp_jit_info")
    return c.value, f.value, l.value, ms.value


def explain(q):
    st, sh = ctypes.c_int32(), ctypes.c_int32()
    A.check(A.lib().tgp_group_agg_explain(ctypes.byref(q), ctypes.byref(st), ctypes.byref(sh)),
            "explain")
    return sh.value


@pytest.fixture(scope="module")
def table(ctx):
    """The fact columns at their generated widths plus narrower and wider copies
    (the layouts a loader that narrows to native widths produces)."""
    h = W.host_fact_columns(N, B.synth_uniform)
    h["price32"] = h["price"].astype(np.int32)
    h["price_u64"] = h["price"].astype(np.uint64)
    h["qty8"] = h["qty"].astype(np.int8)
    h["disc8"] = h["disc"].astype(np.int8)
    h["tax16"] = h["tax"].astype(np.int16)
    h["dt16"] = h["dt"].astype(np.int16)
    h["dt_u32"] = h["dt"].astype(np.uint32)
    h["flaga16"] = h["flaga"].astype(np.uint16)
    rng = np.random.default_rng(3)
    h["neg"] = rng.integers(-2**40, 2**40, N).astype(np.int64)
    d = {k: DeviceColumn.from_numpy(ctx, np.ascontiguousarray(v)) for k, v in h.items()}
    yield h, d
    for c in d.values():
        c.free()


def forms(ctx, q):
    """The run-time compiled kernel in its forms (default choice, replicated
    histograms with static and chunked schedules, per-slice register fold), then
    the generic engine."""
    out = []
    try:
        ctx.tune(A.TUNE_JIT, 2)
        out.append(ctx.group_agg(q))
        ctx.tune(A.TUNE_SPEC_KERNEL, 2)
        out.append(ctx.group_agg(q))
        ctx.tune(A.TUNE_CHUNK, 3)
        out.append(ctx.group_agg(q))
        ctx.tune(A.TUNE_CHUNK, 0)
        ctx.tune(A.TUNE_SPEC_KERNEL, 1)
        out.append(ctx.group_agg(q))
        ctx.tune(A.TUNE_SPEC_KERNEL, 0)
        ctx.tune(A.TUNE_JIT, 1)
        out.append(ctx.group_agg(q))
    finally:
        ctx.tune(A.TUNE_JIT, 0)
        ctx.tune(A.TUNE_SPEC_KERNEL, 0)
        ctx.tune(A.TUNE_CHUNK, 0)
    return out


def check(ctx, table, names, preds, parts, aggs, b=0, e=None, sel=None):
    h, d = table
    q = make_agg_query([d[c] for c in names], preds, parts, aggs, b, e,
                       sel_bitmap=sel[1] if sel else None)
    assert explain(q) == A.SPEC_SHAPE_RUNTIME
    got = forms(ctx, q)
    oq = B.make_query([h[c] for c in names], preds, parts, aggs, b, e,
                      sel[0] if sel else None)
    cnt, sums, oor = B.group_agg(oq)
    for i, g in enumerate(got):
        assert [int(x) for x in g.count] == [int(x) for x in cnt], i
        assert g.sums() == sums, i
        assert g.n_out_of_range == oor, i


def q1_cols(price="price", qty="qty", disc="disc", tax="tax", dt="dt", flaga="flaga"):
    return [dt, qty, price, disc, tax, flaga, "flagb"]


Q1_AGGS = [("sum", 1, [(1, 0, 1)]), ("sum", 1, [(2, 0, 1)]),
           ("sum", 1, [(2, 0, 1), (3, 100, -1)]),
           ("sum", 1, [(2, 0, 1), (3, 100, -1), (4, 100, 1)])]


@pytest.mark.parametrize("group", [[(5, 1, 0, 3)], [(5, 1, 0, 3), (6, 1, 0, 2)], [(0, 30, 0, 85)], []])
def test_q1_price_int32(ctx, table, group):
    """The benchmark's Q1 with price stored as int32."""
    check(ctx, table, q1_cols(price="price32"), [(0, LE, 2435, False)], group, Q1_AGGS)


def test_q1_narrow_columns_reordered(ctx, table):
    """Every column at its narrowest width, in another order."""
    names = ["flagb", "tax16", "qty8", "flaga16", "price32", "dt16", "disc8"]
    aggs = [("sum", 1, [(2, 0, 1)]), ("sum", 1, [(4, 0, 1)]),
            ("sum", 1, [(4, 0, 1), (6, 100, -1)]),
            ("sum", 1, [(4, 0, 1), (6, 100, -1), (1, 100, 1)])]
    check(ctx, table, names, [(5, LE, 2000, True)], [(3, 1, 0, 3), (0, 1, 0, 2)], aggs, 17, N - 5)


def test_q6_narrow_and_unsigned(ctx, table):
    names = ["dt_u32", "disc8", "qty8", "price_u64"]
    preds = [(0, GE, 731, False), (0, LT, 1096, False), (1, GE, 5, True), (1, LE, 7, True),
             (2, LT, 24, True)]
    check(ctx, table, names, preds, [], [("sum", 1, [(3, 0, 1), (1, 0, 1)])])


def test_general_factors_count_nonzero_and_negative_values(ctx, table):
    names = ["neg", "disc", "flaga", "dt"]
    aggs = [("sum", -3, [(0, 5, 7), (1, 2, -1)]), ("count_nonzero", 1), ("sum", 1, [(0, 0, 1)])]
    check(ctx, table, names, [(3, NE, 100, True), (1, GT, 2, True)], [(2, 1, 0, 3)], aggs, 1, N)


def test_selection_bitmap_and_ragged_range(ctx, table):
    h, _ = table
    rng = np.random.default_rng(11)
    b, e = 300, N - 1001
    flags = rng.random(e - b) < 0.3
    words = np.zeros((e - b + 63) // 64, np.uint64)
    idx = np.flatnonzero(flags)
    np.bitwise_or.at(words, idx >> 6, np.left_shift(np.uint64(1), (idx & 63).astype(np.uint64)))
    bm = DeviceBitmap(ctx, e - b)
    try:
        bm.buf.upload(words)
        check(ctx, table, q1_cols(price="price32"), [(0, LE, 2000, False)], [(5, 1, 0, 3)],
              Q1_AGGS, b, e, sel=(words, bm))
    finally:
        bm.free()


def test_many_groups_take_the_workgroup_table(ctx, table):
    """85 months x 3 flags: histogram words beyond a wave's registers."""
    check(ctx, table, q1_cols(qty="qty8"), [(0, GT, 100, True)], [(0, 30, 0, 85), (5, 1, 0, 3)],
          Q1_AGGS[:2])


def test_a_shape_compiles_once(ctx, table):
    h, d = table
    names = ["dt16", "price32"]
    q = make_agg_query([d[c] for c in names], [(0, LT, 1500, True)], [(0, 100, 0, 26)],
                       [("sum", 1, [(1, 0, 1)])])
    try:
        ctx.tune(A.TUNE_JIT, 2)
        r0 = ctx.group_agg(q)
        c0, f0, l0, _ = jit_info()
        r1 = ctx.group_agg(q)
        c1, f1, l1, ms = jit_info()
    finally:
        ctx.tune(A.TUNE_JIT, 0)
    assert (c1, f1) == (c0, f0) and l1 == l0 + 1 and ms > 0
    assert r0.sums() == r1.sums()
    # below 2^20 rows the default runs the generic engine without compiling
    q2 = make_agg_query([d[c] for c in names], [(0, LT, 1499, True)], [(0, 100, 0, 26)],
                        [("sum", 1, [(1, 0, 1), (1, 1, 1)])])
    ctx.group_agg(q2)
    assert jit_info()[0] == c1


def test_default_compiles_in_the_background(ctx):
    """TGP_TUNE_JIT 0 on >= 2^20 rows: the first call starts the compile and runs on
    the generic engine; once the compile is done, calls run the compiled kernel;
    every result equals the oracle."""
    import time
    n = (1 << 20) + 4097
    rng = np.random.default_rng(21)
    h = [rng.integers(-300, 300, n).astype(np.int16), rng.integers(-2**30, 2**30, n).astype(np.int32),
         rng.integers(0, 7, n).astype(np.uint8)]
    d = [DeviceColumn.from_numpy(ctx, x) for x in h]
    preds, parts, aggs = [(0, GE, -17, True)], [(2, 1, 0, 7)], [("sum", 2, [(1, 0, 1), (0, 3, 1)])]
    try:
        q = make_agg_query(d, preds, parts, aggs)
        assert explain(q) == A.SPEC_SHAPE_RUNTIME
        want = B.group_agg(B.make_query(h, preds, parts, aggs))
        c0, f0, l0, _ = jit_info()
        first = ctx.group_agg(q)
        deadline = time.time() + 60
        while jit_info()[0] == c0 and jit_info()[1] == f0 and time.time() < deadline:
            time.sleep(0.05)
        c1, f1, _, _ = jit_info()
        assert (c1, f1) == (c0 + 1, f0)
        later = ctx.group_agg(q)
        assert jit_info()[2] == l0 + 1  # this call ran the compiled kernel
        for g in (first, later):
            assert [int(x) for x in g.count] == [int(x) for x in want[0]]
            assert g.sums() == want[1] and g.n_out_of_range == want[2]
    finally:
        for c in d:
            c.free()


MINMAX_AGGS = [("min", 2), ("max", 1), ("min_u64", 4), ("max_u64", 0),
               ("max", 1, [(2, 0, 1), (3, 100, -1)]), ("sum", 1, [(2, 0, 1)])]


@pytest.mark.parametrize("group", [[], [(5, 1, 0, 3)], [(5, 1, 0, 3), (6, 1, 0, 2)],
                                   [(0, 30, 0, 85)]])
def test_min_max_shapes(ctx, table, group):
    """MIN / MAX (signed and raw u64 order, of a column and of an expression) beside a
    sum: order codes with a maximum in every kernel form (per-slice register fold,
    replicated histograms, workgroup table), decoded after the kernel."""
    check(ctx, table, q1_cols(), [(0, LE, 2435, False)], group, MINMAX_AGGS)


def test_min_max_negative_and_unsigned_values(ctx, table):
    names = ["neg", "price_u64", "flaga", "dt"]
    aggs = [("min", 0), ("max", 0), ("min_u64", 0), ("max_u64", 1),
            ("min", 1, [(0, 0, 1), (2, 1, 1)])]
    check(ctx, table, names, [(3, GT, 100, True)], [(2, 1, 0, 3)], aggs, 5, N - 3)
    # nothing selected: every group empty, values 0
    check(ctx, table, names, [(3, LT, 0, True)], [(2, 1, 0, 3)], aggs)
