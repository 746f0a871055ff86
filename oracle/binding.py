"""ctypes binding of liboracle.so (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  The structs have the product ABI's layout but hold host pointers.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_NP = {0: np.int8, 1: np.uint8, 2: np.int16, 3: np.uint16, 4: np.int32, 5: np.uint32,
       6: np.int64, 7: np.uint64}
_FROM_NP = {np.dtype(v): k for k, v in _NP.items()}


class OColumn(C.Structure):
    _fields_ = [("data", C.c_void_p), ("len", C.c_uint64), ("dtype", C.c_int32),
                ("_pad", C.c_int32)]


class OExprOp(C.Structure):
    _fields_ = [("code", C.c_int32), ("arg", C.c_int32), ("imm", C.c_int64)]


class OPred(C.Structure):
    _fields_ = [("col", C.c_int32), ("op", C.c_int32), ("is_signed", C.c_int32),
                ("_pad", C.c_int32), ("lit", C.c_int64)]


class OPart(C.Structure):
    _fields_ = [("col", C.c_int32), ("divisor", C.c_int32), ("base", C.c_int64),
                ("card", C.c_int64)]


class OAgg(C.Structure):
    _fields_ = [("kind", C.c_int32), ("nfactors", C.c_int32), ("fcol", C.c_int32 * 3),
                ("_pad", C.c_int32), ("coef", C.c_int64), ("fa", C.c_int64 * 3),
                ("fb", C.c_int64 * 3)]


class OQuery(C.Structure):
    _fields_ = [("row_begin", C.c_uint64), ("row_end", C.c_uint64), ("cols", OColumn * 8),
                ("ncols", C.c_int32), ("npreds", C.c_int32), ("nparts", C.c_int32),
                ("naggs", C.c_int32), ("preds", OPred * 8), ("parts", OPart * 2),
                ("aggs", OAgg * 6), ("sel_bitmap", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run make -C oracle")
        L = C.CDLL(LIB_PATH)
        P, U64, I64, I32 = C.c_void_p, C.c_uint64, C.c_int64, C.c_int32
        sig = {
            "orc_mix64": (U64, [U64]),
            "orc_synth_uniform": (None, [P, I32, U64, U64, I64, I64, U64, U64]),
            "orc_synth_cdf": (None, [P, I32, U64, U64, P, U64, P, U64, U64]),
            "orc_widen_cells": (None, [P, I32, U64, P]),
            "orc_filter_cells": (None, [P, U64, I32, I32, I64, P]),
            "orc_bitmap_and_inplace": (None, [P, P, U64]),
            "orc_bitmap_or_inplace": (None, [P, P, U64]),
            "orc_bitmap_not": (None, [P, P, U64]),
            "orc_bitmap_count_range": (U64, [P, U64, U64, U64]),
            "orc_set_bits": (U64, [P, U64, P]),
            "orc_scan_multi": (U64, [P, U64, I32, P, P, P]),
            "orc_query_filter": (I32, [C.POINTER(OQuery), P]),
            "orc_group_agg": (I32, [C.POINTER(OQuery), I32, I32, P, P, P, P]),
            "orc_join_agg": (I32, [P, P, U64, P, P, P, U64, U64, I32, I64, U64, I32, P, P, P]),
            "orc_join_pairs": (U64, [P, U64, P, U64, P, P, U64]),
            "orc_join_hash": (U64, [U64]),
            "orc_xxh3_u64": (U64, [U64]),
            "orc_hll_registers": (None, [P, I32, U64, P, P]),
            "orc_hll_estimate": (U64, [P]),
            "orc_count_distinct": (U64, [P, I32, U64, P, I32]),
            "orc_group_by_keys": (I64, [C.POINTER(OQuery), I32, P, P, U64, P, P, P, P]),
            "orc_simd_level": (I32, []),
            "orc_filter_f64": (None, [P, U64, I32, I32, C.c_double, P]),
            "orc_scan_expr": (I32, [P, I32, P, I32, U64, U64, P]),
            "orc_join_residual": (I64, [P, I32, I32, P, I32, I32, P, P, U64, C.c_uint32, P, P]),
            "orc_set_simd": (None, [I32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data if a is not None and a.size else None


def synth_uniform(dtype, seed, column_id, lo, hi, row_begin, row_end):
    out = np.empty(row_end - row_begin, dtype=_NP[dtype])
    lib().orc_synth_uniform(ptr(out), dtype, seed, column_id, lo, hi, row_begin, row_end)
    return out


def synth_cdf(dtype, seed, column_id, cdf, table, row_begin, row_end):
    out = np.empty(row_end - row_begin, dtype=_NP[dtype])
    cdf = np.ascontiguousarray(cdf, np.float64)
    table = np.ascontiguousarray(table, np.int64)
    lib().orc_synth_cdf(ptr(out), dtype, seed, column_id, ptr(cdf), len(cdf), ptr(table),
                        row_begin, row_end)
    return out


def filter_cells(cells, op, is_signed, lit):
    cells = np.ascontiguousarray(cells, np.uint64)
    bits = np.zeros((len(cells) + 7) // 8, np.uint8)
    lib().orc_filter_cells(ptr(cells), len(cells), op, int(is_signed), lit, ptr(bits))
    return bits


def set_bits(bits, nbits):
    out = np.empty(max(1, nbits), np.uint64)
    n = lib().orc_set_bits(ptr(bits), nbits, ptr(out))
    return out[:n]


def bitmap_not(bits, nbits):
    out = np.zeros_like(bits)
    lib().orc_bitmap_not(ptr(bits), ptr(out), nbits)
    return out


def count_range(bits, nbits, start, end):
    return lib().orc_bitmap_count_range(ptr(bits), nbits, start, end)


def scan_multi(cells, preds):
    """preds: [(target, op {0 Eq, 1 Gt, 2 Lt})] -> (count, mask)"""
    cells = np.ascontiguousarray(cells, np.uint64)
    ops = np.array([p[1] for p in preds] or [0], np.int32)
    tg = np.array([p[0] for p in preds] or [0], np.uint64)
    mask = np.zeros(1, np.uint64)
    c = lib().orc_scan_multi(ptr(cells), len(cells), len(preds), ptr(ops), ptr(tg), ptr(mask))
    return c, int(mask[0])


_MINMAX = ("min", "max", "min_u64", "max_u64")


def make_query(arrays, preds=(), parts=(), aggs=(), row_begin=0, row_end=None, sel_words=None):
    """Same description as turbogp_amd.device.make_agg_query over host arrays.
    Keeps references to the arrays on the returned object."""
    q = OQuery()
    q._keep = [np.ascontiguousarray(a) for a in arrays]
    n = len(q._keep[0]) if q._keep else 0
    q.row_begin, q.row_end = row_begin, n if row_end is None else row_end
    for i, a in enumerate(q._keep):
        q.cols[i] = OColumn(ptr(a), len(a), _FROM_NP[a.dtype], 0)
    q.ncols = len(arrays)
    for i, (ci, op, lit, sg) in enumerate(preds):
        q.preds[i] = OPred(ci, op, int(bool(sg)), 0, lit)
    q.npreds = len(preds)
    for i, (ci, div, base, card) in enumerate(parts):
        q.parts[i] = OPart(ci, div, base, card)
    q.nparts = len(parts)
    for i, ag in enumerate(aggs):
        a = OAgg()
        if ag[0] == "count_nonzero":
            a.kind, a.nfactors, a.coef = 1, 1, 1
            a.fcol[0] = ag[1]
        elif ag[0] in _MINMAX:
            # ('min' | 'max' | 'min_u64' | 'max_u64', col) or (..., coef, [(col, fa, fb), ...])
            a.kind = 2 + _MINMAX.index(ag[0])
            if len(ag) == 2:
                a.nfactors, a.coef = 1, 1
                a.fcol[0], a.fa[0], a.fb[0] = ag[1], 0, 1
            else:
                a.nfactors, a.coef = len(ag[2]), ag[1]
                for k, (ci, fa, fb) in enumerate(ag[2]):
                    a.fcol[k], a.fa[k], a.fb[k] = ci, fa, fb
        else:
            _, coef, factors = ag
            a.kind, a.nfactors, a.coef = 0, len(factors), coef
            for k, (ci, fa, fb) in enumerate(factors):
                a.fcol[k], a.fa[k], a.fb[k] = ci, fa, fb
        q.aggs[i] = a
    q.naggs = len(aggs)
    if sel_words is not None:
        q._keep.append(np.ascontiguousarray(sel_words, np.uint64))
        q.sel_bitmap = ptr(q._keep[-1])
    return q


def _i128(lo, hi):
    v = (int(hi) << 64) | int(lo)
    return v - (1 << 128) if v >> 127 else v


def group_agg(q, mode=0, nthreads=0):
    """-> (count[ngroups], sums[ngroups][naggs] as Python ints or floats, n_out_of_range)"""
    ng = 1
    for k in range(q.nparts):
        ng *= q.parts[k].card
    cnt = np.zeros(ng, np.uint64)
    sums = np.zeros(ng * max(1, q.naggs) * 2, np.uint64)
    f64 = np.zeros(ng * max(1, q.naggs), np.float64)
    oor = np.zeros(1, np.uint64)
    rc = lib().orc_group_agg(C.byref(q), mode, nthreads, ptr(cnt), ptr(sums), ptr(f64), ptr(oor))
    assert rc == 0
    if mode == 0:
        s = [[_i128(sums[(g * q.naggs + a) * 2], sums[(g * q.naggs + a) * 2 + 1])
              for a in range(q.naggs)] for g in range(ng)]
    else:
        s = [[float(f64[g * q.naggs + a]) for a in range(q.naggs)] for g in range(ng)]
    return cnt, s, int(oor[0])


def group_by_keys(q, key_cols, key_nulls=None, signed=None):
    """GROUP BY the key tuple of q.cols[key_cols] (orc_group_by_keys): a dict
    {key tuple (None for a NULL part; signed[k] reads key k as int64): (count, [sums])}
    with exact Python-int sums (MIN/MAX: the value).  key_nulls[k]: bitmap words
    (bit i <-> row row_begin + i) or None."""
    nk = len(key_cols)
    kc = np.asarray(key_cols, np.int32)
    nb = [None if b is None else np.ascontiguousarray(b, np.uint64)
          for b in (key_nulls or [None] * nk)]
    nbp = (C.c_void_p * nk)(*[ptr(b) if b is not None else None for b in nb])
    G = lib().orc_group_by_keys(C.byref(q), nk, ptr(kc), C.cast(nbp, C.c_void_p), 0, None, None,
                                None, None)
    assert G >= 0, G
    keys = np.zeros(max(1, G * nk), np.uint64)
    nulls = np.zeros(max(1, G), np.uint32)
    cnt = np.zeros(max(1, G), np.uint64)
    sums = np.zeros(max(1, G * q.naggs * 2), np.uint64)
    G2 = lib().orc_group_by_keys(C.byref(q), nk, ptr(kc), C.cast(nbp, C.c_void_p), G, ptr(keys),
                                 ptr(nulls), ptr(cnt), ptr(sums))
    assert G2 == G
    signed = signed or [True] * nk
    out = {}
    for g in range(G):
        t = []
        for k in range(nk):
            if (int(nulls[g]) >> k) & 1:
                t.append(None)
            else:
                v = int(keys[g * nk + k])
                t.append(v - (1 << 64) if signed[k] and v >> 63 else v)
        out[tuple(t)] = (int(cnt[g]), [_i128(sums[(g * q.naggs + a) * 2],
                                             sums[(g * q.naggs + a) * 2 + 1])
                                       for a in range(q.naggs)])
    return out


def query_filter(q):
    n = q.row_end - q.row_begin
    bits = np.zeros((n + 7) // 8 + 8, np.uint8)
    assert lib().orc_query_filter(C.byref(q), ptr(bits)) == 0
    return bits[: (n + 7) // 8]


def join_agg(build_keys, build_payload, probe_keys, probe_vals, probe_group, group_by, key_base,
             ngroups, row_begin=0, row_end=None, nthreads=0):
    bk = np.ascontiguousarray(build_keys, np.int64)
    bp = np.ascontiguousarray(build_payload, np.int32) if build_payload is not None else None
    pk = np.ascontiguousarray(probe_keys, np.int64)
    pv = np.ascontiguousarray(probe_vals, np.int64)
    pg = np.ascontiguousarray(probe_group, np.int32) if probe_group is not None else None
    row_end = len(pk) if row_end is None else row_end
    cnt = np.zeros(max(1, ngroups), np.uint64)
    sums = np.zeros(max(1, ngroups) * 2, np.uint64)
    oor = np.zeros(1, np.uint64)
    rc = lib().orc_join_agg(ptr(bk), ptr(bp), len(bk), ptr(pk), ptr(pv), ptr(pg), row_begin,
                            row_end, group_by, key_base, ngroups, nthreads, ptr(cnt), ptr(sums),
                            ptr(oor))
    assert rc == 0
    s = [_i128(sums[2 * g], sums[2 * g + 1]) for g in range(ngroups)]
    return cnt[:ngroups], s, int(oor[0])


def join_pairs(build_keys, probe_keys, capacity=None):
    bk = np.ascontiguousarray(build_keys, np.int64)
    pk = np.ascontiguousarray(probe_keys, np.int64)
    cap = capacity if capacity is not None else max(16, 4 * len(pk))
    op = np.zeros(cap, np.uint64)
    ob = np.zeros(cap, np.uint32)
    n = lib().orc_join_pairs(ptr(bk), len(bk), ptr(pk), len(pk), ptr(op), ptr(ob), cap)
    return op[:min(n, cap)], ob[:min(n, cap)], n


def hash_join(build_cols, probe_cols, join_type=0, row_begin=0, row_end=None,
              build_payload=None):
    """exec/join.rs:68-160 hash_join over one or more key columns (INNER = 0,
    LEFT = 1): for every probe row in order, the build rows whose cells equal the
    probe row's cells in every key column (the reference verifies the full key
    after probing, :102-111), in ascending build row; under LEFT a row without a
    match gives one pair with build 0xFFFFFFFF (the reference fills that row's
    build columns with 0).  Returns (probe rows u64, build payloads u32), the
    payload being the build row unless `build_payload` is given."""
    bcs = [cells(c) for c in (build_cols if isinstance(build_cols, (list, tuple)) else [build_cols])]
    pcs = [cells(c) for c in (probe_cols if isinstance(probe_cols, (list, tuple)) else [probe_cols])]
    assert len(bcs) == len(pcs)
    nb = len(bcs[0])
    row_end = len(pcs[0]) if row_end is None else row_end
    table = {}
    for r, key in enumerate(zip(*[c.tolist() for c in bcs])):
        table.setdefault(key, []).append(r)
    pay = None if build_payload is None else np.asarray(build_payload).astype(np.int64).tolist()
    out_p, out_b = [], []
    probe_lists = [c[row_begin:row_end].tolist() for c in pcs]
    for off, key in enumerate(zip(*probe_lists)):
        rows = table.get(key)
        if rows:
            vals = sorted(pay[r] & 0xFFFFFFFF for r in rows) if pay is not None else rows
            out_p.extend([row_begin + off] * len(vals))
            out_b.extend(vals)
        elif join_type == 1:
            out_p.append(row_begin + off)
            out_b.append(0xFFFFFFFF)
    assert nb >= 0
    return np.array(out_p, np.uint64), np.array(out_b, np.uint32)


def xxh3_u64(v):
    return int(lib().orc_xxh3_u64(int(v) % (1 << 64)))


def hll_registers(col, sel_words=None):
    """16384 HyperLogLog registers of the selected cells (exec/hll.rs)."""
    col = np.ascontiguousarray(col)
    regs = np.zeros(1 << 14, np.uint8)
    sel = None if sel_words is None else np.ascontiguousarray(sel_words, np.uint64)
    lib().orc_hll_registers(ptr(col), _FROM_NP[col.dtype], len(col),
                            ptr(sel) if sel is not None else None, ptr(regs))
    return regs


def hll_estimate(regs):
    regs = np.ascontiguousarray(regs, np.uint8)
    return int(lib().orc_hll_estimate(ptr(regs)))


def count_distinct(col, sel_words=None, mode=0):
    col = np.ascontiguousarray(col)
    sel = None if sel_words is None else np.ascontiguousarray(sel_words, np.uint64)
    return int(lib().orc_count_distinct(ptr(col), _FROM_NP[col.dtype], len(col),
                                        ptr(sel) if sel is not None else None, mode))


def cells(col):
    """u64 cells of a column as the reference stores them (signed types sign-extended)."""
    col = np.asarray(col)
    return col.astype(np.int64).view(np.uint64) if col.dtype.kind == "i" else col.astype(np.uint64)


def select_rows(row_begin, row_end, sel_words=None, order_col=None, descending=False,
                limit=None):
    """dispatch.rs SelectStar/SelectColumn/SelectMulti (:486-733): selected rows of
    [row_begin, row_end) in scan order, stably sorted by the order column's u64 cells
    (`va.cmp(&vb)`, DESC `vb.cmp(&va)`), first `limit`."""
    n = row_end - row_begin
    if sel_words is None:
        rows = np.arange(row_begin, row_end, dtype=np.uint64)
    else:
        bits = np.unpackbits(np.ascontiguousarray(sel_words, np.uint64).view(np.uint8),
                             bitorder="little")[:n].astype(bool)
        rows = (np.flatnonzero(bits) + row_begin).astype(np.uint64)
    if order_col is not None and len(rows):
        keys = cells(order_col)[rows.astype(np.int64)]
        perm = np.argsort(~keys if descending else keys, kind="stable")
        rows = rows[perm]
    return rows if limit is None else rows[:limit]


def sum_f64(bits, sel_words=None, neg_zero_start=True):
    """SUM over f64 cells (their u64 bit patterns) under an optional bitmap, exact
    and rounded once: the sum of the finite values as a Fraction (each double is a
    dyadic rational, converted exactly), rounded to the nearest double by Python's
    int / int division (correctly rounded, ties to even; OverflowError -> +-inf).
    NaN, or +inf with -inf, give NaN; else an infinity gives itself.  A zero total
    is -0.0 only when neg_zero_start and every summed cell is -0.0 (or there are
    none): Iterator::sum over f64 starts from -0.0 (kernel/aggregate.rs:41); the
    dispatch sums start from +0.0 (exec/vectorized.rs:129, :526)."""
    import math
    from fractions import Fraction
    vals = np.ascontiguousarray(bits, np.uint64).view(np.float64)
    if sel_words is not None:
        w = np.ascontiguousarray(sel_words, np.uint64)
        idx = np.arange(len(vals))
        keep = (w[idx >> 6] >> (idx & 63).astype(np.uint64)) & np.uint64(1)
        vals = vals[keep.astype(bool)]
    nan = bool(np.isnan(vals).any())
    pinf, ninf = bool((vals == np.inf).any()), bool((vals == -np.inf).any())
    if nan or (pinf and ninf):
        return math.nan
    if pinf or ninf:
        return math.inf if pinf else -math.inf
    all_neg_zero = bool(np.all((vals == 0) & np.signbit(vals)))
    total = sum((Fraction(float(x)) for x in vals if x != 0), Fraction(0))
    if total == 0:
        return -0.0 if (neg_zero_start and all_neg_zero) else 0.0
    try:
        return total.numerator / total.denominator
    except OverflowError:
        return math.inf if total > 0 else -math.inf


def filter_f64(values, op, eps, lit):
    """exec/bitmap.rs filter_*_f64 over f64 values -> LSB-first bitmap bytes."""
    cells = np.ascontiguousarray(values, np.float64).view(np.uint64)
    bits = np.zeros((len(cells) + 7) // 8, np.uint8)
    lib().orc_filter_f64(ptr(cells), len(cells), op, int(eps), float(lit), ptr(bits))
    return bits


def scan_expr(arrays, prog, row_begin=0, row_end=None, bits=None):
    """The interpreter's per-row eval of a postfix program (code, arg, imm) over host
    column arrays: the selection bits (all rows if bits is None) of [row_begin, row_end)
    narrowed to the rows whose value is truthy."""
    keep = [np.ascontiguousarray(a) for a in arrays]
    n = len(keep[0]) if keep else 0
    row_end = n if row_end is None else row_end
    cols = (OColumn * max(1, len(keep)))()
    for i, a in enumerate(keep):
        cols[i] = OColumn(ptr(a), len(a), _FROM_NP[a.dtype], 0)
    ops = (OExprOp * len(prog))(*[OExprOp(c, a, im) for c, a, im in prog])
    nb = row_end - row_begin
    if bits is None:
        bits = np.zeros((nb + 7) // 8, np.uint8)
        bits[: nb // 8] = 0xFF
        if nb % 8:
            bits[nb // 8] = (1 << (nb % 8)) - 1
    bits = np.ascontiguousarray(bits, np.uint8).copy()
    rc = lib().orc_scan_expr(cols, len(keep), ops, len(prog), row_begin, row_end, ptr(bits))
    if rc != 0:
        raise ValueError("malformed expression program")
    return bits


NO_BUILD_ROW = 0xFFFFFFFF


def join_residual(left_arrays, right_arrays, prog, probe, build, left=False):
    """Filters equi-join pairs (probe row, build row; NO_BUILD_ROW: a LEFT row with no
    match) by a residual program whose columns 0..len(left_arrays)-1 read the probe
    table at the probe row and the rest the build table at the build row; LEFT keeps
    a probe row whose matches all fail as (row, NO_BUILD_ROW)."""
    keep = [np.ascontiguousarray(a) for a in list(left_arrays) + list(right_arrays)]
    cols = (OColumn * max(1, len(keep)))()
    for i, a in enumerate(keep):
        cols[i] = OColumn(ptr(a), len(a), _FROM_NP[a.dtype], 0)
    ops = (OExprOp * len(prog))(*[OExprOp(c, a, im) for c, a, im in prog])
    probe = np.ascontiguousarray(probe, np.uint64)
    build = np.ascontiguousarray(build, np.uint32)
    n = len(probe)
    op, ob = np.zeros(max(1, n), np.uint64), np.zeros(max(1, n), np.uint32)
    m = lib().orc_join_residual(cols, len(keep), len(left_arrays), ops, len(prog), int(left),
                                ptr(probe), ptr(build), n, NO_BUILD_ROW, ptr(op), ptr(ob))
    if m < 0:
        raise ValueError("malformed expression program")
    return op[:m].copy(), ob[:m].copy()
