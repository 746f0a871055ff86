"""Kernel dispatch of SELECT shapes onto the device operators.

Mirrors engine/dispatch.rs: `QueryShape` (:196-222), `classify_query`
(:225-299) and `execute_dispatched` / `execute_shape` (:304-735), for queries
whose WHERE clause is an AND of column-vs-literal comparisons over resident
device columns.  The SQL parser is not part of this path: a query arrives as
the `SelectQuery` structure below, as the reference's parser would produce it.

Results follow `QueryResult` / `ResultColumn` (engine/result.rs:37-131) with
exact values instead of the dispatch path's encodings, as a caller of the
device path needs them:
  - SUM / AVG are exact: an integer when it fits INT8 (type_oid 20), else NUMERIC
    text (type_oid 1700) in `string_values`; AVG is NUMERIC text rounded half
    away from zero (the dispatch path returns `(u64 sum as f64).to_bits()` and a
    u64 truncating average, vectorized.rs:514-522, :567-582);
  - MIN / MAX compare u64 cells, as `min_masked` / `max_masked` do
    (vectorized.rs:541-565);
  - COUNT(DISTINCT) is exact, as `count_distinct_masked` (vectorized.rs:601-610);
  - GROUP BY emits the groups ordered by key (the reference emits them in
    first-seen order; rows compare equal after sorting by key);
  - an aggregate argument that names no column is an error, not column 0
    (dispatch.rs:1233 resolves it with `unwrap_or(0)`).
NULLs (a device bitmap per nullable column, bit set = NULL, as loaded by
tgp_null_bitmap_from_flags): COUNT(*) counts every selected row; COUNT(col),
SUM, AVG, MIN, MAX and COUNT(DISTINCT) skip NULL cells of their columns
(adjust_mask_for_nulls, dispatch.rs:1679-1687; tests/null_semantics.rs), also
per group; a comparison with a NULL cell is false in WHERE (SQL; the dispatch
path compares the stored cell); SELECT results carry a null_mask.
"""
from dataclasses import dataclass, field, replace
from decimal import Decimal
from enum import Enum
from typing import List, Optional, Sequence, Tuple

import struct

import numpy as np

from . import _capi as A
from . import expr as X
from .device import DeviceBitmap, DeviceBuffer, HashTable, make_agg_query, numeric_text

OPS = {"=": A.EQ, "!=": A.NE, "<>": A.NE, "<": A.LT, "<=": A.LE, ">": A.GT, ">=": A.GE}
INT8_OID, NUMERIC_OID, FLOAT8_OID = 20, 1700, 701
# a single GROUP BY key whose value range is at most this wide groups by dense ids
# (tgp_group_agg); wider or sparse keys, several keys and nullable keys go through
# the device hash table (tgp_hash_group_agg_sync)
DENSE_KEY_RANGE = 1 << 24


class Unsupported(Exception):
    """A query shape the device dispatcher does not run: execute_dispatched returns
    None for it, as dispatch.rs:304-333 hands such queries back to the interpreter."""


class QueryShape(Enum):
    CountAll = 0
    CountFilter = 1
    SumCol = 2
    MinMax = 3
    CountDistinct = 4
    AvgCol = 5
    GroupByCount = 6
    GroupBySum = 7
    GroupByOrderByLimit = 8
    SelectStar = 9
    SelectColumn = 10
    SelectMulti = 11
    Complex = 12


@dataclass
class Aggregate:
    """SelectItem::Aggregate: func in COUNT, COUNT_DISTINCT, SUM, AVG, MIN, MAX; arg a
    column name, "*" (COUNT), or a product (coef, [(column, fa, fb), ...]) meaning
    coef * prod(fa + fb * column) for SUM / AVG."""
    func: str
    arg: object
    alias: Optional[str] = None


@dataclass
class Column:
    name: str


@dataclass
class Star:
    pass


@dataclass
class SelectQuery:
    select: List[object]
    # an AND of terms: (column, op, literal), (column, op, Column(other)), or an
    # expression tree of turbogp_amd.expr (arithmetic, AND / OR / NOT, BETWEEN, IN,
    # CASE), which the interpreter evaluates per row (expr.rs:1453-1463)
    where: List[object] = field(default_factory=list)
    group_by: List[str] = field(default_factory=list)
    order_by: List[Tuple[str, bool]] = field(default_factory=list)  # (column, ascending)
    limit: Optional[int] = None
    # HAVING <aggregate or result column> <op> <literal> (accum_grouped.rs HavingLit):
    # an Aggregate (matched to the SELECT list, or computed for the filter alone) or
    # the name of a result column
    having: Optional[Tuple[object, str, object]] = None


@dataclass
class ResultColumn:
    name: str
    values: list
    string_values: Optional[List[str]] = None
    type_oid: int = 0
    null_mask: Optional[List[bool]] = None


@dataclass
class QueryResult:
    columns: List[ResultColumn]
    row_count: int

    def scalar(self):
        c = self.columns[0]
        return c.string_values[0] if c.string_values else c.values[0]


class _NullsByColumn(dict):
    """NULL bitmaps keyed by column name, looked up through the table's name
    resolution (so that `t.id` and `id` find the same column)."""

    def __init__(self, table, d):
        super().__init__(d)
        self._t = table

    def __contains__(self, name):
        try:
            return dict.__contains__(self, self._t.resolve(name))
        except KeyError:
            return False

    def get(self, name, default=None):
        try:
            return dict.get(self, self._t.resolve(name), default)
        except KeyError:
            return default


class Table:
    """Resident device columns by name (all of one length), and the NULL bitmaps
    (DeviceBitmap, bit set = NULL) of the nullable ones."""

    def __init__(self, ctx, columns: dict, nulls: Optional[dict] = None,
                 types: Optional[dict] = None):
        """types: column name -> 'int' (default), 'float' (f64 bits in 8-byte cells,
        ColType::Float) or 'date' (ColType::Date)."""
        self.ctx = ctx
        self.columns = dict(columns)
        self.types = dict(types or {})
        lens = {c.n for c in self.columns.values()}
        assert len(lens) <= 1, "columns of different lengths"
        self.row_count = lens.pop() if lens else 0
        self.names = list(self.columns)
        self.nulls = _NullsByColumn(self, {self.resolve(k): v for k, v in (nulls or {}).items()})

    def resolve(self, name):
        """resolve_col_name (dispatch.rs:1034-1061): the name itself, the query name
        without its table prefix, or a qualified column whose bare name it is."""
        if name in self.columns:
            return name
        parts = name.split(".")
        if len(parts) > 1 and parts[1] in self.columns:
            return parts[1]
        for cn in self.names:
            cparts = cn.split(".")
            if cn == name or (len(cparts) > 1 and cparts[1] == name) or cn.endswith("." + name):
                return cn
        raise KeyError(f"unknown column {name!r}")

    def col(self, name):
        return self.columns[self.resolve(name)]

    def col_type(self, name):
        return self.types.get(self.resolve(name), "int")

    def free(self):
        for c in self.columns.values():
            c.free()
        for b in dict.values(self.nulls):
            b.free()
        self.columns = {}


def classify_query(q: SelectQuery) -> QueryShape:
    """dispatch.rs:225-299."""
    has_where = bool(q.where)
    if q.group_by:
        aggs = [s for s in q.select if isinstance(s, Aggregate)]
        if not aggs:
            return QueryShape.Complex
        func = aggs[0].func.upper()
        if func == "COUNT":
            return (QueryShape.GroupByOrderByLimit if q.order_by and q.limit is not None
                    else QueryShape.GroupByCount)
        if func == "SUM":
            return QueryShape.GroupBySum
        return QueryShape.Complex
    if len(q.select) == 1:
        s = q.select[0]
        if isinstance(s, Aggregate):
            f = s.func.upper()
            if f == "COUNT":
                return QueryShape.CountAll if s.arg == "*" and not has_where else \
                    QueryShape.CountFilter
            return {"COUNT_DISTINCT": QueryShape.CountDistinct, "SUM": QueryShape.SumCol,
                    "AVG": QueryShape.AvgCol, "MIN": QueryShape.MinMax,
                    "MAX": QueryShape.MinMax}.get(f, QueryShape.Complex)
        if isinstance(s, Star):
            return QueryShape.SelectStar
        if isinstance(s, Column):
            return QueryShape.SelectColumn
        return QueryShape.Complex
    if len(q.select) > 1:
        if any(isinstance(s, Aggregate) for s in q.select):
            return QueryShape.Complex
        return QueryShape.SelectMulti
    return QueryShape.Complex


def execute_dispatched(q: SelectQuery, t: Table) -> Optional[QueryResult]:
    """dispatch.rs:304-333: None for shapes the dispatcher does not run, for any
    query with HAVING (dispatch.rs:329-331 sends those to the interpreter), and for
    a recognised shape whose details the device path does not take (Unsupported:
    e.g. COUNT(DISTINCT) per group, a SELECT column that is neither grouped nor
    aggregated, more than four GROUP BY columns).  Malformed input (an unknown
    column, a bad operator) still raises."""
    shape = classify_query(q)
    if shape == QueryShape.Complex or q.having is not None:
        return None
    try:
        return _Executor(q, t).run(shape)
    except Unsupported:
        return None


def execute_grouped(q: SelectQuery, t: Table) -> Optional[QueryResult]:
    """The interpreter's accumulator path for a GROUP BY of one column
    (accum_grouped.rs:55-409 try_accumulator_grouped): the groups and their exact
    aggregates on the device, then HAVING <aggregate> <op> <literal> (HavingLit,
    :148-160, :339-349), then ORDER BY and LIMIT, in that order.  None when the
    device path does not take the query (as try_accumulator_grouped returns None)."""
    if not q.group_by:
        raise ValueError("execute_grouped needs a GROUP BY")
    hidden = False
    if q.having is not None and isinstance(q.having[0], Aggregate):
        # HAVING on an aggregate: its result column, computed for the filter only
        # when the SELECT list does not hold it
        h = q.having[0]
        same = [s for s in q.select if isinstance(s, Aggregate) and
                s.func.upper() == h.func.upper() and s.arg == h.arg]
        if same:
            name = same[0].alias or same[0].func.lower()
        else:
            name, hidden = "__having", True
            q = replace(q, select=list(q.select) + [Aggregate(h.func, h.arg, name)])
        q = replace(q, having=(name, q.having[1], q.having[2]))
    ex = _Executor(q, t)
    try:
        bm, _ = ex.mask()
        r = ex.group_by(bm)
    except Unsupported:
        return None
    finally:
        ex.free()
    if hidden:
        r = QueryResult([c for c in r.columns if c.name != "__having"], r.row_count)
    return r


def _signed_value(c, i):
    """A result cell as a number: NUMERIC text as Decimal, an INT8 cell as a signed
    64-bit integer."""
    if c.string_values:
        return Decimal(c.string_values[i])
    v = int(c.values[i])
    return v - (1 << 64) if v >= (1 << 63) else v


_HAVING_OPS = {"=": lambda a, b: a == b, "!=": lambda a, b: a != b, "<>": lambda a, b: a != b,
               "<": lambda a, b: a < b, "<=": lambda a, b: a <= b, ">": lambda a, b: a > b,
               ">=": lambda a, b: a >= b}


@dataclass
class Join:
    """One JOIN clause: the joined table, the ON equalities as (left column, right
    column) pairs (what extract_join_keys yields, exec/join.rs), and the type:
    INNER, LEFT, RIGHT, FULL or CROSS (executor.rs:1766-1788).  `residual`: the ON
    clause's other conjuncts as one expression (turbogp_amd.expr, AND of them), applied
    to every match as the interpreter's hash_join does (query_interpreter/join.rs:
    1146-1195, eval_non_equi_match :1233-1300; INNER and LEFT only, its JoinType2).
    split_on() divides an ON expression into the two as the interpreter does."""
    table: str
    on: List[Tuple[str, str]] = field(default_factory=list)
    join_type: str = "INNER"
    residual: object = None


def execute_with_join(q: SelectQuery, from_table: str, joins: Sequence[Join],
                      catalog: dict) -> Optional[QueryResult]:
    """engine/executor.rs:1747-1811 execute_with_join: every JOIN is computed on the
    device (hash_join: build on the right table, probe with the running left one; CROSS:
    every left row with every right row, cross_join_into :1886-1948), the joined rows
    are materialised as a device table whose columns are named `table.column`, and the
    query without its JOINs is dispatched on it.  None when the dispatcher does not
    take the query (the reference then runs its row executor).  As JoinResult::into_table,
    the joined table has no NULL bitmaps: the missing side of an outer join's row reads
    0 (exec/join.rs:118-124, :141-152).  A self-join's equal qualified names collide:
    the right table's columns win."""
    base = catalog[from_table]
    running = base
    try:
        for j in joins:
            joined = _join_tables(running, catalog[j.table], from_table, j)
            if running is not base:
                running.free()
            running = joined
        return execute_dispatched(q, running)
    except Unsupported:
        return None
    finally:
        if running is not base:
            running.free()


_KIND = {"int": X.INT, "float": X.FLOAT, "date": X.DATE}


def split_on(on, left: Table, right: Table, right_name: str):
    """An ON expression (turbogp_amd.expr) as (key pairs, residual): the interpreter's
    collect_keys (join.rs:1313-1334: every `a = b` reachable through ANDs with one
    column in each table, as (left column, right column)) and its non-equi conjuncts
    (:1105-1109: the AND-conjuncts that are not `column = column`, ANDed back, or
    None).  As there, `x = y` of two columns of one table is neither."""
    def conjuncts(e):
        if isinstance(e, X.BinOp) and e.op.upper() == "AND":
            return conjuncts(e.left) + conjuncts(e.right)
        return [e]

    def col_in(e, t, tname):
        if not isinstance(e, X.Col):
            return None
        parts = e.name.split(".")
        if t is right and len(parts) == 2:
            return parts[1] if parts[0] == tname and parts[1] in t.columns else None
        try:
            return t.resolve(e.name)
        except KeyError:
            return None

    keys, rest = [], []
    for c in conjuncts(on):
        if isinstance(c, X.BinOp) and c.op == "=" and isinstance(c.left, X.Col) \
                and isinstance(c.right, X.Col):
            lk, rk = col_in(c.left, left, None), col_in(c.right, right, right_name)
            if lk is not None and rk is not None:
                keys.append((lk, rk))
            else:
                rk, lk = col_in(c.left, right, right_name), col_in(c.right, left, None)
                if lk is not None and rk is not None:
                    keys.append((lk, rk))
            continue
        rest.append(c)
    residual = None
    for c in rest:
        residual = c if residual is None else X.BinOp("AND", residual, c)
    return keys, residual


def _residual_pairs(left: Table, right: Table, left_name: str, j: Join, p, b):
    """The pairs (p, b) of left JOIN right kept by j.residual.  Column names resolve as
    the interpreter's combined column map (join.rs:1127-1140): `table.column` of the
    joined table is a right column, other names the left table's, an unqualified
    name present in both is the right one."""
    names = X.columns(j.residual)
    lrefs, rrefs = [], []
    for nm in names:
        parts = nm.split(".")
        side = None
        if len(parts) == 2 and parts[0] == j.table and parts[1] in right.columns:
            side = ("R", parts[1])
        elif len(parts) == 1 and nm in right.columns:
            side = ("R", nm)
        else:
            try:
                side = ("L", left.resolve(nm if len(parts) == 1 or parts[0] != left_name
                                          else parts[1]))
            except KeyError:
                raise Unsupported(f"residual column {nm!r}")
        (rrefs if side[0] == "R" else lrefs).append((nm, side[1]))
    lidx = {nm: i for i, (nm, _) in enumerate(lrefs)}
    ridx = {nm: len(lrefs) + i for i, (nm, _) in enumerate(rrefs)}

    def ref(nm):
        if nm in lidx:
            return lidx[nm], _KIND[left.col_type(dict(lrefs)[nm])]
        return ridx[nm], _KIND[right.col_type(dict(rrefs)[nm])]
    prog = X.compile_expr(j.residual, ref)
    if len(lrefs) + len(rrefs) > A.MAX_COLS or len(prog) > A.MAX_EXPR_OPS:
        raise Unsupported("residual too large for the device program")
    return left.ctx.join_residual([left.columns[c] for _, c in lrefs],
                                  [right.columns[c] for _, c in rrefs], prog, p, b,
                                  left=j.join_type.upper() == "LEFT")


def _join_tables(left: Table, right: Table, left_name: str, j: Join) -> Table:
    """The rows of left JOIN right as a new device table (hash_join, exec/join.rs:68-160)."""
    ctx = left.ctx
    jt = j.join_type.upper()
    u64 = np.uint64
    if jt == "CROSS":
        li = np.repeat(np.arange(left.row_count, dtype=u64), right.row_count)
        ri = np.tile(np.arange(right.row_count, dtype=u64), left.row_count)
    else:
        if not j.on:
            raise ValueError("JOIN needs at least one ON equality (hash_join requires a key)")
        if j.residual is not None and jt not in ("INNER", "LEFT"):
            raise Unsupported(f"{jt} JOIN with non-equi conditions")
        lk = [left.col(a) for a, _ in j.on]
        rk = [right.col(b) for _, b in j.on]
        t = HashTable(ctx, rk if len(rk) > 1 else rk[0])
        try:
            outer_left = jt in ("LEFT", "FULL")
            p, b = t.join_pairs(lk if len(lk) > 1 else lk[0],
                                A.JOIN_LEFT if outer_left else A.JOIN_INNER)
        finally:
            t.free()
        if j.residual is not None:
            p, b = _residual_pairs(left, right, left_name, j, p, b)
        hit = b != A.JOIN_NO_MATCH
        li = p.astype(u64)
        ri = np.where(hit, b.astype(u64), u64(A.GATHER_NO_ROW))
        if jt in ("RIGHT", "FULL"):  # the right rows no left row matched, after the pairs
            matched = np.zeros(right.row_count, bool)
            matched[b[hit].astype(np.int64)] = True
            extra = np.flatnonzero(~matched).astype(u64)
            li = np.concatenate([li, np.full(len(extra), A.GATHER_NO_ROW, u64)])
            ri = np.concatenate([ri, extra])
    n = len(li)
    cols, types = {}, {}
    bufs = [DeviceBuffer(ctx, max(8, 8 * n)) for _ in range(2)]
    try:
        if n:
            bufs[0].upload(li)
            bufs[1].upload(ri)
        try:
            for name in left.names:
                qn = name if "." in name else f"{left_name}.{name}"
                cols[qn] = ctx.gather_column(left.columns[name], bufs[0], n)
                types[qn] = left.col_type(name)
            for name in right.names:
                qn = f"{j.table}.{name}"
                if qn in cols:
                    cols[qn].free()
                cols[qn] = ctx.gather_column(right.columns[name], bufs[1], n)
                types[qn] = right.col_type(name)
        except Exception:
            for c in cols.values():
                c.free()
            raise
    finally:
        for bf in bufs:
            bf.free()
    return Table(ctx, cols, types=types)


def _exact_column(name, vals):
    """INT8 when every value fits, else NUMERIC text."""
    if all(-2**63 <= v < 2**63 for v in vals):
        return ResultColumn(name, [v % 2**64 for v in vals], None, INT8_OID)
    return ResultColumn(name, [0] * len(vals), [str(v) for v in vals], NUMERIC_OID)


def _avg_text(count, total):
    """AVG as NUMERIC text of an exact sum over a count (tgp_int128_avg_to_numeric)."""
    from .device import GroupAggResult
    w = np.array([[[total % 2**64, (total >> 64) % 2**64]]], np.uint64)
    return GroupAggResult(np.array([count], np.uint64), w, 0).avg_numeric(0)[0]


def _arg_columns(arg):
    if isinstance(arg, str):
        return [] if arg == "*" else [arg]
    return [c for c, _, _ in arg[1]]


class _Executor:
    def __init__(self, q, t):
        self.q, self.t, self.ctx = q, t, t.ctx
        self.owned = []

    def free(self):
        for b in self.owned:
            b.free()
        self.owned = []

    # ---- WHERE: AND of terms -> selection bitmap over all rows ----------------------
    def mask(self):
        if not self.q.where:
            return None, self.t.row_count
        signed = (A.I8, A.I16, A.I32, A.I64)
        exprs = [w for w in self.q.where if not isinstance(w, tuple)]
        terms = [w for w in self.q.where if isinstance(w, tuple)]
        for _, op, _ in terms:
            if op not in OPS:  # e.g. LIKE (try_string_like_filter): not a device filter
                raise Unsupported(f"comparison {op!r}")
        # FLOAT columns against a literal: the f64 filters, `=` with the epsilon of
        # apply_comparison (expr.rs:1509-1517: filter_eq_f64_epsilon, filter_ne_f64)
        floats = [w for w in terms if not isinstance(w[2], Column) and self.t.col_type(w[0]) == "float"]
        lits = [w for w in terms if not isinstance(w[2], Column) and w not in floats]
        pairs = [w for w in terms if isinstance(w[2], Column)]
        names = sorted({c for c, _, _ in lits})
        cols = [self.t.col(c) for c in names]
        preds = [(names.index(c), OPS[op], int(lit), self.t.col(c).dtype in signed)
                 for c, op, lit in lits]
        bm, n = self.ctx.scan_filter(cols, preds, 0, self.t.row_count)
        self.owned.append(bm)
        narrowed = False
        for c, op, lit in floats:
            fb, _ = self.ctx.scan_filter_f64(self.t.col(c), OPS[op], float(lit), OPS[op] == A.EQ,
                                             0, self.t.row_count)
            both = bm.combine(fb, A.BIT_AND)
            fb.free()
            self.owned.append(both)
            bm, narrowed = both, True
        # column-vs-column comparisons narrow the selection (expr.rs:1382-1451:
        # signed for signed columns, as INT / DATE compare)
        for c, op, other in pairs:
            a, b = self.t.col(c), self.t.col(other.name)
            self.ctx.compare_columns(a, OPS[op], b, a.dtype in signed, 0, self.t.row_count, bm)
            narrowed = True
        # expression terms: the per-row eval, on the device (tgp_scan_expr)
        kinds = {"int": X.INT, "float": X.FLOAT, "date": X.DATE}
        used_expr = []
        for e in exprs:
            names_e = [self.t.resolve(c) for c in X.columns(e)]
            used_expr += names_e
            prog = X.compile_expr(e, lambda nm: (names_e.index(self.t.resolve(nm)),
                                                 kinds[self.t.col_type(nm)]))
            if len(names_e) > A.MAX_COLS or len(prog) > A.MAX_EXPR_OPS:
                raise Unsupported("expression too large for the device evaluator")
            self.ctx.scan_expr([self.t.columns[c] for c in names_e], prog, 0, self.t.row_count, bm)
            narrowed = True
        if narrowed:
            n = bm.count_ones()
        # a comparison with a NULL cell is not true; an expression term drops the rows
        # where any column it reads is NULL (as the AND of comparisons does)
        used = (names + [c for c, _, _ in floats] + [c for c, _, _ in pairs] +
                [o.name for _, _, o in pairs] + used_expr)
        bm, n = self.without_nulls(bm, n, used)
        return bm, n

    def without_nulls(self, bm, n, columns):
        """The selection `bm` (None = all rows) minus the rows where any of `columns`
        is NULL -> (bitmap or None, selected count)."""
        for c in dict.fromkeys(columns):
            nb = self.t.nulls.get(c)
            if nb is None:
                continue
            if bm is None:
                bm = nb.not_()
            else:
                bm = bm.combine(nb, A.BIT_ANDNOT)
            self.owned.append(bm)
        if bm is not None and self.t.nulls and any(c in self.t.nulls for c in columns):
            n = bm.count_ones()
        return bm, n

    # ---- aggregates without GROUP BY -------------------------------------------------
    def agg_query(self, items, parts=(), bitmap=None, names_extra=()):
        return self.ctx.group_agg(self.build_query(items, parts, bitmap, names_extra))

    def build_query(self, items, parts=(), bitmap=None, names_extra=()):
        """AggQuery of the aggregates `items` over all rows; the columns named in
        names_extra take the first column slots, in that order."""
        names = []

        def slot(c):
            if c not in names:
                names.append(c)
            return names.index(c)

        for c in names_extra:
            slot(c)
        aggs = []
        for func, arg in items:
            if func in ("SUM", "AVG") and any(self.t.col_type(c) == "float"
                                             for c in _arg_columns(arg)):
                # the reference sums FLOAT cells as doubles; the exact integer
                # engine would add their bit patterns
                raise Unsupported(f"{func} over a FLOAT column in this shape")
            if func in ("SUM", "AVG"):
                if isinstance(arg, str):
                    aggs.append(("sum", 1, [(slot(arg), 0, 1)]))
                else:
                    coef, factors = arg
                    aggs.append(("sum", coef, [(slot(c), fa, fb) for c, fa, fb in factors]))
            elif func in ("MIN", "MAX"):
                aggs.append((func.lower() + "_u64", slot(arg)))
            else:
                raise Unsupported(f"aggregate {func}")
        parts = [(slot(c), div, base, card) for c, div, base, card in parts]
        if len(names) > A.MAX_COLS or len(aggs) > A.MAX_AGGS:
            raise Unsupported(f"{len(names)} columns / {len(aggs)} aggregates in one pass")
        cols = [self.t.col(c) for c in names] if names else [next(iter(self.t.columns.values()))]
        return make_agg_query(cols, (), parts, aggs, 0, self.t.row_count, bitmap)

    def run(self, shape):
        q, t = self.q, self.t
        if shape == QueryShape.CountAll:
            return QueryResult([ResultColumn("count", [t.row_count], None, INT8_OID)], 1)
        try:
            bm, nsel = self.mask()
            return self._run(shape, bm, nsel)
        finally:
            self.free()

    def _run(self, shape, bm, nsel):
        q, t, ctx = self.q, self.t, self.ctx
        s0 = q.select[0]
        if shape == QueryShape.CountFilter:
            cols = _arg_columns(s0.arg)
            for c in cols:
                t.col(c)
            _, n = self.without_nulls(bm, nsel, cols)  # COUNT(col) skips NULLs
            return QueryResult([ResultColumn("count", [n], None, INT8_OID)], 1)
        if shape in (QueryShape.SumCol, QueryShape.AvgCol, QueryShape.MinMax):
            func = s0.func.upper()
            for c in _arg_columns(s0.arg):
                t.col(c)
            bm, nsel = self.without_nulls(bm, nsel, _arg_columns(s0.arg))
            if func in ("SUM", "AVG") and isinstance(s0.arg, str) and t.col_type(s0.arg) == "float":
                # sum_masked_f64 / avg_masked_f64 (dispatch.rs:399-403, :428-432): a FLOAT
                # column sums as doubles from +0.0; AVG of no rows reads 0
                total = ctx.sum_f64(t.col(s0.arg), t.row_count, bm, neg_zero_start=False)
                v = total if func == "SUM" else (total / nsel if nsel else 0.0)
                bits = struct.unpack("<Q", struct.pack("<d", v))[0]
                return QueryResult([ResultColumn(s0.alias or func.lower(), [bits], [repr(v)],
                                                 FLOAT8_OID)], 1)
            r = self.agg_query([(func, s0.arg)], bitmap=bm)
            name = s0.alias or func.lower()
            if func == "AVG":
                v = r.avg_numeric(0) if nsel else [None]
                return QueryResult([ResultColumn(name, [0], [v[0] or "0"], NUMERIC_OID)], 1)
            if func == "SUM":
                return QueryResult([_exact_column(name, [r.sums()[0][0]])], 1)
            val = r.sums()[0][0] if nsel else (0)  # MIN/MAX of no rows: 0, as min_masked
            return QueryResult([ResultColumn(name, [val % 2**64], None, INT8_OID)], 1)
        if shape == QueryShape.CountDistinct:
            bm, _ = self.without_nulls(bm, nsel, [s0.arg])
            n = ctx.count_distinct(t.col(s0.arg), 0, t.row_count, bm, A.DISTINCT_EXACT)
            return QueryResult([ResultColumn(s0.alias or "count", [n], None, INT8_OID)], 1)
        if shape in (QueryShape.GroupByCount, QueryShape.GroupBySum,
                     QueryShape.GroupByOrderByLimit):
            return self.group_by(bm)
        return self.select(shape, bm)

    # ---- GROUP BY: dense ids for one narrow key, else the device hash table ------------
    def group_by(self, bm):
        q, t = self.q, self.t
        keys = [t.resolve(k) for k in q.group_by]
        if not 1 <= len(keys) <= A.MAX_GROUP_KEYS or len(set(keys)) != len(keys):
            raise Unsupported(f"GROUP BY of {len(keys)} columns")
        for s in q.select:
            if isinstance(s, Column) and t.resolve(s.name) not in keys:
                # the reference emits the first row's cell of such a column
                raise Unsupported(f"column {s.name!r} is neither grouped nor aggregated")
            if isinstance(s, Aggregate):
                if s.func.upper() == "COUNT_DISTINCT":
                    raise Unsupported("COUNT(DISTINCT) per group")
                for c in _arg_columns(s.arg):
                    t.col(c)
        signed = [t.col(k).dtype in (A.I8, A.I16, A.I32, A.I64) for k in keys]
        dense = None
        if len(keys) == 1 and keys[0] not in t.nulls and t.row_count:
            lo, hi = t.col(keys[0]).minmax()  # int64 readings of the cells
            narrow = (hi - lo + 1 <= DENSE_KEY_RANGE) if signed[0] else \
                (0 <= lo <= hi < DENSE_KEY_RANGE)
            if narrow:
                dense = (lo, hi - lo + 1)

        def groups(items, bitmap):
            """{key tuple: (count, [exact values per item])} of the rows in `bitmap`."""
            if dense is not None:
                lo, card = dense
                r = self.agg_query(items, parts=[(keys[0], 1, lo, card)], bitmap=bitmap,
                                   names_extra=keys)
                sums = r.sums()
                return {(lo + g,): (int(r.count[g]), sums[g]) for g in range(card) if r.count[g]}
            qq = self.build_query(items, bitmap=bitmap, names_extra=keys)
            hg = self.ctx.hash_group_agg(qq, list(range(len(keys))),
                                         [t.nulls.get(k) for k in keys])
            return hg.as_dict(signed)

        base = groups([], bm)
        tuples = sorted(base, key=lambda tp: [(1, 0) if v is None else (0, v) for v in tp])
        # one pass per set of nullable columns the aggregates read: each aggregate
        # sees the selected rows where its own columns are not NULL
        passes = {}
        for s in q.select:
            if isinstance(s, Aggregate) and not (s.func.upper() == "COUNT" and s.arg == "*"):
                ns = tuple(sorted(c for c in _arg_columns(s.arg) if c in t.nulls))
                passes.setdefault(ns, []).append(s)
        results = {}
        for ns, aggs in passes.items():
            items = [(s.func.upper(), s.arg) for s in aggs
                     if s.func.upper() in ("SUM", "AVG", "MIN", "MAX")]
            pbm, _ = self.without_nulls(bm, 0, list(ns)) if ns else (bm, 0)
            results[ns] = (groups(items, pbm), items)
        cols = []
        for s in q.select:
            if isinstance(s, Column):
                k = keys.index(t.resolve(s.name))
                vals = [tp[k] for tp in tuples]
                col = _exact_column(s.name, [0 if v is None else v for v in vals])
                if any(v is None for v in vals):
                    col.null_mask = [v is None for v in vals]
                cols.append(col)
                continue
            f = s.func.upper()
            name = s.alias or f.lower()
            if f == "COUNT" and s.arg == "*":
                cols.append(ResultColumn(name, [base[tp][0] for tp in tuples], None, INT8_OID))
                continue
            ns = tuple(sorted(c for c in _arg_columns(s.arg) if c in t.nulls))
            got, items = results[ns]
            if f == "COUNT":
                cols.append(ResultColumn(name, [got.get(tp, (0, []))[0] for tp in tuples], None,
                                         INT8_OID))
                continue
            k = items.index((f, s.arg))
            if f == "SUM":
                cols.append(_exact_column(name, [got[tp][1][k] if tp in got else 0
                                                 for tp in tuples]))
            elif f == "AVG":
                cols.append(ResultColumn(name, [0] * len(tuples),
                                         [_avg_text(*got[tp][:1], got[tp][1][k]) if tp in got
                                          else "0" for tp in tuples], NUMERIC_OID))
            else:
                cols.append(ResultColumn(name, [(got[tp][1][k] if tp in got else 0) % 2**64
                                                for tp in tuples], None, INT8_OID))
        rows = list(range(len(tuples)))
        names = [c.name for c in cols]
        if q.having is not None:
            hname, hop, hlit = q.having
            if hname not in names:
                raise ValueError(f"HAVING {hname!r} is not a result column")
            if hop not in _HAVING_OPS:
                raise ValueError(f"unsupported HAVING comparison {hop!r}")
            hc = cols[names.index(hname)]
            lit = Decimal(str(hlit))
            rows = [i for i in rows if _HAVING_OPS[hop](_signed_value(hc, i), lit)]
        if q.order_by:
            oname, asc = q.order_by[0]
            if oname not in names:
                raise ValueError(f"ORDER BY {oname!r} is not a result column")
            c = cols[names.index(oname)]
            rows = sorted(rows, key=lambda i: _signed_value(c, i), reverse=not asc)
        if q.limit is not None:
            rows = rows[: q.limit]
        for c in cols:
            c.values = [c.values[i] for i in rows]
            if c.string_values:
                c.string_values = [c.string_values[i] for i in rows]
            if c.null_mask:
                c.null_mask = [c.null_mask[i] for i in rows]
        return QueryResult(cols, len(rows))

    # ---- SELECT rows ---------------------------------------------------------------
    def select(self, shape, bm):
        q, t, ctx = self.q, self.t, self.ctx
        if shape == QueryShape.SelectStar:
            names, order = list(t.names), None  # SelectStar does not read ORDER BY
        else:
            names = [s.name for s in q.select]
            order = q.order_by[0] if q.order_by else None
        for n in names:
            t.col(n)
        rows, cnt = ctx.select_rows(0, t.row_count, bm,
                                    t.col(order[0]) if order else None,
                                    descending=bool(order and not order[1]), limit=q.limit)
        try:
            cols = [ResultColumn(n, [int(v) for v in ctx.gather(t.col(n), rows, cnt)], None,
                                 INT8_OID) for n in names]
            if any(n in t.nulls for n in names) and cnt:
                idx = rows.download(np.uint64, cnt).astype(np.int64)
                for c in cols:
                    nb = t.nulls.get(c.name)
                    if nb is not None:
                        w = nb.words()
                        flags = (w[idx >> 6] >> (idx & 63).astype(np.uint64)) & 1
                        if flags.any():
                            c.null_mask = [bool(x) for x in flags]
        finally:
            rows.free()
        return QueryResult(cols, cnt)
