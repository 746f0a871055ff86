"""Synthetic fact table and the queries of the benchmark configurations.

Fact table (one row per line item; all values seeded and counter-based):
  dt     int32  uniform [0, 2526)            days (named dt: DATE is a keyword
                                            in the reference parser, parser.rs:539-549)
  qty    int64  uniform 1..50
  price  int64  uniform 90 000..10 500 000   cents
  disc   int64  uniform 0..10                hundredths
  tax    int64  uniform 0..8                 hundredths
  flaga  uint8  uniform {0, 1, 2}
  flagb  uint8  uniform {0, 1}
Row i of column c is lo + mulhi64(mix64(SEED ^ (c << 40) ^ i), hi - lo + 1).
"""
import numpy as np

from . import _capi as A
from .device import DeviceColumn, make_agg_query

SEED = 0x74626770_5F4D4933

# name -> (column id, dtype, lo, hi)
FACT_SCHEMA = {
    "dt": (0, A.I32, 0, 2525),
    "qty": (1, A.I64, 1, 50),
    "price": (2, A.I64, 90_000, 10_500_000),
    "disc": (3, A.I64, 0, 10),
    "tax": (4, A.I64, 0, 8),
    "flaga": (5, A.U8, 0, 2),
    "flagb": (6, A.U8, 0, 1),
}
FACT_COLUMNS = list(FACT_SCHEMA)
ROW_BYTES = {k: A.DTYPE_WIDTH[v[1]] for k, v in FACT_SCHEMA.items()}

Q1_DATE_CUTOFF = 2435          # dt <= 2435 (~96.4 % selected)
Q6_WINDOW = (731, 1096)        # 731 <= dt < 1096 (365 days)
Q6_DISC = (5, 7)               # disc BETWEEN 5 AND 7
Q6_QTY = 24                    # qty < 24

Q1_COLS = ["dt", "qty", "price", "disc", "tax", "flaga", "flagb"]
Q6_COLS = ["dt", "disc", "qty", "price"]
Q1_BYTES_PER_ROW = sum(ROW_BYTES[c] for c in Q1_COLS)   # 38
Q6_BYTES_PER_ROW = sum(ROW_BYTES[c] for c in Q6_COLS)   # 28


def make_fact_table(ctx, n, columns=None, stream=None):
    """Generates the fact table on the device (tgp_synth_uniform)."""
    import ctypes as C
    out = {}
    for name in columns or FACT_COLUMNS:
        cid, dt, lo, hi = FACT_SCHEMA[name]
        col = DeviceColumn(ctx, n, dt, name)
        d = col.desc()
        A.check(A.lib().tgp_synth_uniform(ctx.handle, C.byref(d), SEED, cid, lo, hi, 0, n, stream),
                "tgp_synth_uniform")
        out[name] = col
    ctx.sync(stream)
    return out


def q1_spec(date_cutoff=Q1_DATE_CUTOFF, group="ab"):
    """Q1-shape: WHERE dt <= cutoff GROUP BY <group>; COUNT(*), SUM(qty), SUM(price),
    SUM(price*(100-disc)), SUM(price*(100-disc)*(100+tax)).
    group: 'ab' (flaga, flagb), 'a', 'b', 'month' (dt/30), '' (none)."""
    c = {n: i for i, n in enumerate(Q1_COLS)}
    preds = [(c["dt"], A.LE, date_cutoff, False)]
    parts = {"ab": [(c["flaga"], 1, 0, 3), (c["flagb"], 1, 0, 2)],
             "a": [(c["flaga"], 1, 0, 3)], "b": [(c["flagb"], 1, 0, 2)],
             "month": [(c["dt"], 30, 0, 85)], "": []}[group]
    aggs = [("sum", 1, [(c["qty"], 0, 1)]),
            ("sum", 1, [(c["price"], 0, 1)]),
            ("sum", 1, [(c["price"], 0, 1), (c["disc"], 100, -1)]),
            ("sum", 1, [(c["price"], 0, 1), (c["disc"], 100, -1), (c["tax"], 100, 1)])]
    return Q1_COLS, preds, parts, aggs


def q6_spec():
    """Q6-shape: WHERE 731 <= dt < 1096 AND disc BETWEEN 5 AND 7 AND qty < 24;
    SUM(price*disc)."""
    c = {n: i for i, n in enumerate(Q6_COLS)}
    preds = [(c["dt"], A.GE, Q6_WINDOW[0], False), (c["dt"], A.LT, Q6_WINDOW[1], False),
             (c["disc"], A.GE, Q6_DISC[0], True), (c["disc"], A.LE, Q6_DISC[1], True),
             (c["qty"], A.LT, Q6_QTY, True)]
    aggs = [("sum", 1, [(c["price"], 0, 1), (c["disc"], 0, 1)])]
    return Q6_COLS, preds, [], aggs


def build_query(table, spec, row_begin=0, row_end=None, sel_bitmap=None):
    names, preds, parts, aggs = spec
    cols = [table[n] for n in names]
    return make_agg_query(cols, preds, parts, aggs, row_begin, row_end, sel_bitmap)


def host_fact_columns(n, gen):
    """Host copy of the fact table from a generator function
    gen(dtype, seed, column_id, lo, hi, row_begin, row_end) -> np.ndarray."""
    return {name: gen(dt, SEED, cid, lo, hi, 0, n) for name, (cid, dt, lo, hi) in FACT_SCHEMA.items()}


NP_DTYPE = {A.I8: np.int8, A.U8: np.uint8, A.I16: np.int16, A.U16: np.uint16,
            A.I32: np.int32, A.U32: np.uint32, A.I64: np.int64, A.U64: np.uint64}


# ---- hash-join workload (build 2M unique keys, probe 60M Zipf(1.1) foreign keys) ----------
JOIN_BUILD_ROWS = 2_000_000
JOIN_PROBE_ROWS = 60_000_000
JOIN_GKEY_CARD = 1_500_000
JOIN_PAY_CARD = 25
ZIPF_S = 1.1
# column ids of the join tables in the counter-based generator
COL_PAY, COL_FK, COL_VAL, COL_GKEY = 10, 11, 12, 13
_KEY_SALT = 0x6A6F696E5F6B6579  # spreads the permuted build keys over int64


def mix64_np(x):
    """SplitMix64 finaliser over a uint64 numpy array."""
    z = x.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def join_build_side(n_build, seed=SEED):
    """Build keys: a seeded permutation of 0..n-1 passed through mix64 (unique,
    spread over int64); payload int32 uniform 0..24 (counter-based)."""
    rng = np.random.Generator(np.random.PCG64(seed & 0xFFFFFFFF))
    perm = rng.permutation(n_build).astype(np.uint64)
    bkey = mix64_np(perm ^ np.uint64(_KEY_SALT)).view(np.int64)
    return bkey, perm


def zipf_cdf(n, s=ZIPF_S):
    w = np.arange(1, n + 1, dtype=np.float64) ** (-s)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    cdf[-1] = 1.0
    return cdf


def zipf_key_table(bkey, seed=SEED):
    """rank r -> build key of a seeded permutation of the build rows."""
    rng = np.random.Generator(np.random.PCG64((seed >> 32) & 0xFFFFFFFF))
    perm2 = rng.permutation(len(bkey))
    return np.ascontiguousarray(bkey[perm2])


def device_lookup_cdf(ctx, n, dtype, column_id, cdf, table, seed=SEED, stream=None):
    """Device column out[i] = table[#(cdf <= unit(u_i))] (tgp_synth_lookup_cdf)."""
    import ctypes as C
    from .device import DeviceBuffer
    dcdf = DeviceBuffer(ctx, len(cdf) * 8)
    dtab = DeviceBuffer(ctx, len(table) * 8)
    try:
        dcdf.upload(np.ascontiguousarray(cdf, np.float64))
        dtab.upload(np.ascontiguousarray(table, np.int64))
        col = DeviceColumn(ctx, n, dtype)
        d = col.desc()
        A.check(A.lib().tgp_synth_lookup_cdf(ctx.handle, C.byref(d), seed, column_id, dcdf.ptr,
                                             len(cdf), dtab.ptr, 0, n, stream),
                "tgp_synth_lookup_cdf")
        ctx.sync(stream)
        return col
    finally:
        dcdf.free()
        dtab.free()


def device_uniform(ctx, n, dtype, column_id, lo, hi, seed=SEED, stream=None):
    import ctypes as C
    col = DeviceColumn(ctx, n, dtype)
    d = col.desc()
    A.check(A.lib().tgp_synth_uniform(ctx.handle, C.byref(d), seed, column_id, lo, hi, 0, n,
                                      stream), "tgp_synth_uniform")
    return col


# ---- batched workload: many Q1-shape queries over one resident table ------------------------
BATCH_TABLE_ROWS = 256_000_000
BATCH_QUERIES = 4096
BATCH_GROUPS = ("a", "b", "ab", "month")


def batch_query_params(n_queries, table_rows, seed=SEED, lmin=1_000, lmax=4_000_000):
    """Seeded query parameters: row range length log-uniform in [lmin, lmax],
    start uniform, date cutoff uniform in [0, 2526), group key uniform over
    {flaga, flagb, flaga x flagb, dt/30}."""
    rng = np.random.Generator(np.random.PCG64((seed ^ 0xBA7C4) & 0xFFFFFFFF))
    lmax = min(lmax, table_rows)
    lmin = min(lmin, lmax)
    L = np.floor(np.exp(rng.uniform(np.log(lmin), np.log(lmax), n_queries))).astype(np.int64)
    start = np.floor(rng.uniform(0, 1, n_queries) * (table_rows - L + 1)).astype(np.int64)
    cutoff = rng.integers(0, 2526, n_queries)
    group = rng.integers(0, len(BATCH_GROUPS), n_queries)
    return [(int(s), int(s + l), int(c), BATCH_GROUPS[g])
            for s, l, c, g in zip(start, L, cutoff, group)]


def batch_queries(table, params):
    return [build_query(table, q1_spec(date_cutoff=c, group=g), b, e) for (b, e, c, g) in params]


BATCH_KEY_COLUMNS = {"a": ("flaga",), "b": ("flagb",), "ab": ("flaga", "flagb"), "month": ()}
BATCH_COMMON_COLUMNS = ("dt", "qty", "price", "disc", "tax")


def batch_bytes(params):
    """Algorithmic bytes: rows x (dt + group-key columns + 4 summed columns),
    summed over the queries (a row read by several queries counts each time)."""
    key_bytes = {"a": 1, "b": 1, "ab": 2, "month": 0}
    return sum((e - b) * (4 + key_bytes[g] + 32) for (b, e, _, g) in params)


def _union_rows(ranges):
    total, end = 0, -1
    for b, e in sorted(ranges):
        if e <= end:
            continue
        total += e - max(b, end)
        end = e
    return total


def batch_unique_bytes(params):
    """Bytes of the distinct column cells the batch reads: per column, the union
    of the ranges of the queries that read it x the column width."""
    out = 0
    for c in BATCH_COMMON_COLUMNS + ("flaga", "flagb"):
        ranges = [(b, e) for (b, e, _, g) in params
                  if c in BATCH_COMMON_COLUMNS or c in BATCH_KEY_COLUMNS[g]]
        out += _union_rows(ranges) * ROW_BYTES[c]
    return out


# ---- serving mix: open-loop Poisson arrivals over the resident tables ----------------
SERVE_SCAN, SERVE_JOIN, SERVE_FULL = 0.70, 0.25, 0.05


def _log_uniform(rng, lo, hi, n=None):
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), n))).astype(np.int64)


def serve_mix(n, table_rows, build_rows, probe_rows, seed=SEED, lmax=4_000_000,
              build_min=1_000, build_max=2_000_000):
    """Seeded request descriptions (kind, begin, end, build_rows):
    70 % scan-aggregates, half Q1-shape (WHERE dt <= 2435 GROUP BY flaga, flagb)
    and half Q6-shape, over a range of log-uniform length in [1K, lmax];
    25 % join-aggregates GROUP BY the build payload, over a build-side prefix of
    log-uniform length in [build_min, build_max] and a probe range of log-uniform
    length in [1K, lmax]; 5 % full-table Q1-shape aggregates."""
    rng = np.random.Generator(np.random.PCG64((seed ^ 0x5E7E) & 0xFFFFFFFF))
    out = []
    for _ in range(n):
        u = rng.random()
        if u < SERVE_SCAN:
            kind = "q1" if rng.random() < 0.5 else "q6"
            L = int(_log_uniform(rng, 1_000, min(lmax, table_rows)))
            b = int(rng.integers(0, table_rows - L + 1))
            out.append((kind, b, b + L, 0))
        elif u < SERVE_SCAN + SERVE_JOIN:
            nb = int(_log_uniform(rng, min(build_min, build_rows), min(build_max, build_rows)))
            L = int(_log_uniform(rng, 1_000, min(lmax, probe_rows)))
            b = int(rng.integers(0, probe_rows - L + 1))
            out.append(("join", b, b + L, nb))
        else:
            out.append(("full", 0, table_rows, 0))
    return out


def serve_requests(table, build, probe, mix):
    """ServeRequest objects for serve_mix() descriptions over device tables
    (fact table columns; build {bkey, pay}; probe {fk, val, gkey})."""
    from .serve import agg_request, join_request
    out = []
    for kind, b, e, nb in mix:
        if kind == "q1" or kind == "full":
            out.append(agg_request(build_query(table, q1_spec(), b, e)))
        elif kind == "q6":
            out.append(agg_request(build_query(table, q6_spec(), b, e)))
        else:
            out.append(join_request(build["bkey"], probe["fk"], probe["val"],
                                    A.JOIN_GROUP_BUILD_PAYLOAD, ngroups=JOIN_PAY_CARD,
                                    build_payload=build["pay"], build_rows=nb, row_begin=b,
                                    row_end=e))
    return out


def serve_bytes(mix):
    """Algorithmic bytes per request: scan rows x referenced widths; joins:
    build rows x 12 (key, payload) + probe rows x 16 (key, value)."""
    out = []
    for kind, b, e, nb in mix:
        if kind in ("q1", "full"):
            out.append((e - b) * Q1_BYTES_PER_ROW)
        elif kind == "q6":
            out.append((e - b) * Q6_BYTES_PER_ROW)
        else:
            out.append(nb * 12 + (e - b) * 16)
    return out


def poisson_offsets(n, qps, seed=SEED):
    """Arrival offsets (ns, non-decreasing) of an open-loop Poisson stream."""
    rng = np.random.Generator(np.random.PCG64((seed ^ 0xA5517) & 0xFFFFFFFF))
    gaps = rng.exponential(1e9 / qps, n)
    return np.cumsum(gaps).astype(np.int64)


# ---- mixed-shape batch: every query form in one batch --------------------------------------
MIXED_FORMS = ("q1", "q1", "q6", "q6_5preds", "minmax", "bitmap", "price_i32", "wide_groups")


def mixed_batch_params(n_queries, table_rows, seed=SEED):
    """(form, row_begin, row_end, date cutoff, group, literal) per query: the compiled
    Q1 / Q6 shapes, Q6 with a fifth predicate (qty <> literal), MIN / MAX beside sums,
    an input selection bitmap (whole table), price as int32 (another column layout)
    and GROUP BY (dt, flagb): 2526 x 2 groups.  Every range is log-uniform in 1K..4M
    rows; the bitmap queries start on a multiple of 64 rows."""
    rng = np.random.default_rng(seed ^ 0x313D)
    out = []
    for k in range(n_queries):
        form = MIXED_FORMS[k % len(MIXED_FORMS)]
        L = int(_log_uniform(rng, 1_000, min(4_000_000, table_rows)))
        b = int(rng.integers(0, table_rows - L + 1))
        e = b + L
        if form == "bitmap":  # the range starts on a bitmap word: a view of one bitmap
            b &= ~63
        out.append((form, b, e, int(rng.integers(0, 2526)), ["ab", "a", "b", "month", ""][k % 5],
                    int(rng.integers(1, 50))))
    return out


class _BitmapView:
    """The words of a table-wide selection bitmap from row `row` (a multiple of 64) on."""

    class _Buf:
        def __init__(self, ptr):
            self.ptr = ptr

    def __init__(self, bitmap, row):
        assert row % 64 == 0
        self.buf = self._Buf(bitmap.buf.ptr + (row // 64) * 8)


def mixed_batch_queries(table, price_i32, bitmap, params):
    """tgp_agg_query objects of mixed_batch_params() descriptions."""
    q1n, q1p, q1parts, q1aggs = q1_spec()
    q6n, q6p, q6parts, q6aggs = q6_spec()
    out = []
    for form, b, e, cut, grp, lit in params:
        sel = None
        if form == "q1":
            nm, pr, pa, ag = q1_spec(date_cutoff=cut, group=grp)
            cols = [table[x] for x in nm]
        elif form == "q6":
            cols, pr, pa, ag = [table[x] for x in q6n], q6p, q6parts, q6aggs
        elif form == "q6_5preds":
            cols, pr, pa, ag = [table[x] for x in q6n], q6p + [(2, A.NE, lit, True)], q6parts, q6aggs
        elif form == "minmax":
            cols, pr, pa = [table[x] for x in q1n], q1p, q1parts
            ag = [("min", 2), ("max_u64", 1), ("sum", 1, [(2, 0, 1)]),
                  ("max", 1, [(1, 0, 1), (3, 100, -1)])]
        elif form == "bitmap":  # rows [b, e) select by bits b.. of the table-wide bitmap
            cols, pr, pa, ag = [table[x] for x in q1n], q1p, q1parts, q1aggs
            sel = _BitmapView(bitmap, b)
        elif form == "price_i32":
            cols = [table[x] if x != "price" else price_i32 for x in q1n]
            pr, pa, ag = q1p, q1parts, q1aggs
        else:
            cols, pr, ag = [table[x] for x in q1n], q1p, q1aggs
            pa = [(0, 1, 0, 2526), (6, 1, 0, 2)]
        out.append(make_agg_query(cols, pr, pa, ag, b, e, sel))
    return out

