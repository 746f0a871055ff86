#!/usr/bin/env python3
"""Benchmark of the scan / group-aggregate path on one MI355X.

Headline workload (one step = one batch): 4096 filter + GROUP BY queries over a
256M-row fact table resident in HBM (row ranges log-uniform in 1K..4M rows,
date cutoffs uniform, group keys drawn from {flagA, flagB, flagA x flagB,
date/30}; each query computes COUNT(*), SUM(qty), SUM(price),
SUM(price*(100-disc)), SUM(price*(100-disc)*(100+tax)) exactly).  The batch runs
as one persistent-kernel launch replayed from a hipGraph, with launch settings
from the batch autotuner.  value = queries per second of the batched form.

Also measured in the same run:
  - the one-launch-per-query form of the same 4096 queries (tgp_group_agg per
    query, launches queued back to back on one stream), its queries/s and
    per-query completion latencies (p50/p99), and the speedup of the batch;
  - the 60M-row scan-aggregate queries over the table's first 60M rows:
    (a) Q1-shape WHERE dt <= 2435 GROUP BY flagA, flagB (6 groups), and
    (b) Q6-shape WHERE 731 <= dt < 1096 AND disc BETWEEN 5 AND 7 AND qty < 24
        -> SUM(price*disc), in GB/s of algorithmic bytes and as a fraction of
        HBM peak;
  - the oracle's restatement of the reference CPU path on a bounded sample of
    the batch (cpu_baseline).

Algorithmic bytes of a query: rows in its range x native widths of the columns
it reads (date 4 B + group-key columns 1 B each + four 8-byte summed columns).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--tuning wg,chunk,order,cached,grab]
       python bench.py --calibrate   (known-size kernels for scaling PMC byte counters:
                                      k_fill_u64 writes and k_minmax reads exactly 512 MiB)
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
METRIC = "batched queries/s + p99 vs per-query launches; scan GB/s as % of HBM peak"
SCAN_ROWS = 60_000_000


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pct(xs, p):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), p))


def host_cpu():
    """The host CPU as lscpu reports it: model, sockets, physical cores, logical CPUs,
    and the CPUs this process may run on."""
    try:
        with open("/proc/cpuinfo") as f:
            info = f.read()
    except OSError:
        return None
    model, sockets, cores = None, set(), set()
    phys = None
    for line in info.splitlines():
        key, _, val = line.partition(":")
        key, val = key.strip(), val.strip()
        if key == "model name" and model is None:
            model = val
        elif key == "physical id":
            phys = val
            sockets.add(val)
        elif key == "core id":
            cores.add((phys, val))
    logical = info.count("processor\t:")
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else logical
    return (f"{model}; {len(sockets) or 1} socket(s), {len(cores) or logical} cores, "
            f"{logical} logical CPUs, {usable} usable by this process")


def _cells(arr):
    """A column as the reference stores it: one u64 cell per value (csv.rs:144-161;
    signed values sign-extended)."""
    return arr.astype(np.int64 if arr.dtype.kind == "i" else np.uint64)


def _threads():
    return (len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
            else os.cpu_count()) or 1


def cpu_baseline(params, rows_cap, join_tables=None, reps=3):
    """The reference's CPU path restated in the oracle (u64 cells; per-leaf filter
    passes, AVX-512 4x8 where the host has it; tzcnt compaction; 65 536-row chunk
    hash grouping; f64 sum_vec; the join as chained table + materialisation +
    accumulator grouping), on every host core, timed around filter -> aggregate
    (column generation excluded):
      C4  the first batch queries up to `rows_cap` rows, one run, as queries/s
          scaled by rows to the whole batch (the headline `value`);
      C2a / C2b  the two scan-aggregates over 60M rows, median of `reps`;
      C3  the two join aggregates (2M build, 60M probe), one run each."""
    from oracle import binding as B
    from turbogp_amd import workloads as W
    threads = _threads()
    gen = lambda c, b, e: _cells(B.synth_uniform(W.FACT_SCHEMA[c][1], W.SEED, W.FACT_SCHEMA[c][0],
                                                 W.FACT_SCHEMA[c][2], W.FACT_SCHEMA[c][3], b, e))
    # ---- C4 sample
    sample, rows = [], 0
    for p in params:
        if rows + (p[1] - p[0]) > rows_cap and sample:
            break
        sample.append(p)
        rows += p[1] - p[0]
    t4 = 0.0
    for (b, e, cutoff, g) in sample:
        names, preds, parts, aggs = W.q1_spec(date_cutoff=cutoff, group=g)
        q = B.make_query([gen(c, b, e) for c in names], preds, parts, aggs)
        t0 = time.perf_counter()
        B.group_agg(q, mode=1, nthreads=threads)
        t4 += time.perf_counter() - t0
    total_rows = sum(p[1] - p[0] for p in params)
    batch_s = t4 * total_rows / rows
    log(f"cpu C4: {len(sample)} queries, {rows} rows, {t4:.2f} s")
    legs = {}
    # ---- C2a / C2b over 60M rows
    host = {c: gen(c, 0, SCAN_ROWS) for c in W.FACT_SCHEMA}
    for name, spec, bpr in (("c2a_q1_groupby_60M", W.q1_spec(), W.Q1_BYTES_PER_ROW),
                            ("c2b_q6_scan_sum_60M", W.q6_spec(), W.Q6_BYTES_PER_ROW)):
        names, preds, parts, aggs = spec
        q = B.make_query([host[c] for c in names], preds, parts, aggs)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            B.group_agg(q, mode=1, nthreads=threads)
            ts.append(time.perf_counter() - t0)
        t = statistics.median(ts)
        log(f"cpu {name}: {t * 1e3:.1f} ms")
        legs[name] = {"ms": round(t * 1e3, 2), "rows_per_s": round(SCAN_ROWS / t, 1),
                      "GB_s": round(SCAN_ROWS * bpr / 1e9 / t, 2), "runs": reps}
    del host
    # ---- C3 join aggregates
    if join_tables is not None:
        from turbogp_amd.joinload import join_bytes
        build, probe = join_tables
        bk, bp = build["bkey"].to_numpy(), build["pay"].to_numpy()
        fk, val, gk = probe["fk"].to_numpy(), probe["val"].to_numpy(), probe["gkey"].to_numpy()
        jb = join_bytes(len(bk), len(fk))
        for name, gb, ng, key in (("c3_group_by_payload", 0, W.JOIN_PAY_CARD, "by_payload"),
                                  ("c3_group_by_probe_key", 1, W.JOIN_GKEY_CARD, "by_gkey")):
            t0 = time.perf_counter()
            cnt, _, _ = B.join_agg(bk, bp, fk, val, gk, gb, 0, ng, nthreads=threads)
            t = time.perf_counter() - t0
            assert int(cnt.sum()) == len(fk)
            log(f"cpu {name}: {t * 1e3:.1f} ms")
            legs[name] = {"ms": round(t * 1e3, 1), "probes_per_s": round(len(fk) / t, 1),
                          "GB_s": round(jb[key] / 1e9 / t, 2), "runs": 1}
    simd = B.lib().orc_simd_level()
    return {"value": round(len(params) / batch_s, 2), "unit": "queries/s", "cores": threads,
            "kind": "port", "host": host_cpu(), "simd": "avx512" if simd >= 512 else "scalar",
            "sample": f"C4: first {len(sample)} of the {len(params)} batch queries ({rows} rows, "
                      f"{t4:.2f} s), scaled by rows to the whole batch ({total_rows} rows, "
                      f"{batch_s:.1f} s); legs: C2a/C2b over 60M rows, C3 2M x 60M; the "
                      f"reference algorithm (u64 cells, f64 sums) in the oracle, built -O3 "
                      f"-march=x86-64-v3 -fopenmp with the AVX-512 filter chosen at run time",
            "legs": legs}


def measured_traffic(name, rows):
    """HBM bytes per launch of a kernel from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs (scaled by known-size read and write kernels), for the workload they
    were taken on."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            t = json.load(f)[name]
    except (OSError, KeyError, ValueError):
        return None
    if t.get("rows") != rows:
        return None
    return t["fetch_bytes"] + t["write_bytes"]


def bench_join(ctx, args):
    """Hash join: 2M-row build side (unique int64 keys, int32 payload 0..24),
    60M-row probe side (Zipf(1.1) foreign keys, int64 value, int32 key with 1.5M
    values); build, probe count, and the fused probe -> GROUP BY -> SUM, COUNT
    by the build payload (25 groups) and by the probe key (1.5M groups)."""
    from turbogp_amd import Event, GroupAggBuffers, HashTable, _capi as A
    from turbogp_amd import workloads as W
    from turbogp_amd.joinload import join_bytes, make_join_tables
    build, probe = make_join_tables(ctx)
    nb, npr = W.JOIN_BUILD_ROWS, W.JOIN_PROBE_ROWS
    ev = [Event(ctx) for _ in range(2)]

    def timed(fn, reps):
        ts = []
        for _ in range(reps):
            ev[0].record()
            fn()
            ev[1].record()
            ts.append(ev[0].elapsed_ms(ev[1]))
        return statistics.median(ts)

    tables = []
    build_ms = timed(lambda: tables.append(HashTable(ctx, build["bkey"], build["pay"])), 5)
    for t in tables[:-1]:
        t.free()
    t = tables[-1]
    cnt = DeviceCounter(ctx)
    count_ms = timed(lambda: t.probe_count(probe["fk"], result=cnt.buf, fetch=False), 10)
    matches = cnt.value()
    rp = GroupAggBuffers(ctx, W.JOIN_PAY_CARD, 1)
    rg = GroupAggBuffers(ctx, W.JOIN_GKEY_CARD, 1)
    pay_ms = timed(lambda: t.probe_agg(probe["fk"], probe["val"], A.JOIN_GROUP_BUILD_PAYLOAD,
                                       ngroups=W.JOIN_PAY_CARD, result=rp, fetch=False), 10)
    gkey_ms = timed(lambda: t.probe_agg(probe["fk"], probe["val"], A.JOIN_GROUP_PROBE_COL,
                                        ngroups=W.JOIN_GKEY_CARD, group_col=probe["gkey"],
                                        result=rg, fetch=False), 10)
    res_pay = rp.fetch()
    assert matches == npr and int(res_pay.count.sum()) == npr, (matches, res_pay.count.sum())
    sums_hot = (res_pay.sums(), rg.fetch().sums())
    # the same three calls without the hot-key image, for comparison
    ctx.tune(A.TUNE_JOIN_HOT, 1)
    no_hot = {
        "probe_count_ms": timed(lambda: t.probe_count(probe["fk"], result=cnt.buf, fetch=False), 10),
        "group_by_payload_ms": timed(lambda: t.probe_agg(
            probe["fk"], probe["val"], A.JOIN_GROUP_BUILD_PAYLOAD, ngroups=W.JOIN_PAY_CARD,
            result=rp, fetch=False), 10),
        "group_by_probe_key_ms": timed(lambda: t.probe_agg(
            probe["fk"], probe["val"], A.JOIN_GROUP_PROBE_COL, ngroups=W.JOIN_GKEY_CARD,
            group_col=probe["gkey"], result=rg, fetch=False), 10)}
    ctx.tune(A.TUNE_JOIN_HOT, 0)
    no_hot = {k: round(v, 4) for k, v in no_hot.items()}
    assert (rp.fetch().sums(), rg.fetch().sums()) == sums_hot, "hot-key image changed a result"
    log(f"join: count {count_ms:.3f} ms, payload {pay_ms:.3f} ms, probe key {gkey_ms:.3f} ms; "
        f"without the hot-key image {no_hot}")
    hash_group = bench_hash_group(ctx, probe)
    # probes that mostly miss: 10 % of the foreign keys kept, the rest random int64 keys
    rng = np.random.default_rng(W.SEED ^ 0xB100)
    fk = probe["fk"].to_numpy()
    miss = rng.random(npr) >= 0.1
    fk[miss] = rng.integers(-2 ** 63, 2 ** 63 - 1, int(miss.sum()), dtype=np.int64)
    from turbogp_amd import DeviceColumn
    fk10 = DeviceColumn.from_numpy(ctx, fk)
    del fk
    bloom = {"probe_rows": npr, "hit_fraction": 0.1}
    tf = HashTable(ctx, build["bkey"], build["pay"])
    t0 = time.perf_counter()
    tf.add_bloom(8)
    ctx.sync()
    bloom["bloom_build_ms_wall"] = round((time.perf_counter() - t0) * 1e3, 4)
    for name, tab in (("no_filter", t), ("bloom_8_bits_per_key", tf)):
        c10 = timed(lambda: tab.probe_count(fk10, result=cnt.buf, fetch=False), 10)
        m10 = cnt.value()
        p10 = timed(lambda: tab.probe_agg(fk10, probe["val"], A.JOIN_GROUP_BUILD_PAYLOAD,
                                          ngroups=W.JOIN_PAY_CARD, result=rp, fetch=False), 10)
        c100 = timed(lambda: tab.probe_count(probe["fk"], result=cnt.buf, fetch=False), 10)
        p100 = timed(lambda: tab.probe_agg(probe["fk"], probe["val"], A.JOIN_GROUP_BUILD_PAYLOAD,
                                           ngroups=W.JOIN_PAY_CARD, result=rp, fetch=False), 10)
        bloom[name] = {"hit10_probe_count_ms": round(c10, 4), "hit10_matches": int(m10),
                       "hit10_group_by_payload_ms": round(p10, 4),
                       "hit100_probe_count_ms": round(c100, 4),
                       "hit100_group_by_payload_ms": round(p100, 4)}
        log(f"join {name}: 10 % hits count {c10:.3f} ms, payload {p10:.3f} ms; "
            f"100 % hits count {c100:.3f} ms, payload {p100:.3f} ms")
    assert bloom["no_filter"]["hit10_matches"] == bloom["bloom_8_bits_per_key"]["hit10_matches"]
    tf.free()
    fk10.free()
    # the random-read roofline of a probe: 60M uniformly random keys (every one a hit)
    # into unique-key tables of 4 MiB (L2 resident) to 256 MiB of slots
    rate = {}
    for nbk in (1 << 17, 1 << 19, 1 << 21, 1 << 23):
        # keys 0 .. nbk-1 (the table's SplitMix64 slot hash scatters them), probes
        # uniform over the same range
        keys = DeviceColumn.from_numpy(ctx, np.arange(nbk, dtype=np.int64))
        probes = W.device_uniform(ctx, npr, A.I64, 0x9E37, 0, nbk - 1)
        tu = HashTable(ctx, keys)
        ms_u = timed(lambda: tu.probe_count(probes, result=cnt.buf, fetch=False), 5)
        assert cnt.value() == npr
        ctx.tune(A.TUNE_JOIN_HOT, 1)
        ms_nohot = timed(lambda: tu.probe_count(probes, result=cnt.buf, fetch=False), 5)
        ctx.tune(A.TUNE_JOIN_HOT, 0)
        _, cap, _ = tu.info()
        rate[f"table_{cap * 16 >> 20}_MiB"] = {
            "build_keys": nbk, "ms": round(ms_nohot, 4),
            "G_probes_s": round(npr / (ms_nohot / 1e3) / 1e9, 1)}
        log(f"uniform probes into a {cap * 16 >> 20} MiB table: {ms_nohot:.3f} ms "
            f"({npr / (ms_nohot / 1e3) / 1e9:.1f} G/s; {ms_u:.3f} ms with the image policy)")
        tu.free()
        keys.free()
        probes.free()
    jb = join_bytes(nb, npr)
    for e in ev:
        e.free()
    t.free()
    rp.free()
    rg.free()
    cnt.free()
    return {"build_ms": round(build_ms, 4), "probe_count_ms": round(count_ms, 4),
            "probe_count_Gprobes_s": round(npr / (count_ms / 1e3) / 1e9, 2),
            "group_by_payload": {"ms": round(pay_ms, 4), "groups": W.JOIN_PAY_CARD,
                                 "GB_s": round(jb["by_payload"] / 1e9 / (pay_ms / 1e3), 1)},
            "group_by_probe_key": {"ms": round(gkey_ms, 4), "groups": W.JOIN_GKEY_CARD,
                                   "GB_s": round(jb["by_gkey"] / 1e9 / (gkey_ms / 1e3), 1)},
            "matches": int(matches), "without_hot_key_image": no_hot,
            "uniform_probe_rate_by_table_size": rate,
            "hash_group_by_60M": hash_group, "bloom_prefilter": bloom,
            "_tables": (build, probe)}


def bench_hash_group(ctx, probe):
    """GROUP BY arbitrary keys through the device hash table (tgp_hash_group_agg_sync, a
    synchronous call: host wall time, result buffers allocated beforehand) over the
    60M probe rows: by the int32 key (1.5M groups) and by the sparse int64 foreign key
    (the 2M build keys, Zipf-distributed), SUM(val) and COUNT(*)."""
    import ctypes as C
    from turbogp_amd import DeviceBuffer, make_agg_query, _capi as A
    from turbogp_amd import workloads as W
    out = {}
    for name, col, cap in (("by_int32_key_1_5M_groups", "gkey", W.JOIN_GKEY_CARD),
                           ("by_int64_foreign_key", "fk", W.JOIN_BUILD_ROWS)):
        q = make_agg_query([probe[col], probe["val"]], (), (), [("sum", 1, [(1, 0, 1)])])
        hq = A.HashGroupQuery()
        hq.agg, hq.nkeys = q, 1
        hq.key_col[0] = 0
        bufs = [DeviceBuffer(ctx, cap * 8), DeviceBuffer(ctx, cap * 4), DeviceBuffer(ctx, cap * 8),
                DeviceBuffer(ctx, cap * 16)]
        try:
            res = A.HashGroupResult(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, cap)
            ng = C.c_uint64()
            ts = []
            for _ in range(6):
                t0 = time.perf_counter()
                A.check(A.lib().tgp_hash_group_agg_sync(ctx.handle, C.byref(hq), C.byref(res),
                                                        C.byref(ng), None), "hash group")
                ts.append((time.perf_counter() - t0) * 1e3)
            ms = statistics.median(ts[1:])
            out[name] = {"ms": round(ms, 4), "groups": int(ng.value),
                         "rows_per_s": round(W.JOIN_PROBE_ROWS / (ms / 1e3), 1)}
            log(f"hash group {name}: {ms:.3f} ms, {ng.value} groups")
        finally:
            for b in bufs:
                b.free()
    # the int32 key as dense group ids (its range known): the generic engine's
    # partitioned form (rows sorted into 733 partitions of 2048 groups), events
    # around the asynchronous call
    from turbogp_amd import Event, GroupAggBuffers
    q = make_agg_query([probe["gkey"], probe["val"]], (), [(0, 1, 0, W.JOIN_GKEY_CARD)],
                       [("sum", 1, [(1, 0, 1)])])
    res = GroupAggBuffers(ctx, W.JOIN_GKEY_CARD, 1)
    ev = [Event(ctx) for _ in range(2)]
    try:
        ts = []
        for i in range(8):
            ev[0].record()
            ctx.group_agg_async(q, res)
            ev[1].record()
            if i >= 2:
                ts.append(ev[0].elapsed_ms(ev[1]))
        ms = statistics.mean(ts)
        got = res.fetch()
        out["by_int32_key_dense_ids_partitioned"] = {
            "ms": round(ms, 4), "groups": int((got.count != 0).sum()),
            "rows_per_s": round(W.JOIN_PROBE_ROWS / (ms / 1e3), 1)}
        log(f"dense-id group by int32 key: {ms:.3f} ms")
    finally:
        res.free()
        for e in ev:
            e.free()
    return out


class DeviceCounter:
    """One device word for tgp_hash_probe_count."""

    def __init__(self, ctx):
        from turbogp_amd import DeviceBuffer
        self.buf = DeviceBuffer(ctx, 8)

    def value(self):
        return int(self.buf.download(np.uint64, 1)[0])

    def free(self):
        self.buf.free()


def bench_generic(ctx, table, args):
    """Queries of no shape compiled ahead of time, on 60M rows: Q1 with price as
    int32 (another column layout; 6 groups) through the generic tile engine and
    through its shape-specialised kernel compiled at run time (the compile is
    timed apart, before the steps), and Q1 grouped by (dt, flagb) (5052 groups,
    generic engine's result table; too many groups for a specialised form)."""
    from turbogp_amd import Event, GroupAggBuffers, _capi as A
    from turbogp_amd import workloads as W
    import ctypes as C
    price32 = W.device_uniform(ctx, SCAN_ROWS, A.I32, W.FACT_SCHEMA["price"][0], 90_000,
                               10_500_000)
    out = {}
    ev = [Event(ctx) for _ in range(2)]

    def jit_info():
        c, f, l, ms = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_double()
        A.check(A.lib().tgp_jit_info(C.byref(c), C.byref(f), C.byref(l), C.byref(ms)), "jit")
        return c.value, f.value, ms.value

    def timed(q, res):
        for _ in range(args.warmup):
            ctx.group_agg_async(q, res)
        ts = []
        for _ in range(args.steps):
            ev[0].record()
            ctx.group_agg_async(q, res)
            ev[1].record()
            ts.append(ev[0].elapsed_ms(ev[1]))
        return statistics.mean(ts)

    try:
        for name, form, bpr in (("q1_price_i32", "price_i32", 34),
                                ("q1_5052_groups", "wide_groups", 38)):
            q = W.mixed_batch_queries(table, price32, None, [(form, 0, SCAN_ROWS, 0, "ab", 0)])[0]
            g = C.c_uint64()
            A.check(A.lib().tgp_group_agg_ngroups(C.byref(q), C.byref(g)), "ngroups")
            st, sh = C.c_int32(), C.c_int32()
            A.check(A.lib().tgp_group_agg_explain(C.byref(q), C.byref(st), C.byref(sh)), "explain")
            res = GroupAggBuffers(ctx, g.value, q.naggs)
            runs = [("generic", 1)]
            if sh.value == A.SPEC_SHAPE_RUNTIME:  # compiled before the first call returns
                runs.append(("compiled_at_run_time", 2))
            if st.value == 5:  # partitioned by group range; also time the one global table
                runs.append(("global_table", 1))
            for path, mode in runs:
                ctx.tune(A.TUNE_JIT, mode)
                ctx.tune(A.TUNE_AGG_PART, 1 if path == "global_table" else 0)
                extra = {}
                if mode == 2:
                    c0, f0, ms0 = jit_info()
                    t0 = time.perf_counter()
                    ctx.group_agg_async(q, res)
                    ctx.sync()
                    c1, f1, ms1 = jit_info()
                    assert f1 == f0, "run-time compile failed"
                    extra = {"first_call_ms_wall": round((time.perf_counter() - t0) * 1e3, 1),
                             "compile_ms": round(ms1 - ms0, 1), "shapes_compiled": c1 - c0}
                ms = timed(q, res)
                gbs = SCAN_ROWS * bpr / 1e9 / (ms / 1e3)
                out[f"{name}_{path}"] = {"ms": round(ms, 4), "GB_s": round(gbs, 1),
                                         "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                                         "bytes_per_row": bpr, "groups": int(g.value),
                                         "strategy": 3 if path == "global_table"
                                         else int(st.value), **extra}
                log(f"{name} {path}: {ms:.3f} ms, {gbs:.0f} GB/s {extra}")
                if path == "global_table":
                    got = res.fetch()
                    ctx.tune(A.TUNE_AGG_PART, 0)
                    ctx.group_agg_async(q, res)
                    want = res.fetch()
                    assert got.sums() == want.sums() and list(got.count) == list(want.count)
            ctx.tune(A.TUNE_AGG_PART, 0)
            if sh.value == A.SPEC_SHAPE_RUNTIME:
                ctx.tune(A.TUNE_JIT, 2)
                ctx.group_agg_async(q, res)
                want = res.fetch()
                ctx.tune(A.TUNE_JIT, 1)
                ctx.group_agg_async(q, res)
                got = res.fetch()
            ctx.tune(A.TUNE_JIT, 0)
            if sh.value == A.SPEC_SHAPE_RUNTIME:
                assert got.sums() == want.sums() and list(got.count) == list(want.count), name
            res.free()
    finally:
        ctx.tune(A.TUNE_JIT, 0)
        ctx.tune(A.TUNE_AGG_PART, 0)
        for e in ev:
            e.free()
        price32.free()
    return out


def bench_predicates(ctx, table, args):
    """WHERE terms outside the AND-of-comparisons filter, over the table's first 60M
    rows: an expression evaluated per row (price * disc > 4e7, the interpreter's
    fallback, tgp_scan_expr) narrowing an all-rows selection, and a FLOAT column
    against a literal (tgp_scan_filter_f64) writing a fresh bitmap."""
    from turbogp_amd import DeviceColumn, Event, _capi as A
    from turbogp_amd import expr as X
    n = SCAN_ROWS
    ev = [Event(ctx) for _ in range(2)]
    out = {}

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        ts = []
        for _ in range(args.steps):
            ev[0].record()
            fn()
            ev[1].record()
            ts.append(ev[0].elapsed_ms(ev[1]))
        return statistics.mean(ts)
    try:
        cols = [table["price"], table["disc"]]
        prog = X.compile_expr(X.BinOp(">", X.BinOp("*", X.Col("price"), X.Col("disc")),
                                      X.Lit(40_000_000)),
                              lambda nm: ({"price": 0, "disc": 1}[nm], X.INT))
        bm = ctx.scan_expr(cols, prog, 0, n)
        ms = timed(lambda: ctx.scan_expr(cols, prog, 0, n, bm))  # narrowing an all-ones bitmap
        bm.buf.fill_u64(~0, nwords=bm.nwords)
        ctx.scan_expr(cols, prog, 0, n, bm)
        cnt = bm.count_ones()
        bm.free()
        byts = n * (8 + 8) + 2 * n // 8
        out["expr_price_x_disc_gt_4e7"] = {"ms": round(ms, 4), "rows_selected": cnt,
                                           "GB_s": round(byts / 1e9 / (ms / 1e3), 1),
                                           "hbm_frac": round(byts / 1e9 / (ms / 1e3) / HBM_PEAK_GBS, 4),
                                           "bytes": byts}
        log(f"expression price*disc > 4e7 over {n} rows: {ms:.3f} ms, {cnt} rows")
        import ctypes as C
        from turbogp_amd import DeviceBitmap, DeviceBuffer
        fp = (table["price"].buf.download(np.int64, n).astype(np.float64) / 100.0).view(np.uint64)
        fc = DeviceColumn.from_numpy(ctx, fp)
        del fp
        fbm, cntb = DeviceBitmap(ctx, n), DeviceBuffer(ctx, 8)
        d = fc.desc()
        ms = timed(lambda: A.check(A.lib().tgp_scan_filter_f64(
            ctx.handle, C.byref(d), A.LT, 0, 50_000.25, 0, n, fbm.buf.ptr, None, None), "f64"))
        cntb.fill_u64(0)
        A.check(A.lib().tgp_scan_filter_f64(ctx.handle, C.byref(d), A.LT, 0, 50_000.25, 0, n,
                                            fbm.buf.ptr, cntb.ptr, None), "f64")
        sel = int(cntb.download(np.uint64, 1)[0])
        for x in (fbm, cntb, fc):
            x.free()
        byts = n * 8 + n // 8
        out["f64_lt_filter"] = {"ms": round(ms, 4), "rows_selected": sel,
                                "GB_s": round(byts / 1e9 / (ms / 1e3), 1),
                                "hbm_frac": round(byts / 1e9 / (ms / 1e3) / HBM_PEAK_GBS, 4),
                                "bytes": byts}
        log(f"f64 filter over {n} rows: {ms:.3f} ms")
    finally:
        for e in ev:
            e.free()
    return out


def bench_mixed(ctx, table):
    """A 4096-query batch of every query form (compiled Q1 / Q6 shapes, a fifth
    predicate, MIN / MAX, an input selection bitmap, an int32 price column, 5052
    groups) over the 256M-row table: the batch launch against one launch per query,
    results compared."""
    from turbogp_amd import DeviceBitmap, Event, _capi as A
    from turbogp_amd import workloads as W
    from turbogp_amd.batch import Batch
    n = W.BATCH_TABLE_ROWS
    price32 = W.device_uniform(ctx, n, A.I32, W.FACT_SCHEMA["price"][0], 90_000, 10_500_000)
    rng = np.random.default_rng(W.SEED ^ 0xB17)
    words = rng.integers(0, 2 ** 63, (n + 63) // 64, dtype=np.int64).astype(np.uint64)
    bm = DeviceBitmap(ctx, n)
    bm.buf.upload(words)
    params = W.mixed_batch_params(W.BATCH_QUERIES, n)
    qs = W.mixed_batch_queries(table, price32, bm, params)
    b = Batch(ctx, len(qs))
    try:
        # queries with a specialised shape of no ahead-of-time kernel: compiled at
        # planning (as a server's later batches find them compiled)
        ctx.tune(A.TUNE_JIT, 2)
        try:
            b.set_queries(qs)
        finally:
            ctx.tune(A.TUNE_JIT, 0)
        info = b.info()
        log(f"mixed batch: {info}")
        # the shared scan's tuning of the 4096-query batch (autotuning this batch
        # chose settings that timed slower when measured again)
        b.set_tuning(0, True, 128, 3, True, 1)
        tuned = b.tuning()
        ev = [Event(ctx) for _ in range(2)]
        b.run()
        ctx.sync()
        ts = []
        for _ in range(10):
            ev[0].record()
            b.run()
            ev[1].record()
            ts.append(ev[0].elapsed_ms(ev[1]))
        batched = b.results()
        tot, _ = b.run_timed(A.BATCH_PER_QUERY)
        single = b.results()
        same = all(np.array_equal(x.count, y.count) and np.array_equal(x.sums_words, y.sums_words)
                   and x.n_out_of_range == y.n_out_of_range for x, y in zip(batched, single))
        assert same, "mixed batch differs from per-query launches"
        for e in ev:
            e.free()
    finally:
        b.close()
        bm.free()
        price32.free()
    ms = statistics.mean(ts)
    log(f"mixed batch: {ms:.3f} ms, per-query {tot:.3f} ms")
    return {"queries": len(qs), "forms": {f: sum(1 for p in params if p[0] == f)
                                          for f in dict.fromkeys(W.MIXED_FORMS)},
            "batched_queries": info["batched"], "fallback": info["fallback"],
            "tuning": tuned, "run_time_compiled_shapes": "compiled at planning (TGP_TUNE_JIT 2)",
            "ms": round(ms, 4), "queries_per_s": round(len(qs) / (ms / 1e3), 1),
            "per_query_ms": round(tot, 4), "speedup_vs_per_query": round(tot / ms, 2),
            "results_identical_batched_vs_per_query": bool(same)}


def bench_distinct(ctx, table):
    """COUNT(DISTINCT price) and COUNT(DISTINCT qty) over the table's first 60M
    rows: HyperLogLog (the reference's path for >= 100K selected rows) and the
    exact device set; and the 99 999-row exact case of the AUTO switch."""
    from turbogp_amd import Event, _capi as A
    ev = [Event(ctx) for _ in range(2)]

    def timed(fn, reps):
        ts, out = [], None
        for _ in range(reps):
            ev[0].record()
            out = fn()
            ev[1].record()
            ts.append(ev[0].elapsed_ms(ev[1]))
        return statistics.median(ts), out

    res = {}
    n = SCAN_ROWS
    for name in ("price", "qty"):
        col = table[name]
        ms_h, est = timed(lambda: ctx.count_distinct(col, 0, n, mode=A.DISTINCT_HLL), 10)
        ms_e, exact = timed(lambda: ctx.count_distinct(col, 0, n, mode=A.DISTINCT_EXACT), 3)
        gbs = n * 8 / 1e9 / (ms_h / 1e3)
        res[name] = {"rows": n, "hll": {"ms": round(ms_h, 4), "GB_s": round(gbs, 1),
                                        "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "estimate": est},
                     "exact": {"ms": round(ms_e, 4), "count": exact},
                     "hll_relative_error": round(abs(est - exact) / max(1, exact), 5)}
    ms_s, small = timed(lambda: ctx.count_distinct(table["price"], 0, 99_999), 10)
    res["price_99999_rows_auto_exact"] = {"ms": round(ms_s, 4), "count": small}
    for e in ev:
        e.free()
    return res


def bench_select(ctx, table):
    """SELECT price, qty FROM t WHERE <the Q6 predicates> ORDER BY price DESC LIMIT 100 on
    the first 60M rows: scan -> bitmap, compaction + radix select of the first 100 in
    (cell, scan order), gather of two columns; the same without LIMIT (stable radix sort
    of every selected row); without ORDER BY; ORDER BY ... LIMIT 10 over all 60M rows."""
    import time
    from turbogp_amd import workloads as W
    names, preds, _, _ = W.q6_spec()
    cols = [table[c] for c in names]
    n = SCAN_ROWS

    def wall(fn, reps):
        ts, out = [], None
        for _ in range(reps):
            ctx.sync()
            t0 = time.perf_counter()
            out = fn()
            ctx.sync()
            ts.append((time.perf_counter() - t0) * 1e3)
        return statistics.median(ts), out

    res = {}
    bm, sel = ctx.scan_filter(cols, preds, 0, n)
    try:
        for label, order, limit in (("order_by_price_desc_limit_100", table["price"], 100),
                                    ("order_by_price_desc_all_selected", table["price"], None),
                                    ("scan_order_limit_1000", None, 1000)):
            def q():
                rows, cnt = ctx.select_rows(0, n, bm, order, True, limit)
                try:
                    return [ctx.gather(table[c], rows, cnt) for c in ("price", "qty")]
                finally:
                    rows.free()
            ms, vals = wall(q, 5)
            res[label] = {"ms": round(ms, 3), "selected_rows": sel, "returned": len(vals[0])}
    finally:
        bm.free()

    def full():
        rows, cnt = ctx.select_rows(0, n, None, table["price"], False, 10)
        rows.free()
        return cnt
    ms, _ = wall(full, 3)
    res["order_by_price_all_60M_limit_10"] = {"ms": round(ms, 3), "rows_ordered": n}
    res["note"] = "host wall time per synchronous call sequence, including the scan only once"
    return res


def bench_serve(ctx, table, join_tables, args):
    """Serving: an open-loop Poisson stream of the mixed requests (70 % Q1/Q6
    scan-aggregates on log-uniform ranges, 25 % join-aggregates over build
    prefixes of 1K..2M rows, 5 % full-table Q1), batched server against the
    one-launch-per-query server, at 50/70/90 % of the batched capacity."""
    from turbogp_amd import _capi as A
    from turbogp_amd import workloads as W
    from turbogp_amd.serve import Server
    build, probe = join_tables
    n = args.serve_requests
    mix = W.serve_mix(n, W.BATCH_TABLE_ROWS, W.JOIN_BUILD_ROWS, W.JOIN_PROBE_ROWS)
    reqs = W.serve_requests(table, build, probe, mix)
    configs = {"batched": dict(max_batch_queries=1024, batch_slots=3, join_slots=4,
                               heavy_rows=50_000_000),
               "per_query_launch": dict(batch_slots=3, join_slots=4, heavy_rows=50_000_000,
                                        per_query_launch=True)}

    def run(srv, offsets, rq=reqs):
        ids, start = srv.replay(rq, offsets)
        res = [srv.result(i) for i in ids]
        assert all(t["status"] == A.OK for _, t in res)
        lat = [(t["done_ns"] - t["arrival_ns"]) / 1e6 for _, t in res]
        end = max(t["done_ns"] for _, t in res)
        return lat, (end - start) / 1e9, [r for r, _ in res]

    out = {"requests": n, "mix": {k: sum(1 for m in mix if m[0] == k) for k in
                                  ("q1", "q6", "join", "full")}}
    results = {}
    for name, kw in configs.items():
        srv = Server(ctx, **kw)
        lat, span, res = run(srv, np.zeros(n, np.int64))
        results[name] = res
        out[name] = {"capacity_req_s": round(n / span, 1)}
        srv.close()
    same = all(np.array_equal(a.count, b.count) and np.array_equal(a.sums_words, b.sums_words)
               for a, b in zip(results["batched"], results["per_query_launch"]))
    assert same, "batched and per-query serving results differ"
    cap = out["batched"]["capacity_req_s"]
    for name, kw in configs.items():
        srv = Server(ctx, **kw)
        levels = []
        for frac in (0.5, 0.7, 0.9):
            qps = frac * cap
            lat, span, _ = run(srv, W.poisson_offsets(n, qps))
            levels.append({"offered_req_s": round(qps, 1), "fraction_of_batched_capacity": frac,
                           "achieved_req_s": round(n / span, 1),
                           "p50_ms": round(pct(lat, 50), 3), "p99_ms": round(pct(lat, 99), 3)})
        out[name]["open_loop"] = levels
        out[name]["stats"] = srv.stats()
        srv.close()
    # the batchable part of the mix alone (scan-aggregates): capacity of both forms, then
    # both at the same offered loads (50 / 90 % of the one-launch form's capacity) and
    # the batched form at 50 / 90 % of its own
    scan = [m for m in mix if m[0] in ("q1", "q6")]
    rq = W.serve_requests(table, build, probe, scan)
    scap = {}
    for name, kw in configs.items():
        srv = Server(ctx, **kw)
        _, span, _ = run(srv, np.zeros(len(rq), np.int64), rq)
        scap[name] = len(rq) / span
        out[name]["scan_only_capacity_req_s"] = round(scap[name], 1)
        srv.close()
    equal = []
    for frac in (0.5, 0.9):
        qps = frac * scap["per_query_launch"]
        row = {"offered_req_s": round(qps, 1), "fraction_of_per_query_capacity": frac}
        for name, kw in configs.items():
            srv = Server(ctx, **kw)
            lat, span, _ = run(srv, W.poisson_offsets(len(rq), qps), rq)
            row[name] = {"p50_ms": round(pct(lat, 50), 3), "p99_ms": round(pct(lat, 99), 3)}
            srv.close()
        equal.append(row)
    own = []
    srv = Server(ctx, **configs["batched"])
    for frac in (0.5, 0.9):
        qps = frac * scap["batched"]
        lat, span, _ = run(srv, W.poisson_offsets(len(rq), qps), rq)
        own.append({"offered_req_s": round(qps, 1), "fraction_of_batched_capacity": frac,
                    "achieved_req_s": round(len(rq) / span, 1),
                    "p50_ms": round(pct(lat, 50), 3), "p99_ms": round(pct(lat, 99), 3)})
    srv.close()
    out["scan_only"] = {"requests": len(rq), "equal_load": equal, "batched_open_loop": own}
    out["capacity_ratio"] = round(cap / out["per_query_launch"]["capacity_req_s"], 2)
    out["scan_only_capacity_ratio"] = round(out["batched"]["scan_only_capacity_req_s"] /
                                            out["per_query_launch"]["scan_only_capacity_req_s"], 2)
    out["results_identical_batched_vs_per_query"] = bool(same)
    return out


def calibrate():
    """One 512 MiB write (k_fill_u64) and one 512 MiB read (k_minmax) for the
    FETCH_SIZE / WRITE_SIZE scale factors."""
    from turbogp_amd import Context, DeviceColumn, _capi as A
    ctx = Context(0)
    n = 1 << 26  # 64M int64 words = 512 MiB
    col = DeviceColumn(ctx, n, A.I64)
    col.buf.fill_u64(7)
    lo, hi = col.minmax()
    assert lo == hi == 7
    col.free()
    ctx.close()
    print(json.dumps({"calibrate": {"write_bytes_k_fill_u64": n * 8,
                                    "read_bytes_k_minmax": n * 8}}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--per-query-reps", type=int, default=3)
    ap.add_argument("--cpu-rows", type=int, default=750_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--sections", default="batch,scan,join,serve,mixed,distinct,select",
                    help="which parts to run (profiling runs restrict them so a kernel's "
                         "launches all come from one workload)")
    ap.add_argument("--tuning", default="",
                    help="wg_per_cu,chunk_slices,order,cached_loads,items_per_grab for the batch "
                         "instead of "
                         "the autotuner (profiling runs that must time one configuration)")
    args = ap.parse_args()
    if args.calibrate:
        calibrate()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}): this build runs on "
              "a single MI355X; the multi-GPU sharding is a design, not a measured path.",
              file=sys.stderr)
        sys.exit(2)

    from turbogp_amd import Context, Event, GroupAggBuffers, _capi as A
    from turbogp_amd import workloads as W
    from turbogp_amd.batch import Batch

    ctx = Context(0)
    log(f"device {ctx.info.name.decode()} {ctx.info.arch.decode()} CUs={ctx.info.cu_count}")
    n = W.BATCH_TABLE_ROWS
    table = W.make_fact_table(ctx, n)
    log(f"generated {n} rows on the device")

    # ---- batch of 4096 queries -------------------------------------------------
    params = W.batch_query_params(W.BATCH_QUERIES, n)
    queries = W.batch_queries(table, params)
    batch_bytes = W.batch_bytes(params)
    unique_bytes = W.batch_unique_bytes(params)
    b = Batch(ctx, len(queries))
    b.set_queries(queries)
    info = b.info()
    assert info["fallback"] == 0, info
    if args.tuning:
        wg, chunk, order, cached, grab = (int(x) for x in args.tuning.split(","))
        b.set_tuning(wg, True, chunk, order, bool(cached), grab)
        tuned = b.tuning()
        log(f"tuning given: {tuned}")
    else:
        tuned, tune_ms = b.autotune()
        log(f"autotuned {tuned} ({tune_ms:.3f} ms)")
    ev = [Event(ctx) for _ in range(2)]
    for _ in range(args.warmup):
        b.run()
    ctx.sync()
    step_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ev[0].record()
        b.run()
        ev[1].record()
        step_ms.append(ev[0].elapsed_ms(ev[1]))
    ctx.sync()
    wall = time.perf_counter() - t0
    batch_results = b.results()
    ms_batch = statistics.mean(step_ms)
    log(f"batch: {args.steps} steps, mean {ms_batch:.3f} ms (wall {wall:.3f} s)")

    # the one-launch-per-query form of the same queries
    pq_tot, pq_lat, same, ms_pq = [], None, None, None
    for _ in range(args.per_query_reps):
        tot, per = b.run_timed(A.BATCH_PER_QUERY)
        pq_tot.append(tot)
        pq_lat = per if pq_lat is None or tot <= min(pq_tot) else pq_lat
    if pq_tot:
        per_query_results = b.results()
        same = all(np.array_equal(x.count, y.count) and
                   np.array_equal(x.sums_words, y.sums_words) and
                   x.n_out_of_range == y.n_out_of_range
                   for x, y in zip(batch_results, per_query_results))
        assert same, "batched and per-query results differ"
        ms_pq = min(pq_tot)
        log(f"per-query: {ms_pq:.3f} ms; speedup {ms_pq / ms_batch:.2f}x")
    b.close()

    # ---- 60M-row scan-aggregate queries (Q1 and Q6 shapes) ------------------------
    qa = W.build_query(table, W.q1_spec(), 0, SCAN_ROWS)
    qb = W.build_query(table, W.q6_spec(), 0, SCAN_ROWS)
    ra = GroupAggBuffers(ctx, 6, 4)
    rb = GroupAggBuffers(ctx, 1, 1)
    ev3 = [Event(ctx) for _ in range(3)]
    for _ in range(args.warmup):
        ctx.group_agg_async(qa, ra)
        ctx.group_agg_async(qb, rb)
    ctx.sync()
    ta, tb = [], []
    for _ in range(args.steps):
        ev3[0].record()
        ctx.group_agg_async(qa, ra)
        ev3[1].record()
        ctx.group_agg_async(qb, rb)
        ev3[2].record()
        ta.append(ev3[0].elapsed_ms(ev3[1]))
        tb.append(ev3[1].elapsed_ms(ev3[2]))
    res_a = ra.fetch()
    sel_a = int(res_a.count.sum())
    assert 0.95 * SCAN_ROWS < sel_a < 0.975 * SCAN_ROWS and res_a.n_out_of_range == 0, sel_a
    res_b = rb.fetch()
    ms_a, ms_b = statistics.mean(ta), statistics.mean(tb)
    gbs_a = SCAN_ROWS * W.Q1_BYTES_PER_ROW / 1e9 / (ms_a / 1e3)
    gbs_b = SCAN_ROWS * W.Q6_BYTES_PER_ROW / 1e9 / (ms_b / 1e3)
    for e in ev + ev3:
        e.free()

    generic = bench_generic(ctx, table, args)
    sections = set(args.sections.split(","))
    predicates = bench_predicates(ctx, table, args) if "scan" in sections else None
    join = bench_join(ctx, args) if "join" in sections or "serve" in sections else None
    join_tables = join.pop("_tables") if join else None
    serve = bench_serve(ctx, table, join_tables, args) if "serve" in sections else None
    mixed = bench_mixed(ctx, table) if "mixed" in sections else None
    distinct = bench_distinct(ctx, table) if "distinct" in sections else None
    select = bench_select(ctx, table) if "select" in sections else None
    if "join" not in sections:
        join = None

    achieved = batch_bytes / 1e9 / (ms_batch / 1e3)
    # roofline of the batch kernel: the bytes the batch must move at least (the union of
    # the cells its queries read, each once) over the launch time; per-query bytes
    # count a row once per query that covers it and so exceed what the kernel reads
    achieved_unique = unique_bytes / 1e9 / (ms_batch / 1e3)
    frac = achieved_unique / HBM_PEAK_GBS
    assert frac <= 1.0, f"roofline fraction {frac} above 1: the byte count is wrong"
    traffic = measured_traffic("k_batch_4096q_256M", n)
    nq = len(queries)
    out = {
        "metric": METRIC,
        "value": round(nq / (ms_batch / 1e3), 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_batch, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64/int128",
        "data": "synthetic",
        "config": {"workload": "4096 filter+group-by queries over a 256M-row resident fact "
                               "table (ranges log-uniform 1K..4M rows, uniform date cutoffs, "
                               "group keys from {flagA, flagB, flagA x flagB, date/30}), one "
                               "persistent-kernel hipGraph launch per batch",
                   "table_rows": n, "queries": nq, "batch_rows": int(sum(e - s for s, e, _, _ in params)),
                   "algorithmic_bytes_per_batch": int(batch_bytes), "tuning": tuned},
        "batch": {"ms": round(ms_batch, 4),
                  "per_query_GB_s": round(achieved, 1),
                  "unique_bytes": int(unique_bytes),
                  "unique_GB_s": round(unique_bytes / 1e9 / (ms_batch / 1e3), 1),
                  "unique_hbm_frac": round(unique_bytes / 1e9 / (ms_batch / 1e3) / HBM_PEAK_GBS, 4),
                  "p50_latency_ms": round(pct(step_ms, 50), 4),
                  "p99_latency_ms": round(pct(step_ms, 99), 4)},
        "per_query_launches": None if ms_pq is None else {
            "ms": round(ms_pq, 4), "queries_per_s": round(nq / (ms_pq / 1e3), 1),
            "per_query_GB_s": round(batch_bytes / 1e9 / (ms_pq / 1e3), 1),
            "p50_latency_ms": round(pct(pq_lat, 50), 4),
            "p99_latency_ms": round(pct(pq_lat, 99), 4),
            "latency_note": "completion time of each query after the 4096 launches were "
                            "queued together; batched queries all complete at the batch end"},
        "speedup_vs_per_query": None if ms_pq is None else round(ms_pq / ms_batch, 2),
        "results_identical_batched_vs_per_query": same,
        "scan_60M": {
            "q1_groupby": {"ms": round(ms_a, 4), "GB_s": round(gbs_a, 1),
                           "hbm_frac": round(gbs_a / HBM_PEAK_GBS, 4),
                           "bytes_per_row": W.Q1_BYTES_PER_ROW},
            "q6_scan_sum": {"ms": round(ms_b, 4), "GB_s": round(gbs_b, 1),
                            "hbm_frac": round(gbs_b / HBM_PEAK_GBS, 4),
                            "bytes_per_row": W.Q6_BYTES_PER_ROW,
                            "selected": int(res_b.count[0])},
            **generic, **(predicates or {})},
        "join_2M_x_60M": join,
        "count_distinct_60M": distinct,
        "select_60M": select,
        "batch_mixed_4096q_256M": mixed,
        "serving": serve,
        "roofline": {"bound": "hbm", "achieved": round(achieved_unique, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(frac, 4), "traffic": traffic,
                     "bytes_per_launch": int(unique_bytes),
                     "traffic_over_bytes": None if traffic is None else
                     round(traffic / unique_bytes, 3),
                     "per_query_bytes_per_launch": int(batch_bytes),
                     "kernel": "persistent batch kernel (k_batch_shared, the shared scan, or "
                               "k_batch_items / k_batch, as autotuned): the cells the batch reads, "
                               "each once (per column, the union of the row ranges of the "
                               "queries that read it, x its width) / mean launch time (hipEvents "
                               "on the batch stream, graph replay incl. the result-zeroing "
                               "fill); traffic = HBM bytes per launch from FETCH_SIZE / "
                               "WRITE_SIZE; per_query_bytes counts a row once per query that "
                               "covers it"},
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(params, args.cpu_rows, join_tables)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
