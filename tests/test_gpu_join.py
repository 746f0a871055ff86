"""GPU parity of the hash join (build / probe count / marks / pairs / fused
probe-aggregate) against the oracle and the reference's known answers."""
import numpy as np
import pytest

from oracle import binding as B
from turbogp_amd import DeviceBuffer, DeviceColumn, HashTable, _capi as A
from turbogp_amd import workloads as W

pytestmark = pytest.mark.gpu
I64MIN = np.iinfo(np.int64).min


def dev(ctx, arr):
    return DeviceColumn.from_numpy(ctx, np.ascontiguousarray(arr))


def build(ctx, keys, payload=None):
    kc = dev(ctx, np.array(keys, np.int64))
    pc = dev(ctx, np.array(payload, np.int32)) if payload is not None else None
    t = HashTable(ctx, kc, pc)
    return t, [kc] + ([pc] if pc else [])


def pairs(ctx, bkeys, pkeys):
    t, keep = build(ctx, bkeys)
    pc = dev(ctx, np.array(pkeys, np.int64))
    try:
        pi, pp = t.probe_pairs(pc)
        return list(zip(pi.tolist(), pp.tolist())), t.probe_count(pc)
    finally:
        t.free()
        for c in keep + [pc]:
            c.free()


def test_basic_insert_probe(ctx):  # join_hash_table.rs test_basic_insert_probe
    build_keys = [i * 7 for i in range(100)]
    p, n = pairs(ctx, build_keys, build_keys + [1, 999])
    assert n == 100 and p == [(i, i) for i in range(100)]


def test_duplicate_keys(ctx):  # join_hash_table.rs test_duplicate_keys
    p, n = pairs(ctx, [5, 42, 42, 42], [42])
    assert n == 3 and p == [(0, 1), (0, 2), (0, 3)]


def test_hash_kernel_kats(ctx):  # kernel/hash.rs:231-285
    p, _ = pairs(ctx, [1, 2, 3, 2, 4, 2, 5], [2, 99])
    assert [b for _, b in p] == [1, 3, 5]
    _, n = pairs(ctx, [1, 2, 3, 2, 4], [2, 4, 99])
    assert n == 3


def test_semi_anti_marks(ctx):  # exec/join.rs test_semi_join / test_anti_join
    t, keep = build(ctx, [2, 4])
    left = dev(ctx, np.array([1, 2, 3, 4], np.int64))
    bm = t.probe_mark(left)
    anti = bm.not_()
    assert list(bm.iter_set_bits()) == [1, 3] and list(anti.iter_set_bits()) == [0, 2]
    for x in [bm, anti, left] + keep:
        x.free()
    t.free()


def test_large_load_and_info(ctx):  # join_hash_table.rs test_large_load
    n = 100_000
    t, keep = build(ctx, np.arange(n))
    nr, cap, dup = t.info()
    assert nr == n and cap == 262_144 and not dup
    assert t.footprint() == A.lib().tgp_hash_table_bytes_for(n)
    pc = dev(ctx, np.arange(n + 1000, dtype=np.int64))
    pi, pp = t.probe_pairs(pc)
    assert np.array_equal(pi, np.arange(n)) and np.array_equal(pp, np.arange(n))
    for x in [pc] + keep:
        x.free()
    t.free()


@pytest.mark.parametrize("seed", range(3))
def test_random_pairs_vs_oracle(ctx, seed):
    rng = np.random.default_rng(seed)
    nb, npb = int(rng.integers(1, 5000)), int(rng.integers(1, 50000))
    bk = rng.integers(-300, 3000, nb).astype(np.int64)
    bk[rng.integers(0, nb, 3)] = I64MIN
    pk = rng.integers(-400, 3100, npb).astype(np.int64)
    pk[::53] = I64MIN
    got, _ = pairs(ctx, bk, pk)
    oi, ob, n = B.join_pairs(bk, pk)
    assert got == list(zip(oi.tolist(), ob.tolist()))


def test_build_into_caller_memory_reused(ctx):
    """Tables built one after another in the same caller buffer (largest
    first, then smaller ones over its stale slots) answer as freshly
    allocated ones do; a buffer too small is an error."""
    rng = np.random.default_rng(11)
    mem = DeviceBuffer(ctx, A.lib().tgp_hash_table_bytes_for(20_000))
    try:
        for nb in (20_000, 3_000, 17):
            bk = rng.integers(-500, 9_000, nb).astype(np.int64)
            bk[rng.integers(0, nb, 2)] = I64MIN  # sentinel keys take the side list
            pk = rng.integers(-600, 9_100, 40_000).astype(np.int64)
            pk[::97] = I64MIN
            kc, pc = dev(ctx, bk), dev(ctx, pk)
            t = HashTable(ctx, kc, memory=mem)
            try:
                assert t.info()[0] == nb
                pi, pp = t.probe_pairs(pc)
                oi, ob, n = B.join_pairs(bk, pk)
                assert list(zip(pi.tolist(), pp.tolist())) == list(zip(oi.tolist(), ob.tolist()))
                assert t.probe_count(pc) == n
            finally:
                t.free()
                kc.free()
                pc.free()
        kc = dev(ctx, np.arange(40_000, dtype=np.int64))  # 131072 slots, not 65536
        try:
            with pytest.raises(RuntimeError, match="bytes given"):
                HashTable(ctx, kc, memory=mem)
        finally:
            kc.free()
    finally:
        mem.free()


@pytest.mark.parametrize("group_by", [A.JOIN_GROUP_BUILD_PAYLOAD, A.JOIN_GROUP_PROBE_COL,
                                      A.JOIN_GROUP_NONE])
@pytest.mark.parametrize("dups", [False, True])
def test_probe_agg_vs_oracle(ctx, group_by, dups):
    rng = np.random.default_rng(10 * group_by + dups)
    nb, npb = 20_000, 300_000
    bk = (rng.integers(0, nb // 2, nb) if dups else rng.permutation(nb) * 7919).astype(np.int64)
    bk[5] = I64MIN
    bp = rng.integers(-1, 26, nb).astype(np.int32)   # -1 and 25 fall outside [0, 25)
    pk = bk[rng.integers(0, nb, npb)]
    pk[::11] += 3                                     # some misses
    pv = rng.integers(-2**62, 2**62, npb).astype(np.int64)
    pg = rng.integers(0, 70_000, npb).astype(np.int32)
    ng = {A.JOIN_GROUP_BUILD_PAYLOAD: 25, A.JOIN_GROUP_PROBE_COL: 65_000,
          A.JOIN_GROUP_NONE: 1}[group_by]
    kb = 0
    b, e = 123, npb - 77
    t, keep = build(ctx, bk, bp)
    cols = [dev(ctx, pk), dev(ctx, pv), dev(ctx, pg)]
    try:
        got = t.probe_agg(cols[0], cols[1], group_by, ng, kb, cols[2], b, e)
    finally:
        t.free()
        for c in keep + cols:
            c.free()
    cnt, sums, oor = B.join_agg(bk, bp, pk, pv, pg, group_by, kb, ng, b, e)
    assert got.n_out_of_range == oor
    assert np.array_equal(got.count, cnt)
    assert [s[0] for s in got.sums()] == sums


def test_c3_slice_vs_oracle(ctx):
    """The join workload's shape at 10K build / 100K probe rows (Zipf 1.1 keys)."""
    nb, npb = 10_000, 100_000
    bkey, _ = W.join_build_side(nb)
    pay = B.synth_uniform(A.I32, W.SEED, W.COL_PAY, 0, 24, 0, nb)
    cdf = W.zipf_cdf(nb)
    ktab = W.zipf_key_table(bkey)
    fk = B.synth_cdf(A.I64, W.SEED, W.COL_FK, cdf, ktab, 0, npb)
    val = B.synth_uniform(A.I64, W.SEED, W.COL_VAL, 90_000, 10_500_000, 0, npb)
    gkey = B.synth_uniform(A.I32, W.SEED, W.COL_GKEY, 0, 14_999, 0, npb)
    t, keep = build(ctx, bkey, pay)
    cols = [dev(ctx, fk), dev(ctx, val), dev(ctx, gkey)]
    try:
        r1 = t.probe_agg(cols[0], cols[1], A.JOIN_GROUP_BUILD_PAYLOAD, 25)
        r2 = t.probe_agg(cols[0], cols[1], A.JOIN_GROUP_PROBE_COL, 15_000, 0, cols[2])
    finally:
        t.free()
        for c in keep + cols:
            c.free()
    c1, s1, _ = B.join_agg(bkey, pay, fk, val, gkey, 0, 0, 25)
    c2, s2, _ = B.join_agg(bkey, pay, fk, val, gkey, 1, 0, 15_000)
    assert np.array_equal(r1.count, c1) and [s[0] for s in r1.sums()] == s1
    assert np.array_equal(r2.count, c2) and [s[0] for s in r2.sums()] == s2
    assert int(r1.count.sum()) == npb  # every foreign key matches exactly one build row


def test_device_zipf_generator_matches_oracle(ctx):
    nb, npb = 5000, 200_000
    bkey, _ = W.join_build_side(nb)
    cdf = W.zipf_cdf(nb)
    ktab = W.zipf_key_table(bkey)
    want = B.synth_cdf(A.I64, W.SEED, W.COL_FK, cdf, ktab, 0, npb)
    got = W.device_lookup_cdf(ctx, npb, A.I64, W.COL_FK, cdf, ktab)
    try:
        assert np.array_equal(got.to_numpy(), want)
    finally:
        got.free()


@pytest.mark.parametrize("group_by,ng", [(A.JOIN_GROUP_PROBE_COL, 20), (A.JOIN_GROUP_PROBE_COL, 1_000),
                                         (A.JOIN_GROUP_PROBE_COL, 2_047), (A.JOIN_GROUP_PROBE_COL, 2_049),
                                         (A.JOIN_GROUP_PROBE_COL, 4_095), (A.JOIN_GROUP_PROBE_COL, 4_097),
                                         (A.JOIN_GROUP_PROBE_COL, 300_001),
                                         (A.JOIN_GROUP_PROBE_COL, 2_000_001),
                                         (A.JOIN_GROUP_BUILD_PAYLOAD, 700),
                                         (A.JOIN_GROUP_BUILD_PAYLOAD, 5_000)])
def test_probe_agg_every_table_form(ctx, group_by, ng):
    """Per-wave replicated LDS tables (few groups), one workgroup table, the
    radix-partitioned form and global atomics must all match the oracle."""
    rng = np.random.default_rng(ng)
    nb, npb = 30_000, 400_000
    bk = rng.integers(0, nb // 2, nb).astype(np.int64)      # duplicate keys
    bp = rng.integers(-2, ng + 2, nb).astype(np.int32)
    pk = bk[rng.integers(0, nb, npb)]
    pk[::13] -= 1_000_000                                  # misses
    pv = rng.integers(-2**62, 2**62, npb).astype(np.int64)
    pg = rng.integers(-5, ng + 5, npb).astype(np.int32)    # some groups out of range
    kb = 3 if group_by == A.JOIN_GROUP_PROBE_COL else 0
    b, e = 17, npb - 1
    t, keep = build(ctx, bk, bp)
    cols = [dev(ctx, pk), dev(ctx, pv), dev(ctx, pg)]
    got = []
    try:
        for atomics in (0, 1):
            ctx.tune(A.TUNE_JOIN_ATOMICS, atomics)
            got.append(t.probe_agg(cols[0], cols[1], group_by, ng, kb, cols[2], b, e))
    finally:
        ctx.tune(A.TUNE_JOIN_ATOMICS, 0)
        t.free()
        for c in keep + cols:
            c.free()
    cnt, sums, oor = B.join_agg(bk, bp, pk, pv, pg, group_by, kb, ng, b, e)
    for r in got:
        assert r.n_out_of_range == oor
        assert np.array_equal(r.count, cnt)
        assert [s[0] for s in r.sums()] == sums


def test_exec_join_kats_device(ctx):
    """exec/join.rs:294-353 through the device: inner, left, multi-key, semi,
    anti and no-match joins (build on the right table, probe with the left)."""
    def cols(*arrs):
        return [dev(ctx, np.array(a, np.int64)) for a in arrs]

    keep = []
    try:
        (r,), (l,) = cols([2, 3, 4]), cols([1, 2, 3])
        keep += [r, l]
        t = HashTable(ctx, r)
        p, b = t.join_pairs(l)
        assert list(p) == [1, 2] and list(b) == [0, 1]                       # test_inner_join
        t.free()
        (r2,) = cols([2])
        keep.append(r2)
        t = HashTable(ctx, r2)
        p, b = t.join_pairs(l, A.JOIN_LEFT)                                  # test_left_join
        assert list(p) == [0, 1, 2] and list(b) == [A.JOIN_NO_MATCH, 0, A.JOIN_NO_MATCH]
        t.free()
        ra, rb, la, lb = cols([1, 2], [2, 1], [1, 1, 2], [1, 2, 1])          # test_multi_key_join
        keep += [ra, rb, la, lb]
        t = HashTable(ctx, [ra, rb])
        p, b = t.join_pairs([la, lb])
        assert list(p) == [1, 2] and list(b) == [0, 1]
        bm = t.join_mark([la, lb])
        assert list(bm.iter_set_bits()) == [1, 2]
        bm.free()
        t.free()
        rs, ls = cols([2, 4], [1, 2, 3, 4])                                  # test_semi / anti
        keep += [rs, ls]
        t = HashTable(ctx, rs)
        bm = t.join_mark(ls)
        nb = bm.not_()
        assert [int(x) + 1 for x in bm.iter_set_bits()] == [2, 4]
        assert [int(x) + 1 for x in nb.iter_set_bits()] == [1, 3]
        bm.free()
        nb.free()
        t.free()
        rn, ln = cols([3, 4], [1, 2])                                        # test_no_match
        keep += [rn, ln]
        t = HashTable(ctx, rn)
        assert len(t.join_pairs(ln)[0]) == 0
        t.free()
    finally:
        for c in keep:
            c.free()


@pytest.mark.parametrize("nkeys,pack_bits", [(1, 0), (2, 0), (2, 3), (3, 0), (3, 5), (4, 1)])
@pytest.mark.parametrize("join_type", [A.JOIN_INNER, A.JOIN_LEFT])
def test_multi_key_join_vs_oracle(ctx, nkeys, pack_bits, join_type):
    """Multi-column INNER / LEFT pairs and semi-join marks against the oracle,
    with mixed key dtypes, duplicate build keys, misses, and (pack_bits > 0) the
    packed key cut to a few bits so that distinct keys collide and only the
    full-key check tells them apart."""
    rng = np.random.default_rng(100 * nkeys + 10 * pack_bits + join_type)
    nb, npb = 6000, 40_000
    dtypes = [np.int64, np.int32, np.uint8, np.uint64][:nkeys]
    bcols = [rng.integers(0, 6, nb).astype(dt) for dt in dtypes]
    bcols[0][::7] = rng.integers(-50, 50, len(bcols[0][::7]))               # negative cells
    pick = rng.integers(0, nb, npb)
    pcols = [c[pick].copy() for c in bcols]
    pcols[-1][::5] = pcols[-1][::5] + 1                                     # some misses
    if nkeys >= 2:
        pcols[1] = pcols[1].astype(np.int64)    # int32 build key against an int64 probe key
    bdev = [dev(ctx, c) for c in bcols]
    pdev = [dev(ctx, c) for c in pcols]
    ctx.tune(A.TUNE_JOIN_PACK_BITS, pack_bits)
    try:
        t = HashTable(ctx, bdev if nkeys > 1 else bdev[0])
    finally:
        ctx.tune(A.TUNE_JOIN_PACK_BITS, 0)
    try:
        b, e = 13, npb - 7
        p, pay = t.join_pairs(pdev if nkeys > 1 else pdev[0], join_type, b, e)
        wp, wb = B.hash_join(bcols, pcols, join_type, b, e)
        assert np.array_equal(p, wp) and np.array_equal(pay, wb), (len(p), len(wp))
        bm = t.join_mark(pdev if nkeys > 1 else pdev[0], b, e)
        try:
            want = np.unique(wp[wb != A.JOIN_NO_MATCH]).astype(np.int64) - b
            assert np.array_equal(bm.iter_set_bits().astype(np.int64), want)
        finally:
            bm.free()
    finally:
        t.free()
        for c in bdev + pdev:
            c.free()


def test_multi_key_table_rejects_single_key_calls(ctx):
    a, b2 = dev(ctx, np.arange(10, dtype=np.int64)), dev(ctx, np.arange(10, dtype=np.int64))
    t = HashTable(ctx, [a, b2])
    try:
        with pytest.raises(RuntimeError, match="key columns"):
            t.probe_count(a)
        with pytest.raises(RuntimeError, match="key columns"):
            t.join_pairs(a)
        with pytest.raises(RuntimeError, match="key columns"):
            t.probe_agg(a, a, A.JOIN_GROUP_NONE)
    finally:
        t.free()
        a.free()
        b2.free()


@pytest.mark.parametrize("ngroups", [25, 3000, 9000])
def test_probe_agg_duplicate_heavy_build_vs_oracle(ctx, ngroups):
    """A build side with many rows per key gives probe rows many matches each, so the
    additions per accumulator word are not bounded by the probe rows: such tables take
    the carry-exact int128 form, in LDS (25 groups) and in the global table (3000 and
    9000 payload groups), and equal the oracle."""
    rng = np.random.default_rng(ngroups)
    nb, npb = 40_000, 200_000
    bk = rng.integers(0, 300, nb).astype(np.int64)       # ~133 build rows per key
    bp = rng.integers(0, ngroups, nb).astype(np.int32)
    pk = bk[rng.integers(0, nb, npb)]
    pk[::7] += 1000                                       # misses
    pv = rng.integers(-2**62, 2**62, npb).astype(np.int64)
    t, keep = build(ctx, bk, bp)
    cols = [dev(ctx, pk), dev(ctx, pv)]
    try:
        assert t.info()[2] == 1  # has duplicates
        got = t.probe_agg(cols[0], cols[1], A.JOIN_GROUP_BUILD_PAYLOAD, ngroups, 0)
    finally:
        t.free()
        for c in keep + cols:
            c.free()
    cnt, sums, oor = B.join_agg(bk, bp, pk, pv, None, A.JOIN_GROUP_BUILD_PAYLOAD, 0, ngroups)
    assert got.n_out_of_range == oor == 0
    assert np.array_equal(got.count, cnt)
    assert [s[0] for s in got.sums()] == sums


def test_probe_agg_past_2_32_additions_per_word(ctx):
    """One payload group gets more than 2^32 additions of 2^32 - 1 (4352 build rows of
    one key x 2^20 probe rows): a split low-word sum would wrap past 2^64; the exact
    form gives the closed form count x value.  3000 other payload groups keep the
    table out of LDS."""
    dup, npb = 4352, 1 << 20
    bk = np.concatenate([np.full(dup, 77, np.int64), np.arange(1000, 4000, dtype=np.int64)])
    bp = np.concatenate([np.zeros(dup, np.int32), np.arange(1, 3001, dtype=np.int32)])
    pk = np.full(npb, 77, np.int64)
    v = (1 << 32) - 1
    pv = np.full(npb, v, np.int64)
    t, keep = build(ctx, bk, bp)
    cols = [dev(ctx, pk), dev(ctx, pv)]
    try:
        got = t.probe_agg(cols[0], cols[1], A.JOIN_GROUP_BUILD_PAYLOAD, 3001, 0)
    finally:
        t.free()
        for c in keep + cols:
            c.free()
    m = dup * npb
    assert m > 2 ** 32
    assert int(got.count[0]) == m and int(got.count[1:].sum()) == 0
    assert got.sums()[0][0] == m * v


@pytest.mark.parametrize("group_by,ngroups", [(A.JOIN_GROUP_BUILD_PAYLOAD, 25),
                                              (A.JOIN_GROUP_BUILD_PAYLOAD, 4000),
                                              (A.JOIN_GROUP_PROBE_COL, 70_000),
                                              (A.JOIN_GROUP_NONE, 1)])
@pytest.mark.parametrize("dups", [False, True])
def test_probe_agg_over_a_build_prefix(ctx, group_by, ngroups, dups):
    """build_rows = nb joins with the first nb build rows of a table built over all of
    them: equal to a table built on the prefix alone (LDS, global and partitioned
    forms), with unique and with duplicate build keys, and to the oracle."""
    rng = np.random.default_rng(40 + ngroups + dups)
    nb_all, npb = 30_000, 400_000
    bk = (rng.integers(0, nb_all // 3, nb_all) if dups else rng.permutation(nb_all) * 13 + 5)
    bk = bk.astype(np.int64)
    bp = rng.integers(0, min(ngroups, 4000), nb_all).astype(np.int32)
    pk = bk[rng.integers(0, nb_all, npb)]
    pk[::9] -= 1                                   # misses
    pv = rng.integers(-2**40, 2**40, npb).astype(np.int64)
    pg = rng.integers(0, ngroups, npb).astype(np.int32)
    full, keep = build(ctx, bk, bp)
    cols = [dev(ctx, pk), dev(ctx, pv), dev(ctx, pg)]
    try:
        for nb in (1, 777, 12_345, nb_all):
            part, keep2 = build(ctx, bk[:nb], bp[:nb])
            try:
                want = part.probe_agg(cols[0], cols[1], group_by, ngroups, 0, cols[2])
            finally:
                part.free()
                for c in keep2:
                    c.free()
            got = full.probe_agg(cols[0], cols[1], group_by, ngroups, 0, cols[2],
                                 build_rows=nb)
            assert np.array_equal(got.count, want.count), nb
            assert got.sums() == want.sums() and got.n_out_of_range == want.n_out_of_range, nb
            cnt, sums, oor = B.join_agg(bk[:nb], bp[:nb], pk, pv, pg, group_by, 0, ngroups)
            assert np.array_equal(got.count, cnt) and [x[0] for x in got.sums()] == sums, nb
    finally:
        full.free()
        for c in keep + cols:
            c.free()


def test_probe_agg_batch_equals_single_calls(ctx):
    """tgp_hash_probe_agg_batch: queries over two tables (unique and duplicate build
    keys), every grouping form, row ranges from empty through one row to ranges
    over many 16K-row work items, build prefixes, groups out of range, and queries
    of 4000 groups that run alone within the call, all equal one
    tgp_hash_probe_agg call each, bit for bit, and a sample equals the oracle."""
    rng = np.random.default_rng(77)
    nb, npb = 50_000, 300_007
    bk_u = (rng.permutation(nb) * 7 + 3).astype(np.int64)
    bk_d = rng.integers(0, nb // 4, nb).astype(np.int64)
    bp = rng.integers(-3, 40, nb).astype(np.int32)
    pk = np.concatenate([bk_u, bk_d])[rng.integers(0, 2 * nb, npb)]
    pk[::11] += 1
    pv = rng.integers(-2**45, 2**45, npb).astype(np.int64)
    pg = rng.integers(-5, 120, npb).astype(np.int32)
    tu, keep_u = build(ctx, bk_u, bp)
    td, keep_d = build(ctx, bk_d, bp)
    cols = [dev(ctx, pk), dev(ctx, pv), dev(ctx, pg)]
    specs = []
    forms = [(A.JOIN_GROUP_BUILD_PAYLOAD, 25, 0), (A.JOIN_GROUP_BUILD_PAYLOAD, 40, -3),
             (A.JOIN_GROUP_PROBE_COL, 100, 0), (A.JOIN_GROUP_NONE, 1, 0),
             (A.JOIN_GROUP_PROBE_COL, 4000, -5)]
    ranges = [(0, npb), (0, 0), (17, 18), (16_384 - 5, 3 * 16_384 + 9), (1000, 250_000),
              (npb - 1, npb)]
    for t, tk in ((tu, bk_u), (td, bk_d)):
        for gb, ng, kb in forms:
            for b, e in ranges:
                lim = int(rng.choice([0, 1, 999, nb]))
                specs.append((t, tk, gb, ng, kb, b, e, lim))
    try:
        qs = [t.agg_query(cols[0], cols[1], gb, ng, kb, cols[2], b, e, lim)
              for t, _, gb, ng, kb, b, e, lim in specs]
        got = ctx.probe_agg_batch(qs)
        again = ctx.probe_agg_batch(qs[::-1])[::-1]
        for i, (t, tk, gb, ng, kb, b, e, lim) in enumerate(specs):
            want = t.probe_agg(cols[0], cols[1], gb, ng, kb, cols[2], b, e, build_rows=lim)
            for g in (got[i], again[i]):
                assert np.array_equal(g.count, want.count), i
                assert g.sums() == want.sums() and g.n_out_of_range == want.n_out_of_range, i
            if i % 7 == 0:
                n = lim or nb
                cnt, sums, oor = B.join_agg(tk[:n], bp[:n], pk, pv, pg, gb, kb, ng, b, e)
                assert np.array_equal(got[i].count, cnt) and \
                    [x[0] for x in got[i].sums()] == sums and got[i].n_out_of_range == oor, i
        assert ctx.probe_agg_batch([]) == []
    finally:
        for t in (tu, td):
            t.free()
        for c in keep_u + keep_d + cols:
            c.free()


@pytest.mark.parametrize("dups", [False, True])
def test_bloom_filter_changes_no_result(ctx, dups):
    """A blocked bloom filter in front of the table (tgp_hash_table_add_bloom): every
    build key passes it (no false negatives: probing the build keys finds them all),
    and counts, marks, pairs and every probe aggregation (alone and batched, with
    build prefixes) over probes that mostly miss equal those of the unfiltered table;
    the INT64_MIN key lives beside the table and is still found."""
    rng = np.random.default_rng(91 + dups)
    nb, npb = 40_000, 500_003
    bk = (rng.integers(-2**62, 2**62, nb // 4)[rng.integers(0, nb // 4, nb)] if dups
          else rng.permutation(nb).astype(np.int64) * 7919 - 2**40)
    bk = bk.astype(np.int64)
    bk[5] = I64MIN
    bp = rng.integers(0, 25, nb).astype(np.int32)
    pk = rng.integers(-2**63, 2**63 - 1, npb, dtype=np.int64)
    hit = rng.random(npb) < 0.1
    pk[hit] = bk[rng.integers(0, nb, int(hit.sum()))]
    pk[::1001] = I64MIN
    pv = rng.integers(-2**40, 2**40, npb).astype(np.int64)
    pg = rng.integers(0, 300, npb).astype(np.int32)
    plain, keep1 = build(ctx, bk, bp)
    filt, keep2 = build(ctx, bk, bp)
    cols = [dev(ctx, pk), dev(ctx, pv), dev(ctx, pg)]
    bkc = dev(ctx, bk)
    try:
        filt.add_bloom(8)
        assert filt.probe_count(bkc) == plain.probe_count(bkc) >= nb
        assert plain.probe_count(cols[0]) == filt.probe_count(cols[0])
        a, b = plain.join_mark(cols[0]), filt.join_mark(cols[0])
        try:
            assert np.array_equal(a.words(), b.words())
        finally:
            a.free()
            b.free()
        pa, pb = plain.probe_pairs(cols[0]), filt.probe_pairs(cols[0])
        assert all(np.array_equal(x, y) for x, y in zip(pa, pb))
        forms = [(A.JOIN_GROUP_BUILD_PAYLOAD, 25), (A.JOIN_GROUP_PROBE_COL, 300),
                 (A.JOIN_GROUP_NONE, 1)]
        qs = []
        for gb, ng in forms:
            for lim in (0, 777):
                want = plain.probe_agg(cols[0], cols[1], gb, ng, 0, cols[2], build_rows=lim)
                got = filt.probe_agg(cols[0], cols[1], gb, ng, 0, cols[2], build_rows=lim)
                assert np.array_equal(got.count, want.count) and got.sums() == want.sums()
                assert got.n_out_of_range == want.n_out_of_range
                qs.append((filt.agg_query(cols[0], cols[1], gb, ng, 0, cols[2], 0, None, lim), want))
        for (q, want), got in zip(qs, ctx.probe_agg_batch([q for q, _ in qs])):
            assert np.array_equal(got.count, want.count) and got.sums() == want.sums()
        cnt, sums, oor = B.join_agg(bk, bp, pk, pv, pg, A.JOIN_GROUP_BUILD_PAYLOAD, 0, 25)
        got = filt.probe_agg(cols[0], cols[1], A.JOIN_GROUP_BUILD_PAYLOAD, 25)
        assert np.array_equal(got.count, cnt) and [x[0] for x in got.sums()] == sums
    finally:
        for t in (plain, filt):
            t.free()
        for c in keep1 + keep2 + cols + [bkc]:
            c.free()


@pytest.mark.parametrize("bloom", [False, True])
def test_hot_key_cache_changes_no_result(ctx, bloom):
    """The hot-key image (TGP_TUNE_JOIN_HOT = 2 forces it on small probes; the match
    count and the match bits of the radix-partitioned GROUP BY use it): skewed
    foreign keys, hot keys absent from the table, a hot INT64_MIN key (never
    cached: it lives beside the table), build prefixes, probe ranges shorter than
    the sample, and every probe-aggregate form (per-wave LDS tables, one workgroup
    table, the radix-partitioned form over match bits, global atomics) give the
    same results with the image as without it, and equal the oracle."""
    rng = np.random.default_rng(5 + bloom)
    nb, npb = 60_000, 700_001
    bk = (rng.permutation(nb).astype(np.int64) * 104_729 - 2**41)
    bk[11] = I64MIN
    bp = rng.integers(-2, 4_003, nb).astype(np.int32)
    rank = np.arange(1, nb + 1, dtype=np.float64)
    w = rank ** -1.1
    pk = bk[rng.choice(nb, npb, p=w / w.sum())]
    absent = np.array([7, -9, 2**60 + 1], np.int64)          # hot keys with no build row
    sel = rng.random(npb) < 0.06
    pk[sel] = absent[rng.integers(0, 3, int(sel.sum()))]
    pk[rng.random(npb) < 0.02] = I64MIN
    pv = rng.integers(-2**50, 2**50, npb).astype(np.int64)
    pg = rng.integers(-3, 70_000, npb).astype(np.int32)
    t, keep = build(ctx, bk, bp)
    cols = [dev(ctx, pk), dev(ctx, pv), dev(ctx, pg)]
    forms = [(A.JOIN_GROUP_BUILD_PAYLOAD, 25, 0, 0), (A.JOIN_GROUP_BUILD_PAYLOAD, 4_000, 0, 0),
             (A.JOIN_GROUP_PROBE_COL, 70_000, 0, 0), (A.JOIN_GROUP_PROBE_COL, 70_000, 0, 1),
             (A.JOIN_GROUP_NONE, 1, 0, 0)]
    try:
        if bloom:
            t.add_bloom(8)
        for b, e in ((0, npb), (333, 100_333), (npb - 5, npb)):
            counts = []
            for mode in (1, 2):
                ctx.tune(A.TUNE_JOIN_HOT, mode)
                counts.append(t.probe_count(cols[0], b, e))
            assert counts[0] == counts[1], (b, e)
            for gb, ng, kb, atomics in forms:
                for lim in (0, 20_000):
                    ctx.tune(A.TUNE_JOIN_ATOMICS, atomics)
                    res = []
                    for mode in (1, 2):
                        ctx.tune(A.TUNE_JOIN_HOT, mode)
                        res.append(t.probe_agg(cols[0], cols[1], gb, ng, kb, cols[2], b, e,
                                               build_rows=lim))
                    assert np.array_equal(res[0].count, res[1].count), (gb, ng, lim, b)
                    assert res[0].sums() == res[1].sums(), (gb, ng, lim, b)
                    assert res[0].n_out_of_range == res[1].n_out_of_range, (gb, ng, lim, b)
                    if b == 0 and lim in (0, 20_000) and atomics == 0:
                        n = lim or nb
                        cnt, sums, oor = B.join_agg(bk[:n], bp[:n], pk, pv, pg, gb, kb, ng, b, e)
                        assert np.array_equal(res[1].count, cnt), (gb, ng, lim)
                        assert [x[0] for x in res[1].sums()] == sums and \
                            res[1].n_out_of_range == oor, (gb, ng, lim)
    finally:
        ctx.tune(A.TUNE_JOIN_HOT, 0)
        ctx.tune(A.TUNE_JOIN_ATOMICS, 0)
        t.free()
        for c in keep + cols:
            c.free()
