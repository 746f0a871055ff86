"""CPU checks of the oracle's join restatement against the reference's own tests
(join_hash_table.rs:243-303, kernel/hash.rs:227-285, exec/join.rs:295-364) and
brute force."""
import numpy as np
import pytest

from oracle import binding as B


def pairs_of(build, probe):
    pi, pb, n = B.join_pairs(np.array(build, np.int64), np.array(probe, np.int64))
    return [(int(a), int(b)) for a, b in zip(pi, pb)], n


def test_basic_insert_probe():  # join_hash_table.rs test_basic_insert_probe
    build = [i * 7 for i in range(100)]
    pairs, n = pairs_of(build, build + [1, 999])
    assert n == 100 and pairs == [(i, i) for i in range(100)]


def test_duplicate_keys():  # join_hash_table.rs test_duplicate_keys: 42 -> {1, 2, 3}
    pairs, n = pairs_of([5, 42, 42, 42], [42])
    assert n == 3 and sorted(b for _, b in pairs) == [1, 2, 3]


def test_large_load():  # join_hash_table.rs test_large_load
    n = 100_000
    build = np.arange(n, dtype=np.int64)
    probe = np.arange(n + 1000, dtype=np.int64)
    pi, pb, m = B.join_pairs(build, probe)
    assert m == n and np.array_equal(pi, np.arange(n)) and np.array_equal(pb, np.arange(n))


def test_hash_distribution():  # join_hash_table.rs test_hash_distribution (pins CRC32C)
    buckets = np.zeros(256, np.int64)
    for i in range(100_000):
        buckets[B.lib().orc_join_hash(i) % 256] += 1
    assert buckets.min() > 300 and buckets.max() < 500


def test_crc32c_known_answer():
    # CRC-32C of the 8 little-endian bytes of 0 and of 1 (RFC 3720 polynomial);
    # the reference's _mm_crc32_u64(0, key) computes this with a zero seed.
    import binascii  # noqa: F401  (zlib's crc32 is a different polynomial)

    def crc32c(data):
        crc = 0
        for byte in data:
            crc ^= byte
            for _ in range(8):
                crc = (crc >> 1) ^ (0x82F63B78 & -(crc & 1))
        return crc
    for key in (0, 1, 0xDEADBEEF, 2**63 + 5):
        want = (crc32c(key.to_bytes(8, "little")) * 0x8648DBDB00000001) % 2**64
        assert B.lib().orc_join_hash(key) == want


def test_hash_kernel_kats():  # kernel/hash.rs hash_table_build_and_probe / kernel_build_and_probe
    pairs, _ = pairs_of([1, 2, 3, 2, 4, 2, 5], [2, 99])
    assert [b for _, b in pairs] == [1, 3, 5]
    _, n = pairs_of([1, 2, 3, 2, 4], [2, 4, 99])
    assert n == 3


def test_exec_join_kats():  # exec/join.rs test_inner_join / semi / anti / no_match
    left = [1, 2, 3]
    pairs, n = pairs_of([2, 3, 4], left)
    assert n == 2 and [left[p] for p, _ in pairs] == [2, 3]
    left = [1, 2, 3, 4]
    pairs, _ = pairs_of([2, 4], left)
    semi = sorted({left[p] for p, _ in pairs})
    assert semi == [2, 4] and [x for x in left if x not in semi] == [1, 3]
    assert pairs_of([3, 4], [1, 2])[1] == 0


def brute_join_agg(bk, bp, pk, pv, pg, group_by, key_base, ngroups):
    from collections import defaultdict
    rows = defaultdict(list)
    for r, k in enumerate(bk):
        rows[int(k)].append(r)
    cnt = [0] * ngroups
    sm = [0] * ngroups
    oor = 0
    for i, k in enumerate(pk):
        for r in rows.get(int(k), []):
            key = int(bp[r]) if group_by == 0 else int(pg[i]) if group_by == 1 else key_base
            g = key - key_base
            if not 0 <= g < ngroups:
                oor += 1
                continue
            cnt[g] += 1
            sm[g] += int(pv[i])
    return cnt, sm, oor


@pytest.mark.parametrize("group_by", [0, 1, 2])
def test_join_agg_vs_brute_force(group_by):
    rng = np.random.default_rng(group_by)
    nb, npb = 3000, 20000
    bk = rng.integers(-500, 2500, nb).astype(np.int64)      # duplicates on purpose
    bk[:3] = np.iinfo(np.int64).min                          # the GPU table's sentinel value
    bp = rng.integers(-2, 30, nb).astype(np.int32)
    pk = rng.integers(-600, 2600, npb).astype(np.int64)
    pk[::97] = np.iinfo(np.int64).min
    pv = rng.integers(-2**62, 2**62, npb).astype(np.int64)
    pg = rng.integers(0, 50, npb).astype(np.int32)
    ng = {0: 25, 1: 45, 2: 1}[group_by]
    cnt, sm, oor = B.join_agg(bk, bp, pk, pv, pg, group_by, 0, ng)
    bc, bs, boor = brute_join_agg(bk, bp, pk, pv, pg, group_by, 0, ng)
    assert list(map(int, cnt)) == bc and sm == bs and oor == boor


def test_hash_join_oracle_reference_kats():
    """exec/join.rs:294-353: inner, left, multi-key and no-match joins
    (hash_join builds on the right table and probes with the left)."""
    p, b = B.hash_join(np.array([2, 3, 4]), np.array([1, 2, 3]))            # test_inner_join
    assert len(p) == 2 and [[1, 2, 3][i] for i in p] == [2, 3]
    p, b = B.hash_join(np.array([2]), np.array([1, 2, 3]), join_type=1)     # test_left_join
    assert len(p) == 3 and list(p) == [0, 1, 2] and list(b) == [0xFFFFFFFF, 0, 0xFFFFFFFF]
    la, lb = np.array([1, 1, 2]), np.array([1, 2, 1])                       # test_multi_key_join
    ra, rb = np.array([1, 2]), np.array([2, 1])
    p, b = B.hash_join([ra, rb], [la, lb])
    assert len(p) == 2 and list(p) == [1, 2] and list(b) == [0, 1]
    p, _ = B.hash_join(np.array([3, 4]), np.array([1, 2]))                   # test_no_match
    assert len(p) == 0


def test_hash_join_oracle_matches_the_c_restatement():
    """Single-key pairs: the key-tuple dictionary against the C restatement of
    JoinHashTable (CRC32-tagged chains) on random keys with duplicates."""
    rng = np.random.default_rng(5)
    bk = rng.integers(-300, 300, 2000).astype(np.int64)
    pk = rng.integers(-400, 400, 5000).astype(np.int64)
    p1, b1 = B.hash_join(bk, pk)
    p2, b2, n = B.join_pairs(bk, pk)
    assert n == len(p1) and np.array_equal(p1, p2) and np.array_equal(b1, b2)
