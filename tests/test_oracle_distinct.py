"""CPU checks of the COUNT(DISTINCT) oracle: XXH3 against the xxhash library,
the reference's HyperLogLog tests (exec/hll.rs:130-188), exact counts against
numpy, and the interpreter's exact/HLL switch."""
import numpy as np
import pytest
import xxhash

from oracle import binding as B

M64 = (1 << 64) - 1


def test_xxh3_matches_the_xxhash_library():
    rng = np.random.default_rng(5)
    vals = [0, 1, 2, 255, 256, 2**32 - 1, 2**32, 2**63 - 1, 2**63, M64]
    vals += [int(x) for x in rng.integers(0, 2**64, 20000, dtype=np.uint64)]
    for v in vals:
        assert B.xxh3_u64(v) == xxhash.xxh3_64_intdigest(v.to_bytes(8, "little")), v


def test_registers_follow_the_library_hash():
    # register = top 14 bits of xxh3, rank = leading zeros of the other 50 + 1
    vals = np.arange(-5000, 15000, 7, dtype=np.int64)
    want = np.zeros(1 << 14, np.uint8)
    for v in vals:
        h = xxhash.xxh3_64_intdigest((int(v) % 2**64).to_bytes(8, "little"))
        w = (h << 14) & M64
        rank = (64 - w.bit_length()) + 1
        want[h >> 50] = max(want[h >> 50], rank)
    assert np.array_equal(B.hll_registers(vals), want)


def test_hll_reference_tests():  # exec/hll.rs:130-188
    assert B.hll_estimate(np.zeros(1 << 14, np.uint8)) == 0                  # test_hll_empty
    assert 1 <= B.hll_estimate(B.hll_registers(np.array([42], np.uint64))) <= 2  # single
    assert B.count_distinct(np.arange(1000, dtype=np.uint64)) == 1000          # small_exact
    big = np.arange(1_000_000, dtype=np.uint64)
    assert abs(B.count_distinct(big) - 1e6) / 1e6 < 0.02                       # large
    dup = np.concatenate([np.arange(500_000), np.arange(500_000)]).astype(np.uint64)
    assert abs(B.count_distinct(dup) - 5e5) / 5e5 < 0.02                       # duplicates
    # merge: the bytewise max of two halves' registers is the registers of the whole
    r1 = B.hll_registers(np.arange(500_000, dtype=np.uint64))
    r2 = B.hll_registers(np.arange(500_000, 1_000_000, dtype=np.uint64))
    assert np.array_equal(np.maximum(r1, r2), B.hll_registers(big))
    assert abs(B.hll_estimate(np.maximum(r1, r2)) - 1e6) / 1e6 < 0.02


def test_estimate_closed_forms():
    m = 1 << 14
    # every register 1: raw = alpha m^2 / (m/2) = 2 alpha m <= 2.5 m, no zero register
    alpha = 0.7213 / (1 + 1.079 / m)
    assert B.hll_estimate(np.ones(m, np.uint8)) == int(alpha * m * m / (m * 0.5))
    # half the registers 0, half 1: linear counting m ln(m / (m/2)) = m ln 2
    r = np.zeros(m, np.uint8)
    r[::2] = 1
    assert B.hll_estimate(r) == int(m * np.log(2.0))


@pytest.mark.parametrize("dtype", [np.int8, np.uint8, np.int16, np.int32, np.uint32, np.int64,
                                   np.uint64])
def test_exact_counts_match_numpy(dtype):
    rng = np.random.default_rng(1)
    info = np.iinfo(dtype)
    col = rng.integers(info.min, info.max, 50_000, dtype=dtype, endpoint=True)
    assert B.count_distinct(col, mode=1) == len(np.unique(col))
    sel = rng.integers(0, 2**63, (len(col) + 63) // 64).astype(np.uint64)
    bits = np.unpackbits(sel.view(np.uint8), bitorder="little")[: len(col)].astype(bool)
    assert B.count_distinct(col, sel, mode=1) == len(np.unique(col[bits]))


def test_auto_switches_at_100k_selected():
    # aggregate.rs:1387: exact below 100 000 selected cells, HyperLogLog from there
    col = np.arange(200_000, dtype=np.int64) % 150_001
    assert B.count_distinct(col[:99_999]) == 99_999
    regs = B.hll_registers(col[:100_000])
    assert B.count_distinct(col[:100_000]) == B.hll_estimate(regs)
    assert B.count_distinct(col[:100_000], mode=1) == 100_000


def test_library_estimate_matches_the_oracle():
    # tgp_hll_estimate is host code: its table of 2^-r must give the oracle's pow() sums
    from turbogp_amd import _capi as A
    rng = np.random.default_rng(9)
    for k in range(100):
        r = rng.integers(0, int(rng.integers(1, 66)), 1 << 14).astype(np.uint8)
        if k % 3 == 0:
            r[rng.random(1 << 14) < 0.3] = 0
        assert A.lib().tgp_hll_estimate(r.ctypes.data) == B.hll_estimate(r)
