"""The Kernel trait / KernelTable mirror, exercised like the reference's own
tests (kernel/mod.rs:377-402, kernel/scan.rs:601-884, kernel/hash.rs:227-285)."""
import ctypes as C

import numpy as np
import pytest

from turbogp_amd import DeviceColumn, _capi as A
from turbogp_amd.kernel import KernelParams, KernelTable, MemoryTier, Operator, PredicateOp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def table(ctx):
    t = KernelTable(ctx)
    yield t
    t.close()


def run_scan_eq(k, cells, target):
    cells = np.ascontiguousarray(cells, np.uint64)
    out = np.zeros(64, np.uint8)
    return k.execute(cells, out, KernelParams(target_u64=target, cell_count=len(cells))).count


def run_scan_multi(k, cells, preds):
    cells = np.ascontiguousarray(cells, np.uint64)
    kw = {"cell_count": len(cells), "predicate_count": min(3, len(preds))}
    names = [("target_u64", "pred1_op"), ("target2_u64", "pred2_op"), ("target3_u64", "pred3_op")]
    for (t, op), (tn, on) in zip(preds, names):
        kw[tn], kw[on] = t, op
    return k.execute(cells, np.zeros(64, np.uint8), KernelParams(**kw)).count


def test_kernel_table_has_kernels(table):  # mod.rs kernel_table_has_kernels
    assert table.list()


def test_select_returns_a_kernel(table):  # mod.rs select_returns_a_kernel
    k = table.select(Operator.ScanEqU64, MemoryTier.L3)
    assert k is not None and k.operator() == Operator.ScanEqU64 and k.target() == 100
    assert table.select(Operator.HashProbe, MemoryTier.Hbm).name() == "hash_probe_gfx950_hbm"
    assert table.select(Operator.SimilarityHamming, MemoryTier.L3) is None


def test_scan_eq_kats(table):  # scan.rs scalar_scan_eq_finds_matches, empty_input_returns_zero
    k = table.select(Operator.ScanEqU64, MemoryTier.L3)
    assert run_scan_eq(k, [1, 2, 3, 2, 4, 2, 5], 2) == 3
    assert run_scan_eq(k, [], 42) == 0
    cells = np.arange(1000, dtype=np.uint64) % 7
    assert run_scan_eq(k, cells, 3) == int((cells == 3).sum())
    r = k.execute(cells, None, KernelParams(target_u64=3, cell_count=1000))
    assert r.mask == sum(1 << i for i in range(64) if cells[i] == 3)


def test_scan_range_kats(table):  # scan.rs scalar_scan_range_finds_matches / avx512 range
    k = table.select(Operator.ScanRangeU64, MemoryTier.L3)
    cells = np.array([1, 5, 10, 15, 20, 25, 30], np.uint64)
    assert k.execute(cells, None, KernelParams(low_u64=10, high_u64=20,
                                               cell_count=7)).count == 3
    cells = np.arange(1000, dtype=np.uint64)
    assert k.execute(cells, None, KernelParams(low_u64=100, high_u64=200,
                                               cell_count=1000)).count == 101


def test_scan_multi_kats(table):  # scan.rs multi_predicate_scalar_*
    k = table.select(Operator.ScanMultiPredicate, MemoryTier.L3)
    E, G, L = PredicateOp.Eq, PredicateOp.Gt, PredicateOp.Lt
    assert run_scan_multi(k, np.arange(21), [(5, E), (2, G), (10, L)]) == 1
    assert run_scan_multi(k, np.full(100, 5), [(5, E), (0, G), (100, L)]) == 100
    assert run_scan_multi(k, np.arange(1000), [(1000, G), (1000, L), (5, E)]) == 0
    assert run_scan_multi(k, np.zeros(0), [(5, E), (2, G), (10, L)]) == 0
    assert run_scan_multi(k, np.arange(100), [(7, E)]) == 1
    assert run_scan_multi(k, np.arange(100), [(50, G), (60, L)]) == 9
    assert run_scan_multi(k, np.arange(100), [(10, G), (20, L)]) == 9
    assert run_scan_multi(k, np.arange(1234), [(100, G), (200, L)]) == 99


def test_scan_on_hbm_resident_input(ctx, table):
    k = table.select(Operator.ScanEqU64, MemoryTier.Hbm)
    cells = (np.arange(1 << 20, dtype=np.uint64) * 2654435761) % 1000
    col = DeviceColumn.from_numpy(ctx, cells)
    try:
        r = k.execute(col.buf.ptr, None, KernelParams(target_u64=7, cell_count=len(cells)))
        assert r.count == int((cells == 7).sum())
    finally:
        col.free()


def test_hash_kernel_build_and_probe(table):  # hash.rs hash_kernel_build_and_probe
    keys = np.array([1, 2, 3, 2, 4], np.uint64)
    out = np.zeros(8, np.uint64)
    kb = table.select(Operator.HashBuild, MemoryTier.Ddr5)
    r = kb.execute(keys, out, KernelParams(cell_count=5))
    assert r.count == 5 and out[0] != 0
    probe = np.array([int(out[0]), 2, 4, 99], np.uint64)
    kp = table.select(Operator.HashProbe, MemoryTier.L3)
    assert kp.execute(probe, None, KernelParams(cell_count=3)).count == 3
    A.check(A.lib().tgp_hash_table_destroy(table.ctx.handle, C.c_void_p(int(out[0]))),
            "tgp_hash_table_destroy")


def test_execute_errors_are_status_codes(ctx, table):
    k = table.select(Operator.ScanEqU64, MemoryTier.L3)
    A.lib().tgp_debug_inject_failure(ctx.handle, 1)
    with pytest.raises(A.TgpError) as e:
        run_scan_eq(k, [1, 2, 3], 2)
    assert e.value.code == A.ERR_INJECTED
    assert run_scan_eq(k, [1, 2, 3], 2) == 1
