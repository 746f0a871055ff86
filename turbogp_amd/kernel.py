"""Python view of the Kernel trait / KernelTable mirror (src/kernel/mod.rs:47-242)
exported by the C ABI (tgp_kernel_table_*, tgp_kernel_execute).

The names, argument meanings and result layout follow the reference:
  KernelTable.new(ctx).select(Operator.ScanEqU64, MemoryTier.L3) -> Kernel | None
  kernel.execute(input, output, KernelParams(...)) -> KernelResult
"""
import ctypes as C
import enum

import numpy as np

from . import _capi as A


class Operator(enum.IntEnum):  # kernel/mod.rs:47-74
    ScanEqU64 = 0
    ScanRangeU64 = 1
    ScanMultiPredicate = 2
    HashBuild = 3
    HashProbe = 4
    AggregateSumF64 = 5
    AggregateCountDistinct = 6
    SimilarityHamming = 7
    LeapfrogJoin = 8


class PredicateOp(enum.IntEnum):  # kernel/mod.rs:84-92
    Eq = 0
    Gt = 1
    Lt = 2


class MemoryTier(enum.IntEnum):  # memory/tier.rs:11-29
    L1L2 = 0
    L3 = 1
    Ddr5 = 2
    Hbm = 3
    Cxl = 4
    Nvme = 5
    NvmeOf = 6
    Network = 7


TARGET_GFX950 = 100


def KernelParams(target_u64=0, low_u64=0, high_u64=2**64 - 1, max_distance=0, cell_count=0,
                 target2_u64=0, target3_u64=0, pred1_op=PredicateOp.Eq, pred2_op=PredicateOp.Eq,
                 pred3_op=PredicateOp.Eq, predicate_count=0):
    """KernelParams with the reference's defaults (kernel/mod.rs:121-137)."""
    p = A.KernelParams()
    p.target_u64, p.low_u64, p.high_u64 = target_u64, low_u64, high_u64
    p.max_distance, p.cell_count = max_distance, cell_count
    p.target2_u64, p.target3_u64 = target2_u64, target3_u64
    p.pred1_op, p.pred2_op, p.pred3_op = int(pred1_op), int(pred2_op), int(pred3_op)
    p.predicate_count = predicate_count
    return p


class Kernel:
    def __init__(self, table, handle):
        self.table = table
        self.handle = handle
        op, tg, tier, name = C.c_int32(), C.c_int32(), C.c_int32(), C.c_char_p()
        A.check(A.lib().tgp_kernel_info(handle, C.byref(op), C.byref(tg), C.byref(tier),
                                        C.byref(name)), "tgp_kernel_info")
        self._op, self._target, self._tier = Operator(op.value), tg.value, MemoryTier(tier.value)
        self._name = name.value.decode()

    def operator(self):
        return self._op

    def target(self):
        return self._target

    def tier(self):
        return self._tier

    def name(self):
        return self._name

    def execute(self, input_ptr, output_ptr, params):
        """input_ptr/output_ptr: integer addresses (host or device) or numpy arrays."""
        def addr(x):
            if x is None:
                return None
            if isinstance(x, np.ndarray):
                return x.ctypes.data if x.size else None
            return int(x)
        r = A.KernelResult()
        A.check(A.lib().tgp_kernel_execute(self.handle, addr(input_ptr), addr(output_ptr),
                                           C.byref(params), C.byref(r)), "tgp_kernel_execute")
        return r


class KernelTable:
    """KernelTable::new() (kernel/mod.rs:191-206) on a device context."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = C.c_void_p()
        A.check(A.lib().tgp_kernel_table_new(ctx.handle, C.byref(h)), "tgp_kernel_table_new")
        self.handle = h.value

    def select(self, op, tier):
        k = C.c_void_p()
        A.check(A.lib().tgp_kernel_table_select(self.handle, int(op), int(tier), C.byref(k)),
                "tgp_kernel_table_select")
        return Kernel(self, k.value) if k.value else None

    def list(self):
        n = C.c_int32()
        A.check(A.lib().tgp_kernel_table_len(self.handle, C.byref(n)), "tgp_kernel_table_len")
        out = []
        for i in range(n.value):
            op, tg, tier, name = C.c_int32(), C.c_int32(), C.c_int32(), C.c_char_p()
            A.check(A.lib().tgp_kernel_table_entry(self.handle, i, C.byref(op), C.byref(tg),
                                                   C.byref(tier), C.byref(name)),
                    "tgp_kernel_table_entry")
            out.append((Operator(op.value), tg.value, MemoryTier(tier.value),
                        name.value.decode()))
        return out

    def close(self):
        if self.handle:
            A.lib().tgp_kernel_table_destroy(self.handle)
            self.handle = None
