"""tgp_bitmap_to_indices_noalloc_sync (compaction into caller memory with the block
offsets in a fixed device array, no allocation) gives what tgp_bitmap_to_indices_sync
gives, through the C ABI: two calls in a row, the second on a smaller input (the
fixed scratch keeps the first call's words), u64 and u32 indices, a capacity error,
and a size past the fixed maximum rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

from turbogp_amd import DeviceBitmap, DeviceBuffer, _capi as A

pytestmark = pytest.mark.gpu


def _compact(ctx, fn, bm, nbits, base, index_bytes, cap):
    out = DeviceBuffer(ctx, max(8, cap * index_bytes))
    try:
        cnt = C.c_uint64()
        rc = fn(ctx.handle, bm.buf.ptr, nbits, base, index_bytes, out.ptr, cap, C.byref(cnt),
                None)
        vals = out.download(np.uint64 if index_bytes == 8 else np.uint32, cnt.value) \
            if rc == A.OK else None
        return rc, cnt.value, vals
    finally:
        out.free()


@pytest.mark.parametrize("index_bytes", [8, 4])
def test_noalloc_compaction_equals_the_allocating_one(ctx, index_bytes):
    lib = A.lib()
    rng = np.random.default_rng(index_bytes)
    for nbits, density, base in ((40_000_017, 0.3, 5), (1_000_003, 0.7, 1 << 20)):
        print(f"[noalloc] nbits={nbits} index_bytes={index_bytes}", flush=True)
        words = rng.integers(0, 2 ** 63, (nbits + 63) // 64, dtype=np.int64).view(np.uint64)
        words &= rng.integers(0, 2 ** 63, len(words), dtype=np.int64).view(np.uint64) \
            if density < 0.5 else ~np.uint64(0)
        bm = DeviceBitmap(ctx, nbits)
        try:
            bm.buf.upload(words)
            cap = nbits
            want = _compact(ctx, lib.tgp_bitmap_to_indices_sync, bm, nbits, base, index_bytes,
                            cap)
            got = _compact(ctx, lib.tgp_bitmap_to_indices_noalloc_sync, bm, nbits, base,
                           index_bytes, cap)
            assert want[0] == got[0] == A.OK
            assert want[1] == got[1] > 0
            assert np.array_equal(want[2], got[2])
            # too small an output: the count, and TGP_ERR_CAPACITY
            small = _compact(ctx, lib.tgp_bitmap_to_indices_noalloc_sync, bm, nbits, base,
                             index_bytes, got[1] - 1)
            assert small[0] == A.ERR_CAPACITY and small[1] == got[1]
        finally:
            bm.free()


def test_noalloc_rejects_sizes_past_its_maximum(ctx):
    cnt = C.c_uint64()
    out = DeviceBuffer(ctx, 64)
    bm = DeviceBitmap(ctx, 64)
    try:
        # rejected on the host before any kernel reads the (too small) bitmap
        rc = A.lib().tgp_bitmap_to_indices_noalloc_sync(ctx.handle, bm.buf.ptr, (1 << 30) + 1,
                                                        0, 8, out.ptr, 8, C.byref(cnt), None)
        assert rc == A.ERR_INVALID
        assert b"maximum" in A.lib().tgp_last_error()
        rc = A.lib().tgp_bitmap_to_indices_noalloc_sync(ctx.handle, bm.buf.ptr, 0, 0, 8,
                                                        out.ptr, 8, C.byref(cnt), None)
        assert rc == A.OK and cnt.value == 0
    finally:
        out.free()
        bm.free()
