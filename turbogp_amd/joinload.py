"""Join workload (build 2M unique keys, probe 60M Zipf(1.1) foreign keys) on the device.

Build side: bkey int64 (a seeded permutation of 0..n-1 through mix64), pay int32
uniform 0..24.  Probe side: fk int64 = the build key of a Zipf(1.1) rank (the
rank -> key map is a second seeded permutation), val int64 uniform
90 000..10 500 000, gkey int32 uniform [0, 1.5M).
"""
import numpy as np

from . import _capi as A
from . import workloads as W
from .device import DeviceColumn


def make_join_tables(ctx, n_build=W.JOIN_BUILD_ROWS, n_probe=W.JOIN_PROBE_ROWS,
                     gkey_card=W.JOIN_GKEY_CARD):
    bkey, _ = W.join_build_side(n_build)
    build = {"bkey": DeviceColumn.from_numpy(ctx, bkey, "bkey"),
             "pay": W.device_uniform(ctx, n_build, A.I32, W.COL_PAY, 0, W.JOIN_PAY_CARD - 1)}
    cdf = W.zipf_cdf(n_build)
    ktab = W.zipf_key_table(bkey)
    probe = {"fk": W.device_lookup_cdf(ctx, n_probe, A.I64, W.COL_FK, cdf, ktab),
             "val": W.device_uniform(ctx, n_probe, A.I64, W.COL_VAL, 90_000, 10_500_000),
             "gkey": W.device_uniform(ctx, n_probe, A.I32, W.COL_GKEY, 0, gkey_card - 1)}
    ctx.sync()
    return build, probe


# algorithmic bytes of the two join queries (probe rows x referenced widths +
# build rows x referenced widths)
def join_bytes(n_build, n_probe):
    return {"by_payload": n_probe * 16 + n_build * 12, "by_gkey": n_probe * 20 + n_build * 8}
