"""turbogp_amd -- MI355X (gfx950) operators for turboGP's scan / group-aggregate /
hash-join path, behind the C ABI declared in include/turbogp_gfx950.h."""
from . import _capi  # noqa: F401
from .device import (Context, DeviceBitmap, DeviceBuffer, DeviceColumn, Event,  # noqa: F401
                     GroupAggBuffers, GroupAggResult, HashTable, int128, make_agg_query)

__all__ = ["Context", "DeviceBitmap", "DeviceBuffer", "DeviceColumn", "GroupAggBuffers",
           "GroupAggResult", "HashTable", "int128", "make_agg_query"]
