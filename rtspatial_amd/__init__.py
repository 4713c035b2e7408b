"""rtspatial_amd: MI355X-native LBVH spatial index with the LibRTS API.

The compute path is librtspatial_amd.so (hand-written HIP kernels for gfx950
behind the C ABI in include/rts_c.h); this package is its Python host mirror.
"""
from ._native import RtsError, lib  # noqa: F401
from .spatial_index import (  # noqa: F401
    Config,
    Predicate,
    Queue,
    SharedValue,
    SpatialIndex,
    pack_updates,
    profile_enable,
    profile_read,
    sort_pairs,
)

__all__ = [
    "Config",
    "Predicate",
    "Queue",
    "SharedValue",
    "SpatialIndex",
    "pack_updates",
    "sort_pairs",
    "profile_enable",
    "profile_read",
    "RtsError",
    "lib",
]
