"""amcl3d_amd: MI355X-native particle-filter measurement update for amcl3d.

The package is a thin host layer over ``lib/libamcl3d_gpu.so`` (hand-written
HIP for gfx950, C ABI in ``include/amcl3d_gpu.h``).  It mirrors the
reference's ``Grid3d`` and ``ParticleFilter`` classes and adds the batched
(fleet) and served forms.  There is no CPU implementation here: every call
runs on the GPU and fails loudly when the library is missing.
"""
from ._lib import (MAP_PARTICLE_LANES, MAP_POINT_LANES, ORDER_INDEX, ORDER_SPATIAL, Amcl3dError, DeviceBuffer,
                   LibraryMissing, RESAMPLE_FP64, RESAMPLE_REFERENCE, device_info, get_cloud_tile, get_deal, get_mapping,
                   get_order, get_tile, load, set_cloud_tile, set_deal, set_mapping, set_order, set_tile, synchronize)
from .grid3d import Grid3d
from .particle_filter import Graph, ParticleFilter

__all__ = ["Amcl3dError", "DeviceBuffer", "Graph", "Grid3d", "LibraryMissing", "MAP_PARTICLE_LANES", "MAP_POINT_LANES",
           "ORDER_INDEX", "ORDER_SPATIAL", "ParticleFilter", "RESAMPLE_FP64", "RESAMPLE_REFERENCE", "device_info",
           "get_cloud_tile", "get_deal", "get_mapping", "get_order", "get_tile", "load", "set_cloud_tile", "set_deal",
           "set_mapping", "set_order", "set_tile", "synchronize"]
