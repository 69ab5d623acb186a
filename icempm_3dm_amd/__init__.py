"""icempm_3dm_amd — MI355X-native single-partition 3D MPM step (HIP/gfx950).

Drop-in for the one-partition step of igorg520b/iceMPM_3Dm's
GPU_Implementation6 / Model3D: brick radix sort, LDS-tiled P2G, sparse grid
update with a rigid indenter, G2P with the Drucker-Prager return mapping in
FP64. The compute runs in ``lib/libicempm_3dm.so`` through the C-ABI of
``include/icempm_3dm.h``; this package is its host-side mirror.
"""
from .capi import IcempmError, Indenter, SimParams3D, device_count, lib  # noqa: F401
from .gpu import GPUImplementation6  # noqa: F401
from .hssoa import HostSideSOA  # noqa: F401
from .model import Model3D  # noqa: F401

__all__ = ["IcempmError", "Indenter", "SimParams3D", "GPUImplementation6", "HostSideSOA", "Model3D",
           "device_count", "lib"]
